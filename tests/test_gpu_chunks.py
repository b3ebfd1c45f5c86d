"""Time-chunked runs of the fused path against the oracle and the golden
digests.

A run cut into time chunks (each chunk re-walks a look-back window of the
largest gate delay before its start, carries every gate output's tail, and
adds its share of the SAIF durations) must give exactly the waveforms, SAIF
text and gate-event count of the one-chunk run, for any chunk length; times
of 2^30 ps and beyond (each chunk in its own time base) must match the
oracle's 64-bit times.
"""
import json
import os
import random
import re
import sys

import pytest

import gls
import golden_cases
import oracle_bind

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = json.load(open(os.path.join(ROOT, "tests", "golden", "oracle_golden.json")))


@pytest.fixture(scope="module")
def ctx():
    c = gls.Context(0)
    yield c
    c.close()


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def check_golden(r, name):
    g = GOLDEN[name]
    st = r.stats()
    assert st["n_events"] == g["n_events"]
    assert "%016x" % r.waves_hash() == g["waves_hash"]
    assert oracle_bind.sha256(r.saif_text()) == g["saif_sha256"]


@pytest.mark.parametrize("name,chunks", [("c0_small", (997, 5000, 33333)),
                                         ("c1_small", (101, 4096, 25000)),
                                         ("c0_full", (150000, 700001))])
def test_chunk_length_invariance(ctx, name, chunks):
    # fixed chunk lengths from just above the largest delay (50 ps) to a
    # third of the window: identical to the one-chunk run's digests
    d, s = golden_cases.make(name)
    start, end = golden_cases.CASES[name]["window"]
    for L in chunks:
        say("chunks", name, L)
        r = ctx.simulate(d, s, start, end, keep_waves=True, chunk_ps=L)
        assert r.stats()["chunks"] > 1, L
        check_golden(r, name)
        r.close()
    r = ctx.simulate(d, s, start, end, keep_waves=True, chunk_ps=gls.NO_CHUNKS)
    assert r.stats()["chunks"] == 1
    check_golden(r, name)


@pytest.mark.parametrize("seed", [11, 12, 13, 14])
def test_chunked_random_designs_against_oracle(ctx, seed, tmp_path):
    # the three generator configurations (config 2: log-normal level widths)
    # with random chunk lengths and segment sizes, against a live oracle run
    rng = random.Random(seed)
    cfg = seed % 3
    cells = rng.randint(100, 600)
    levels = rng.randint(3, 30) if cfg != 2 else rng.randint(10, 60)
    pis = rng.randint(4, 40)
    t_end = rng.randint(5000, 80000)
    gap = rng.choice([20.0, 60.0, 200.0])
    say("chunked random", dict(seed=seed, cfg=cfg, cells=cells, levels=levels, pis=pis, t_end=t_end))
    d = gls.Design.generate(cfg, 2000 + seed, cells, levels, pis)
    s = gls.Stimulus.generate(d, 99 + seed, t_end, gap, 0.05)
    paths = [str(tmp_path / n) for n in ("lib.vlib", "n.gv", "n.sdf", "s.vcd")]
    d.write(*paths[:3])
    s.write_vcd(paths[3])
    start, end = rng.randint(0, t_end // 3), t_end + rng.randint(1, 500)
    want = oracle_bind.run(*paths, start, end, str(tmp_path))
    for L, seg in ((rng.randint(51, 400), 0), (rng.randint(400, 9000), 2), (t_end // 2, 32)):
        r = ctx.simulate(d, s, start, end, keep_waves=True, chunk_ps=L, seg_events=seg)
        assert r.stats()["n_events"] == want["n_events"], L
        assert r.waves_hash() == want["waves_hash"], L
        assert r.saif_text() == want["saif"], L
        r.close()


def test_config2_generator_against_oracle(ctx, tmp_path):
    # the 2M-cell configuration's generator (log-normal level widths, 60
    # levels, gap 200 ps) at a size the oracle finishes in seconds, one chunk
    # and many
    d = gls.Design.generate(2, 0x5EED0002, 3000, 60, 64)
    s = gls.Stimulus.generate(d, 0x5EED0002, 40000, 200.0, 0.01)
    paths = [str(tmp_path / n) for n in ("lib.vlib", "n.gv", "n.sdf", "s.vcd")]
    d.write(*paths[:3])
    s.write_vcd(paths[3])
    want = oracle_bind.run(*paths, 0, 40001, str(tmp_path))
    for L in (0, 3001):
        say("config 2", L)
        r = ctx.simulate(d, s, 0, 40001, keep_waves=True, chunk_ps=L)
        assert r.stats()["n_events"] == want["n_events"]
        assert r.waves_hash() == want["waves_hash"]
        assert r.saif_text() == want["saif"]


def shifted_vcd(src, dst, offset):
    """Copy a VCD moving every timestamp after #0 by `offset`."""
    out = []
    for line in open(src):
        m = re.match(r"#(\d+)\s*$", line)
        if m and int(m.group(1)) > 0:
            line = "#%d\n" % (int(m.group(1)) + offset)
        out.append(line)
    open(dst, "w").write("".join(out))


def test_times_beyond_30_bits_against_oracle(ctx, tmp_path):
    # stimulus at 3e9 ps and beyond (past 2^30 and 2^31): every chunk is
    # simulated in its own time base; waveforms (64-bit times) and SAIF equal
    # the oracle's
    off = 3_000_000_000
    d = gls.Design.generate(1, 4242, 300, 12, 24)
    s = gls.Stimulus.generate(d, 4243, 30000, 100.0, 0.02)
    paths = [str(tmp_path / n) for n in ("lib.vlib", "n.gv", "n.sdf", "s0.vcd")]
    d.write(*paths[:3])
    s.write_vcd(paths[3])
    vcd = str(tmp_path / "s.vcd")
    shifted_vcd(paths[3], vcd, off)
    s2 = gls.Stimulus.from_vcd(vcd)
    start, end = off - 5000, off + 30500
    want = oracle_bind.run(paths[0], paths[1], paths[2], vcd, start, end, str(tmp_path), waves=True)
    # automatic chunks; 12 fixed chunks of 2.5e8 ps; automatic chunks under a
    # 2 MiB pool (the active stretch near 3e9 is cut finer, chunks are run
    # again shorter)
    for L, pool in ((0, 0), (250_000_000, 0), (0, 2 << 20)):
        say("large times", L, pool)
        r = ctx.simulate(d, s2, start, end, keep_waves=True, chunk_ps=L, pool_bytes=pool)
        st = r.stats()
        say("  chunks", st["chunks"], "retries", st["chunk_retries"])
        assert st["chunks"] >= 2  # 0 .. 3e9 cannot be one 30-bit chunk
        assert r.stats()["n_events"] == want["n_events"]
        assert r.saif_text() == want["saif"]
        assert r.waves_hash() == want["waves_hash"]
        for net in list(want["waves"])[:40]:
            assert r.waveform64(net) == want["waves"][net], net
    with pytest.raises(gls.GlsError) as e:
        ctx.simulate(d, s2, start, end, chunk_ps=gls.NO_CHUNKS)
    assert e.value.status == "GLS_E_RANGE"


def test_small_pool_is_cut_into_chunks(ctx):
    # a fixed 2 MiB pool cannot hold the one-chunk run: automatic chunks
    # shrink until each fits, with the golden results
    d, s = golden_cases.make("c1_small")
    start, end = golden_cases.CASES["c1_small"]["window"]
    r = ctx.simulate(d, s, start, end, keep_waves=True, pool_bytes=2 << 20)
    st = r.stats()
    say("small pool", st["chunks"], st["chunk_retries"], st["pool_bytes"])
    assert st["chunks"] > 1 and st["pool_bytes"] == 2 << 20
    check_golden(r, "c1_small")


def test_no_chunks_pool_exhaustion_is_reported(ctx):
    # one chunk only, the pool clamped to 4 KiB: the run fails with
    # GLS_E_OOM and the context stays usable
    d, s = golden_cases.make("c0_small")
    with pytest.raises(gls.GlsError) as e:
        ctx.simulate(d, s, 0, 200001, pool_bytes=4096, chunk_ps=gls.NO_CHUNKS)
    assert e.value.status == "GLS_E_OOM"
    r = ctx.simulate(d, s, 0, 200001, keep_waves=True)
    check_golden(r, "c0_small")


def test_chunked_graph_replay(ctx):
    # the chunk and level sequence captured once and replayed: same SAIF and
    # events as the plain run, replay after replay with eager work between
    d, s = golden_cases.make("c1_small")
    start, end = golden_cases.CASES["c1_small"]["window"]
    g = GOLDEN["c1_small"]
    r = ctx.prepare(d, s, start, end, graph=True, chunk_ps=9000)
    for i in range(3):
        r.execute()
        st = r.stats()
        assert st["n_events"] == g["n_events"], i
        assert oracle_bind.sha256(r.saif_text()) == g["saif_sha256"], i
        assert st["chunks"] > 1
        other = ctx.simulate(d, s, start, end, chunk_ps=4000)
        assert oracle_bind.sha256(other.saif_text()) == g["saif_sha256"]
        other.close()
    assert r.stats()["eval_launches"] == 1  # the last execution was a replay
    # waveforms are kept by graph replays of one-chunk runs only
    with pytest.raises(gls.GlsError) as e:
        ctx.prepare(d, s, start, end, graph=True, chunk_ps=9000, keep_waves=True).execute()
    assert e.value.status == "GLS_E_UNSUPPORTED"


def test_chunked_batch_matches_plain_runs(ctx):
    # a batch of instances of different windows run in shared chunks: each
    # instance equals its plain one-chunk run
    specs = [("c0_small", None), ("c1_small", None)]
    ds, ss, ons, offs, want = [], [], [], [], []
    for name, _ in specs:
        d, s = golden_cases.make(name)
        on, off = golden_cases.CASES[name]["window"]
        ds.append(d)
        ss.append(s)
        ons.append(on)
        offs.append(off)
        want.append(GOLDEN[name])
    d3 = gls.Design.generate(2, 777, 800, 30, 16)
    s3 = gls.Stimulus.generate(d3, 778, 150000, 300.0, 0.01)
    ref3 = ctx.simulate(d3, s3, 1000, 150001, keep_waves=True)
    ds.append(d3)
    ss.append(s3)
    ons.append(1000)
    offs.append(150001)
    for L in (0, 20000):
        say("chunked batch", L)
        b = ctx.prepare_batch(ds, ss, ons, offs, keep_waves=True, chunk_ps=L)
        b.execute()
        for i, g in enumerate(want):
            assert "%016x" % b.instance_waves_hash(i) == g["waves_hash"], (L, i)
            assert oracle_bind.sha256(b.instance_saif_text(i)) == g["saif_sha256"], (L, i)
        assert b.instance_waves_hash(2) == ref3.waves_hash(), L
        assert b.instance_saif_text(2) == ref3.saif_text(), L
        b.close()


def test_chunked_toggle_counts(ctx):
    # SAIF toggle counts are owned by the chunk of each entry's time
    d, s = golden_cases.make("c1_small")
    start, end = golden_cases.CASES["c1_small"]["window"]
    one = ctx.simulate(d, s, start, end, saif_tc=True, chunk_ps=gls.NO_CHUNKS).saif_text()
    assert "(TC " in one
    for L in (333, 12000):
        assert ctx.simulate(d, s, start, end, saif_tc=True, chunk_ps=L).saif_text() == one, L
