"""The measured configurations as parity cases on the GPU.

* 200k cells (config 1: 40 levels, 1024 inputs, gap 500 ps) over its full
  0-20 Mps window: T0 + T1 + TX = DURATION on every record; one chunk,
  fixed 5 Mps chunks and a captured graph's replays give the same SAIF
  text and gate events (a 6,000-cell netlist of the same generator is
  pinned to the oracle's digests in test_gpu_parity.py).

* 2M cells (config 2: log-normal level widths, 60 levels, 4096 inputs, gap
  200 ps) over its full 0-20 Mps window: too large for the oracle, so it is
  checked through properties that hold at any size: T0 + T1 + TX = DURATION
  for every SAIF record, and the same SAIF text and gate-event count for two
  different time-chunk schedules (the chunk cut must not change a result).
* The batching mix: one netlist of the 200k-cell generator (scaled) driven
  by several stimulus sets with mean gaps U[100, 5000] ps, plus netlists of
  depth 10-60, all in one batched run: each instance equals its own plain
  run, and the instances the oracle finishes in seconds equal the oracle.
* The serving mix: a seeded Poisson stream of jobs with log-uniform netlist
  sizes and 30% netlist reuse: every job's SAIF equals the job run alone.
"""
import math
import os
import random
import sys
import time

import pytest

import gls
import oracle_bind

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = gls.Context(0)
    yield c
    c.close()


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def saif_records(text):
    """(name, T0, T1, TX) of every net record, and the DURATION."""
    recs, name, duration = [], None, None
    for line in text.splitlines():
        s = line.strip()
        if s.startswith("(DURATION"):
            duration = int(s.replace("(", " ").replace(")", " ").split()[1])
        elif s.startswith("(T0"):
            p = s.replace("(", " ").replace(")", " ").split()
            recs.append((name, int(p[1]), int(p[3]), int(p[5])))
        elif s.startswith("(") and not s.startswith(("(INSTANCE", "(NET", "(TIMESCALE")):
            name = s[1:]
    return recs, duration


def test_c2_full_window_properties(ctx):
    # 2M cells, 60 levels, 4096 inputs, 0-20 Mps: T0+T1+TX = DURATION on
    # every record; automatic chunks and fixed 0.5 Mps chunks give the same
    # SAIF text and gate events
    t0 = time.time()
    d = gls.Design.generate(2, 0x5EED0002, 2000000, 60, 4096)
    s = gls.Stimulus.generate(d, 0x5EED0002, 20000000, 200.0, 0.01)
    info = d.info()
    say("c2 generated", info["n_instances"], "cells", info["n_levels"], "levels",
        s.info()["n_events"], "stimulus events", "%.0fs" % (time.time() - t0))
    assert info["n_instances"] == 2000000 and info["n_levels"] == 60
    results = []
    for L in (0, 500000):
        t1 = time.time()
        r = ctx.simulate(d, s, 0, 20000001, chunk_ps=L)
        st = r.stats()
        text = r.saif_text()
        r.close()
        say("c2 run chunk_ps", L, "chunks", st["chunks"], "retries", st["chunk_retries"],
            "events", st["n_events"], "%.0fs" % (time.time() - t1))
        assert st["chunks"] > 1
        results.append((st["n_events"], text))
    (ev_a, saif_a), (ev_b, saif_b) = results
    assert ev_a == ev_b and ev_a > 10 ** 12
    assert saif_a == saif_b
    recs, duration = saif_records(saif_a)
    assert duration == 20000001
    assert len(recs) >= info["n_instances"]
    bad = [r for r in recs if r[1] + r[2] + r[3] != duration]
    assert not bad, bad[:5]


def test_c1_full_window_properties(ctx):
    t0 = time.time()
    d = gls.Design.generate(1, 0x5EED0001, 200000, 40, 1024)
    s = gls.Stimulus.generate(d, 0x5EED0001, 20000000, 500.0, 0.01)
    say("c1 generated", "%.0fs" % (time.time() - t0))
    r = ctx.simulate(d, s, 0, 20000001)
    st = r.stats()
    text = r.saif_text()
    r.close()
    say("c1", st["n_events"], "events", st["chunks"], "chunks")
    assert st["n_events"] > 5 * 10 ** 10
    recs, duration = saif_records(text)
    assert duration == 20000001 and len(recs) >= 200000
    bad = [x for x in recs if x[1] + x[2] + x[3] != duration]
    assert not bad, bad[:5]
    r = ctx.simulate(d, s, 0, 20000001, chunk_ps=5000000)
    assert r.stats()["chunks"] >= 4
    assert r.stats()["n_events"] == st["n_events"] and r.saif_text() == text
    r.close()
    g = ctx.prepare(d, s, 0, 20000001, graph=True)
    for rep_ in range(3):
        g.execute()
        assert g.stats()["n_events"] == st["n_events"], rep_
        assert g.saif_text() == text, rep_
    assert g.stats()["eval_launches"] == 1
    g.close()


def c3_mix():
    """One 20k-cell netlist of the 200k-cell generator (config 1, 40 levels)
    with 6 stimulus sets (0-200 kps, mean gap U[100, 5000] ps) and 2 short
    ones (0-10 kps), plus 4 netlists of log-uniform size (500-50k cells)
    and depth U{10..60}; (design, stimulus, dumpoff) per instance."""
    rng = random.Random(0x5EED0003)
    main = gls.Design.generate(1, 0x5EED0003, 20000, 40, 103)
    insts = []
    for i in range(8):
        t_end = 200000 if i < 6 else 10000
        s = gls.Stimulus.generate(main, 0x5EED0003 + 1 + i, t_end, rng.uniform(100, 5000), 0.01)
        insts.append((main, s, t_end + 1))
    for i in range(4):
        c = int(round(10 ** rng.uniform(math.log10(500), math.log10(50000))))
        d = gls.Design.generate(1, 0x5EED1003 + i, c, rng.randint(10, 60), max(16, c // 200))
        s = gls.Stimulus.generate(d, 0x5EED2003 + i, 200000, rng.uniform(100, 5000), 0.01)
        insts.append((d, s, 200001))
    return insts


def test_c3_mix_batched_equals_plain_and_oracle(ctx, tmp_path):
    insts = c3_mix()
    n = len(insts)
    on, off = [0] * n, [x[2] for x in insts]
    plain = []
    for k, (d, s, end) in enumerate(insts):
        r = ctx.simulate(d, s, 0, end, keep_waves=True)
        plain.append((r.stats()["n_events"], r.waves_hash(), r.saif_text()))
        r.close()
    say("c3 plain runs", [p[0] for p in plain])
    # eager batch with waveforms, then the captured graph replayed
    b = ctx.prepare_batch([x[0] for x in insts], [x[1] for x in insts], on, off, keep_waves=True)
    b.execute()
    for k in range(n):
        assert b.instance_waves_hash(k) == plain[k][1], k
        assert b.instance_saif_text(k) == plain[k][2], k
        assert b.instance_events(k) == plain[k][0], k
    assert b.stats()["n_events"] == sum(p[0] for p in plain)
    b.close()
    g = ctx.prepare_batch([x[0] for x in insts], [x[1] for x in insts], on, off, graph=True)
    for rep in range(2):
        g.execute()
        for k in range(n):
            assert g.instance_saif_text(k) == plain[k][2], (rep, k)
            assert g.instance_events(k) == plain[k][0], (rep, k)
    assert g.stats()["eval_launches"] == 1  # the second execution was a replay
    g.close()
    say("c3 batch == plain")
    # the instances the oracle finishes in seconds, against the oracle
    checked = 0
    for k, (d, s, end) in enumerate(insts):
        if plain[k][0] > 3000000:
            continue
        dd = tmp_path / ("i%d" % k)
        dd.mkdir()
        paths = [str(dd / x) for x in ("lib.vlib", "n.gv", "n.sdf", "s.vcd")]
        d.write(*paths[:3])
        s.write_vcd(paths[3])
        want = oracle_bind.run(*paths, 0, end, str(dd))
        say("c3 oracle", k, want["n_events"], "%.1fs" % want["sim_seconds"])
        assert want["n_events"] == plain[k][0], k
        assert want["waves_hash"] == plain[k][1], k
        assert want["saif"] == plain[k][2], k
        checked += 1
    assert checked >= 3


def test_c4_poisson_stream_alone_equals_mix(ctx):
    # 16 jobs: netlists log-uniform in [1k, 100k] cells, depth 10-60,
    # windows U[50k, 1M] ps, 30% of jobs reuse a netlist already submitted,
    # Poisson arrivals; every job's SAIF equals the job run alone
    rng = random.Random(0x5EED0004)
    designs, jobs = [], []
    for i in range(16):
        if designs and rng.random() < 0.3:
            d = rng.choice(designs)
        else:
            c = int(10 ** rng.uniform(3, 5))
            d = gls.Design.generate(1, 0x5EED4000 + i, c, rng.randint(10, 60), max(16, c // 200))
            designs.append(d)
        t_end = int(rng.uniform(5e4, 1e6))
        s = gls.Stimulus.generate(d, 0x5EED5000 + i, t_end, 500.0, 0.01)
        jobs.append((d, s, 0, t_end + 1))
    alone = []
    t_alone = []
    for d, s, on, off in jobs:
        t0 = time.perf_counter()
        r = ctx.simulate(d, s, on, off)
        t_alone.append(time.perf_counter() - t0)
        alone.append(r.saif_text())
        r.close()
    say("c4", len(jobs), "jobs,", len(designs), "netlists, mean alone %.1f ms" % (1e3 * sum(t_alone) / len(t_alone)))
    rate = 1.5 / (sum(t_alone) / len(t_alone))  # over-subscribed: jobs overlap
    srv = gls.Server(ctx, max_running=4)
    ids = []
    for d, s, on, off in jobs:
        time.sleep(rng.expovariate(rate))
        ids.append(srv.submit(d, s, on, off))
    for k, j in enumerate(ids):
        assert srv.wait(j, 120.0) == 0, k
        assert srv.saif_text(j) == alone[k], k
    st = srv.stats()
    say("c4 server", st)
    assert st["done"] == len(jobs) and st["failed"] == 0
    srv.close()
