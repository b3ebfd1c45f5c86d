"""Gates of 5 and 6 distinct inputs (GEN_AO221_D1, GEN_AO222_D1; the
reference's GPU path takes up to 6, src/simulate.cu:25) against the oracle:
random netlists of the generator's wide-cell mix (configuration 3), in one
time chunk and in many, alone and in a batch with narrow netlists.  The
hand-derived known answers for both cells run in test_gpu_parity.py."""
import os
import random
import sys

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


def write(d, s, tmp_path, tag):
    dd = tmp_path / tag
    dd.mkdir()
    paths = [str(dd / n) for n in ("lib.vlib", "n.gv", "n.sdf", "s.vcd")]
    d.write(*paths[:3])
    s.write_vcd(paths[3])
    return paths, str(dd)


@pytest.mark.parametrize("seed", [21, 22, 23, 24])
def test_wide_cells_against_oracle(ctx, seed, tmp_path):
    rng = random.Random(seed)
    cells = rng.randint(150, 700)
    levels = rng.randint(4, 30)
    pis = rng.randint(8, 48)
    t_end = rng.randint(5000, 60000)
    gap = rng.choice([25.0, 80.0, 300.0])
    say("wide", dict(seed=seed, cells=cells, levels=levels, pis=pis, t_end=t_end, gap=gap))
    d = gls.Design.generate(3, 3000 + seed, cells, levels, pis)
    s = gls.Stimulus.generate(d, 300 + seed, t_end, gap, 0.05)
    paths, dd = write(d, s, tmp_path, "o")
    start, end = rng.randint(0, t_end // 4), t_end + rng.randint(1, 400)
    want = oracle_bind.run(*paths, start, end, dd)
    for L, seg in ((0, 0), (rng.randint(60, 900), 2), (t_end // 3, 32)):
        r = ctx.simulate(d, s, start, end, keep_waves=True, chunk_ps=L, seg_events=seg)
        assert r.stats()["n_events"] == want["n_events"], L
        assert r.waves_hash() == want["waves_hash"], L
        assert r.saif_text() == want["saif"], L
        r.close()


def test_wide_cells_in_a_batch(ctx):
    # wide-cell netlists batched with narrow ones: each instance equals its
    # plain run, eager and replayed from a captured graph
    insts = []
    for k in range(4):
        cfg = 3 if k % 2 == 0 else 0
        d = gls.Design.generate(cfg, 700 + k, 300 + 100 * k, 6 + 4 * k, 24)
        s = gls.Stimulus.generate(d, 710 + k, 40000, 90.0, 0.03)
        insts.append((d, s, 0, 40001))
    plain = []
    for d, s, on, off in insts:
        r = ctx.simulate(d, s, on, off, keep_waves=True)
        plain.append((r.waves_hash(), r.saif_text()))
        r.close()
    args = ([i[0] for i in insts], [i[1] for i in insts], [i[2] for i in insts], [i[3] for i in insts])
    b = ctx.prepare_batch(*args, keep_waves=True).execute()
    for k in range(len(insts)):
        assert b.instance_waves_hash(k) == plain[k][0], k
        assert b.instance_saif_text(k) == plain[k][1], k
    b.close()
    g = ctx.prepare_batch(*args, graph=True, chunk_ps=7000)
    for rep in range(2):
        g.execute()
        for k in range(len(insts)):
            assert g.instance_saif_text(k) == plain[k][1], (rep, k)
    g.close()


def test_two_pass_path_refuses_wide_cells(ctx):
    d = gls.Design.generate(3, 55, 300, 8, 16)
    s = gls.Stimulus.generate(d, 56, 5000, 100.0, 0.0)
    with pytest.raises(gls.GlsError) as e:
        ctx.simulate(d, s, 0, 5001, path=gls.PATH_TWO_PASS)
    assert e.value.status == "GLS_E_UNSUPPORTED"
