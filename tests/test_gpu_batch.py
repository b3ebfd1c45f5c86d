"""Batched form, autotuner and serving loop (through the C ABI).

Bar: every instance of a batch and every served job produces exactly the
waveforms and SAIF text of its own plain run, and the plain runs are pinned
to the oracle (golden digests here, live oracle runs in test_gpu_parity.py).
"""
import json
import os
import random
import sys
import time

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


def mixed_instances():
    """Golden workloads plus netlists of other sizes and depths, stimuli with
    different gaps and windows; one netlist appears with two stimuli."""
    insts = []
    for name in ("c0_small", "c1_small"):
        d, s = golden_cases.make(name)
        insts.append((d, s) + tuple(golden_cases.CASES[name]["window"]) + (name,))
    rng = random.Random(7)
    for k in range(4):
        cfg = k % 2
        d = gls.Design.generate(cfg, 900 + k, rng.randint(80, 900), rng.randint(3, 30),
                                rng.randint(8, 64))
        for rep in range(1 + (k == 1)):
            t_end = rng.randint(5000, 80000)
            s = gls.Stimulus.generate(d, 40 + 10 * k + rep, t_end, rng.choice([20.0, 150.0, 600.0]),
                                      0.03)
            start = rng.randint(0, t_end // 4)
            insts.append((d, s, start, t_end + rng.randint(1, 300), None))
    return insts


def test_batch_matches_plain_runs(ctx):
    insts = mixed_instances()
    say("batch", len(insts), "instances")
    run = ctx.prepare_batch([i[0] for i in insts], [i[1] for i in insts], [i[2] for i in insts],
                            [i[3] for i in insts], keep_waves=True).execute()
    assert run.instances() == len(insts)
    for k, (d, s, on, off, name) in enumerate(insts):
        say("instance", k, name)
        plain = ctx.simulate(d, s, on, off, keep_waves=True)
        assert run.instance_saif_text(k) == plain.saif_text(), k
        assert run.instance_waves_hash(k) == plain.waves_hash(), k
        assert run.instance_events(k) == plain.stats()["n_events"], k
        if name:
            g = GOLDEN[name]
            assert "%016x" % run.instance_waves_hash(k) == g["waves_hash"]
            assert oracle_bind.sha256(run.instance_saif_text(k)) == g["saif_sha256"]
    # every instance's events are counted once
    total = sum(ctx.simulate(d, s, on, off).stats()["n_events"] for d, s, on, off, _ in insts)
    assert run.stats()["n_events"] == total


def test_batch_in_instance_groups(ctx):
    # a fixed 64 MiB pool splits the batch into instance groups (each runs
    # its own time chunks from time 0): every instance still equals its
    # plain run, eagerly (waveforms too) and replayed from a captured graph
    insts = mixed_instances()
    args = ([i[0] for i in insts], [i[1] for i in insts], [i[2] for i in insts],
            [i[3] for i in insts])
    plain = []
    for d, s, on, off, _ in insts:
        r = ctx.simulate(d, s, on, off, keep_waves=True)
        plain.append((r.waves_hash(), r.saif_text(), r.stats()["n_events"]))
        r.close()
    b = ctx.prepare_batch(*args, keep_waves=True, pool_bytes=64 << 20).execute()
    say("instance groups: chunks", b.stats()["chunks"])
    assert b.stats()["chunks"] >= 3
    for k in range(len(insts)):
        assert b.instance_waves_hash(k) == plain[k][0], k
        assert b.instance_saif_text(k) == plain[k][1], k
        assert b.instance_events(k) == plain[k][2], k
    b.close()
    g = ctx.prepare_batch(*args, graph=True, pool_bytes=64 << 20)
    for rep in range(3):
        g.execute()
        for k in range(len(insts)):
            assert g.instance_saif_text(k) == plain[k][1], (rep, k)
            assert g.instance_events(k) == plain[k][2], (rep, k)
    assert g.stats()["eval_launches"] == 1
    g.close()


def test_batch_segment_and_ticket_variants(ctx):
    insts = mixed_instances()[:4]
    args = ([i[0] for i in insts], [i[1] for i in insts], [i[2] for i in insts],
            [i[3] for i in insts])
    ref = ctx.prepare_batch(*args, keep_waves=True).execute()
    for seg, hbm, pf in ((2, False, True), (48, True, False), (1 << 20, False, True)):
        say("batch variant", seg, hbm, pf)
        r = ctx.prepare_batch(*args, keep_waves=True, seg_events=seg, hbm_stacks=hbm,
                              prefetch=pf).execute()
        for k in range(len(insts)):
            assert r.instance_waves_hash(k) == ref.instance_waves_hash(k), (seg, k)
            assert r.instance_saif_text(k) == ref.instance_saif_text(k), (seg, k)


def test_autotune_keeps_results(ctx):
    d, s = golden_cases.make("c1_small")
    on, off = golden_cases.CASES["c1_small"]["window"]
    run = ctx.prepare(d, s, on, off, keep_waves=True)
    rep = run.autotune()
    say("autotune", rep["seg_events"], rep["level_blocks"], rep["prefetch"], rep["best_ms"])
    assert rep["tried"] and rep["seg_events"] in {t["seg_events"] for t in rep["tried"]}
    g = GOLDEN["c1_small"]
    assert "%016x" % run.waves_hash() == g["waves_hash"]
    assert oracle_bind.sha256(run.saif_text()) == g["saif_sha256"]
    run.execute()  # later executions use the tuned setting
    assert "%016x" % run.waves_hash() == g["waves_hash"]


def test_two_pass_rejects_batches(ctx):
    insts = mixed_instances()[:2]
    with pytest.raises(gls.GlsError) as e:
        o = gls.Context._opts(path=gls.PATH_TWO_PASS)
        import ctypes
        n = 2
        D = (ctypes.c_void_p * n)(*[i[0].ptr for i in insts])
        S = (ctypes.c_void_p * n)(*[i[1].ptr for i in insts])
        on = (ctypes.c_uint64 * n)(*[i[2] for i in insts])
        off = (ctypes.c_uint64 * n)(*[i[3] for i in insts])
        h = ctypes.c_void_p()
        gls._check(gls.load().gls_prepare_batch(ctx.ptr, n, D, S, on, off, ctypes.byref(o),
                                                ctypes.byref(h)))
    assert e.value.status == "GLS_E_UNSUPPORTED"


def test_server_results_same_alone_and_in_mix(ctx):
    insts = mixed_instances()
    # alone: one job at a time through the plain path
    alone = [ctx.simulate(d, s, on, off).saif_text() for d, s, on, off, _ in insts]
    srv = gls.Server(ctx, max_running=4)
    jobs = []
    for rep in range(2):  # the second round reuses resident netlists
        for d, s, on, off, _ in insts:
            jobs.append(srv.submit(d, s, on, off))
    for k, j in enumerate(jobs):
        st = srv.wait(j)
        assert st == 0, (k, st)
        assert srv.saif_text(j) == alone[k % len(insts)], k
    stats = srv.stats()
    say("server", stats)
    assert stats["done"] == len(jobs) and stats["failed"] == 0
    assert stats["p99_ms"] >= stats["p50_ms"] > 0
    assert stats["peak_running"] >= 2  # jobs overlapped on the device
    # the second round takes the first round's device buffers from the arena
    assert stats["arena_hits"] > 0
    infos = [srv.info(j) for j in jobs]
    assert any(i["resident"] for i in infos[len(insts):])
    assert any(i["overlap"] for i in infos)
    srv.close()


def test_server_batches_queued_jobs(ctx):
    # one worker, up to 8 queued jobs admitted together run as one batch
    # (level l of every job in one launch): each job's SAIF and gate events
    # equal the job run alone
    insts = mixed_instances()
    alone = []
    for d, s, on, off, _ in insts:
        r = ctx.simulate(d, s, on, off)
        alone.append((r.saif_text(), r.stats()["n_events"]))
        r.close()
    srv = gls.Server(ctx, max_running=1, max_batch=8)
    jobs = [srv.submit(d, s, on, off) for rep in range(2) for d, s, on, off, _ in insts]
    for k, j in enumerate(jobs):
        assert srv.wait(j) == 0, k
        assert srv.saif_text(j) == alone[k % len(insts)][0], k
        assert srv.info(j)["n_events"] == alone[k % len(insts)][1], k
    st = srv.stats()
    say("server batches", st)
    assert st["done"] == len(jobs) and st["failed"] == 0
    assert st["peak_running"] >= 2  # queued jobs ran together
    srv.close()


def test_server_rejects_jobs_over_budget(ctx):
    d, s = golden_cases.make("c0_small")
    srv = gls.Server(ctx, device_budget=1 << 20, max_running=1)
    j = srv.submit(d, s, 0, 200001)
    assert srv.wait(j) == -3  # GLS_E_OOM: the estimate can never fit
    srv.close()


def test_server_oversize_job_does_not_block_the_queue(ctx):
    # one worker, an oversize job at the head of the queue and a small job
    # behind it: the oversize job fails with GLS_E_OOM, the small one still
    # runs (the admission scan goes on past a job it fails), and close()
    # (which drains) returns
    d, s = golden_cases.make("c0_small")
    on, off = golden_cases.CASES["c0_small"]["window"]
    ref = ctx.simulate(d, s, on, off).saif_text()
    s_big = gls.Stimulus.generate(d, 5, 20_000_000, 100.0, 0.01)
    srv = gls.Server(ctx, device_budget=256 << 20, max_running=1)
    big = srv.submit(d, s_big, 0, 20_000_001)
    small = srv.submit(d, s, on, off)
    assert srv.wait(small, 120.0) == 0
    assert srv.saif_text(small) == ref
    assert srv.wait(big, 10.0) == -3
    srv.close()


def test_server_cancel_and_deadline(ctx):
    """A queued job cancelled, a running job cancelled, a deadline already
    past and one far away, and a finished job cancelled: the cancelled and
    late jobs end with GLS_E_CANCELLED, the others keep their exact SAIF."""
    d, s = golden_cases.make("c0_small")
    on, off = golden_cases.CASES["c0_small"]["window"]
    ref = ctx.simulate(d, s, on, off).saif_text()
    # long enough to be caught while running: 50k cells, 40 levels, 5 Mps
    d_big = gls.Design.generate(1, 0x5EED0001, 50000, 40, 1024)
    s_big = gls.Stimulus.generate(d_big, 3, 5_000_000, 500.0, 0.01)
    srv = gls.Server(ctx, max_running=1)
    big = srv.submit(d_big, s_big, 0, 5_000_001)
    queued = srv.submit(d, s, on, off)
    late = srv.submit(d, s, on, off, deadline_s=1e-6)
    timely = srv.submit(d, s, on, off, deadline_s=600.0)
    # one worker: these wait behind the big job
    assert srv.cancel(queued) == -6
    assert srv.info(queued)["status"] == -6
    t0 = time.time()
    while srv.info(big)["status"] == 1 and time.time() - t0 < 60:
        time.sleep(0.001)
    assert srv.cancel(big) == 2  # running: stops before its next level
    assert srv.wait(big) == -6
    assert srv.wait(late) == -6
    assert srv.wait(timely) == 0
    assert srv.saif_text(timely) == ref
    assert srv.cancel(timely) == 0  # finished: keeps its result
    assert srv.saif_text(timely) == ref
    again = srv.submit(d, s, on, off)  # the cancelled job's memory is back
    assert srv.wait(again) == 0 and srv.saif_text(again) == ref
    st = srv.stats()
    say("server with cancels", st)
    assert st["cancelled"] == 3 and st["failed"] == 3 and st["done"] == 2
    srv.close()


def test_graph_replays_match_eager(ctx):
    d, s = golden_cases.make("c1_small")
    on, off = golden_cases.CASES["c1_small"]["window"]
    g = GOLDEN["c1_small"]
    d2, s2 = golden_cases.make("c0_small")
    r = ctx.prepare(d, s, on, off, keep_waves=True, graph=True)
    for i in range(4):
        say("graph execution", i)
        r.execute()  # the first one runs eagerly and captures, later ones replay
        assert "%016x" % r.waves_hash() == g["waves_hash"], i
        assert oracle_bind.sha256(r.saif_text()) == g["saif_sha256"], i
        if i:
            assert r.stats()["eval_launches"] == 1  # one graph launch
        # eager work and synchronisation between the replays
        e = ctx.simulate(d2, s2, 0, 200001, keep_waves=True)
        assert "%016x" % e.waves_hash() == GOLDEN["c0_small"]["waves_hash"]
        e.close()
    r.close()


def test_graph_batch_replays(ctx):
    insts = mixed_instances()
    args = ([i[0] for i in insts], [i[1] for i in insts], [i[2] for i in insts],
            [i[3] for i in insts])
    ref = ctx.prepare_batch(*args, keep_waves=True).execute()
    b = ctx.prepare_batch(*args, keep_waves=True, graph=True)
    for i in range(3):
        say("graph batch execution", i)
        b.execute()
        for k in range(len(insts)):
            assert b.instance_waves_hash(k) == ref.instance_waves_hash(k), (i, k)
            assert b.instance_saif_text(k) == ref.instance_saif_text(k), (i, k)


def test_batch_toggle_counts_match_plain(ctx):
    insts = mixed_instances()[:5]
    args = ([i[0] for i in insts], [i[1] for i in insts], [i[2] for i in insts],
            [i[3] for i in insts])
    b = ctx.prepare_batch(*args, saif_tc=True).execute()
    for k, (d, s, on, off, _) in enumerate(insts):
        plain = ctx.simulate(d, s, on, off, saif_tc=True).saif_text()
        assert "(TC " in plain
        assert b.instance_saif_text(k) == plain, k


def test_run_outlives_its_context():
    # a run freed after its context (as a garbage-collected handle can be)
    # still reads its results and frees its buffers: the context's stream
    # and buffer cache live until the run is gone
    c = gls.Context(0)
    d, s = golden_cases.make("c0_small")
    start, end = golden_cases.CASES["c0_small"]["window"]
    r = c.simulate(d, s, start, end)
    text = r.saif_text()
    c.close()
    assert r.saif_text() == text
    assert oracle_bind.sha256(text) == GOLDEN["c0_small"]["saif_sha256"]
    r.close()


def test_cached_buffers_make_room_for_new_netlists():
    # a finished run's pool stays cached by its context; a later batch of a
    # netlist not yet resident needs more room for its program image than
    # is left free, so the upload fails once, the cache is freed and the
    # upload runs again (the failure must not be reported again by the
    # batch's first kernel launch); results equal those on a fresh context
    c = gls.Context(0)
    try:
        free, total = c.mem_info()
        d0 = gls.Design.generate(0, 91, 300, 8, 16)
        s0 = gls.Stimulus.generate(d0, 92, 20000, 200.0, 0.01)
        big = (free - (512 << 20)) // 4096 * 4096
        say("cached pool", big >> 30, "GiB of", free >> 30)
        r = c.simulate(d0, s0, 0, 20001, pool_bytes=big)
        assert r.stats()["pool_bytes"] == big
        r.close()
        f2, _ = c.mem_info()
        assert f2 >= big  # cached blocks count as free
        d1 = gls.Design.generate(1, 93, 500000, 40, 2560)
        stims = [gls.Stimulus.generate(d1, 97 + k, 4000, 150.0, 0.01) for k in range(24)]
        say("batch of 24 x 500k cells")
        b = c.prepare_batch([d1] * 24, stims, [0] * 24, [4001] * 24).execute()
        got = [b.instance_saif_text(k) for k in (0, 11, 23)]
        b.close()
    finally:
        c.close()
    fresh = gls.Context(0)
    try:
        for k, i in enumerate((0, 11, 23)):
            r = fresh.simulate(d1, stims[i], 0, 4001)
            assert r.saif_text() == got[k], i
            r.close()
    finally:
        fresh.close()
