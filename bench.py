"""Benchmark: gate waveform events/s of the plain per-level HIP path.

Workload (one "step" = one full simulation of every level over the whole
window, inputs already resident in HBM): the 200,000-cell configuration —
generic 12-cell library, 1,024 primary inputs, 40 levels, Pareto(1.5)
fanout capped at 256, SDF delays U{5..50} ps, Poisson stimulus with mean gap
500 ps over 0..20,000,000 ps, 1% X toggles.  Synthetic data from the seeded
generator (seed 0x5EED0001).

Prints one JSON line (rank 0).  Only one GPU is used; --gpus N > 1 exits.
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "gate-level-simulator_amd"))

METRIC = "gate waveform events/sec, SAIF bit-exact; batched vs plain; p99 job latency"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(budget_events):
    """The oracle (serial restatement of the reference CPU simulator) on one
    core, on a bounded sample of the same workload: a 64-cell, 64-input
    netlist from the same generator over the same full 20 Mps window (40k
    events per input), stopped after `budget_events` gate events."""
    import gls
    exe = os.path.join(ROOT, "oracle", "_build", "cpu_ref")
    if not os.path.exists(exe):
        return None
    d = gls.Design.generate(1, 0x5EED0001, 64, 1, 64)
    s = gls.Stimulus.generate(d, 0x5EED0001, 20000000, 500.0, 0.01)
    with tempfile.TemporaryDirectory() as t:
        paths = [os.path.join(t, n) for n in ("lib.vlib", "n.gv", "n.sdf", "s.vcd")]
        d.write(*paths[:3])
        s.write_vcd(paths[3])
        cmd = [exe] + paths + ["0", "20000001", os.path.join(t, "o.saif"), str(budget_events)]
        try:
            cmd = ["taskset", "-c", "0"] + cmd
            out = subprocess.run(cmd, check=True, capture_output=True, text=True, timeout=600).stdout
        except (subprocess.CalledProcessError, FileNotFoundError):
            out = subprocess.run(cmd[3:], check=True, capture_output=True, text=True, timeout=600).stdout
    r = json.loads(out.strip().splitlines()[-1])
    return {"value": round(r["events_per_s"], 1), "unit": "gate events/s", "cores": 1, "kind": "port",
            "sample": f"first {r['n_events']} gate events of a 64-cell, 64-input netlist from the same "
                      f"generator over the full 0..20 Mps window (40k events per input); "
                      f"oracle/cpu_ref, 1 thread, {r['sim_seconds']:.1f} s"}


def measured_traffic(args, launches):
    """HBM bytes per launch of the fused level kernel, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary of this same default workload
    (counters scaled by a plain copy of known size).  None for any other
    workload or path: counters cannot be read from inside this process."""
    if (args.workload != "c1" or args.path != "fused" or args.cells != 200000 or args.t_end != 20000000
            or args.seg_events or args.level_blocks or args.no_prefetch
            or args.hbm_stacks):
        return None
    p = os.path.join(ROOT, "profiles", "k_level_200k_cells_hbm_counters.txt")
    try:
        for line in open(p):
            if line.startswith("traffic per launch (mean)"):
                return round(float(line.split()[4]) * 1e9)
    except OSError:
        pass
    return None


def pool_bytes_for(d, s):
    """Event-pool size for one instance: stimulus plus each gate's output
    reservation (its inputs' events; internal nets toggle a few times more
    often than the inputs), with headroom."""
    gates = d.info()["n_gates"]
    sig, ev = s.info()["n_signals"], s.info()["n_events"]
    return 4 * (ev + 12 * gates * (ev // max(1, sig) + 1) + (1 << 20))


def batched_vs_plain(ctx, n_sets, scale, budget_frac=0.6):
    """Batched form against plain runs on the batching workload: one netlist
    from the 200k-cell generator at `scale` of its size driven by `n_sets`
    seeded stimulus sets (0-2 Mps windows, mean gap U[100, 5000] ps), plus
    n_sets/4 netlists with log-uniform cell counts (5k-500k x scale) and
    depths U{10..60}.  Plain: every instance prepared and executed alone, one
    after the other.  Batched: the instances in as few prepared runs as fit
    `budget_frac` of free HBM (each run: level l of all its instances in one
    launch, the level sequence captured as one hipGraph), autotuned on the
    first run, then executed.  Every instance's SAIF is compared with its
    plain run."""
    import random
    import gls
    rng = random.Random(0x5EED0003)
    cells = max(1000, int(200000 * scale))
    main_d = gls.Design.generate(1, 0x5EED0003, cells, 40, max(16, int(1024 * scale)))
    insts = []
    for i in range(n_sets):
        s = gls.Stimulus.generate(main_d, 0x5EED0003 + 1 + i, 2000000, rng.uniform(100, 5000), 0.01)
        insts.append((main_d, s))
        if i % 64 == 63:
            log(f"bench: batch stimulus sets {i + 1}/{n_sets} generated")
    for i in range(max(1, n_sets // 4)):
        c = int(round(10 ** rng.uniform(3.7, 5.7) * scale))
        c = max(500, c)
        d = gls.Design.generate(1, 0x5EED1003 + i, c, rng.randint(10, 60), max(16, int(1024 * c / 200000)))
        s = gls.Stimulus.generate(d, 0x5EED2003 + i, 2000000, rng.uniform(100, 5000), 0.01)
        insts.append((d, s))
        if i % 16 == 15:
            log(f"bench: batch netlists {i + 1}/{max(1, n_sets // 4)} generated")
    # plain: one instance at a time (the first run of each pays nothing extra:
    # its netlist image is uploaded in prepare, outside the timed execute)
    log(f"bench: batch workload generated: {len(insts)} instances")
    plain_s, plain_ev, saifs, used = 0.0, 0, [], []
    for k, (d, s) in enumerate(insts):
        r = ctx.prepare(d, s, 0, 2000001)
        r.execute()
        t0 = time.perf_counter()
        r.execute()
        plain_s += time.perf_counter() - t0
        st = r.stats()
        plain_ev += st["n_events"]
        used.append(st["pool_used_bytes"])
        saifs.append(r.saif_text())
        r.close()
        if k % 8 == 7:
            log(f"bench: plain instance {k + 1}/{len(insts)} done")
    # groups of instances whose waveforms fit the budget together
    free, _ = ctx.mem_info()
    budget = int(free * budget_frac)
    groups, cur, cur_b = [], [], 0
    for k in range(len(insts)):
        need = int(used[k] * 1.25) + (8 << 20)
        if cur and cur_b + need > budget:
            groups.append((cur, cur_b))
            cur, cur_b = [], 0
        cur.append(k)
        cur_b += need
    groups.append((cur, cur_b))
    log(f"bench: {len(groups)} batched run(s) within {budget / 2**30:.0f} GiB")
    batch_s, batch_ev, same, replayed, tune, knobs = 0.0, 0, True, True, None, {}
    for gi, (ks, pool) in enumerate(groups):
        b = ctx.prepare_batch([insts[k][0] for k in ks], [insts[k][1] for k in ks], [0] * len(ks),
                              [2000001] * len(ks), pool_bytes=pool, pool_grow=True, graph=True,
                              **knobs)
        if tune is None:
            log("bench: batch prepared, autotuning")
            tune = b.autotune()
            knobs = {"seg_events": tune["seg_events"], "level_blocks": tune["level_blocks"],
                     "prefetch": bool(tune["prefetch"])}
            log(f"bench: autotune picked {tune['seg_events']} events/segment, "
                f"{tune['level_blocks']} workgroups, prefetch {tune['prefetch']}")
        b.execute()  # eager run with the tuned setting, captured as one graph
        t0 = time.perf_counter()
        b.execute()  # graph replay: every level of every instance, no host round trip
        batch_s += time.perf_counter() - t0
        replayed = replayed and b.stats()["eval_launches"] == 1
        batch_ev += b.stats()["n_events"]
        same = same and all(b.instance_saif_text(j) == saifs[k] for j, k in enumerate(ks))
        b.close()
        log(f"bench: batched run {gi + 1}/{len(groups)} done ({len(ks)} instances)")
    return {"workload": f"{n_sets} stimulus sets on a {cells}-cell netlist + {len(insts) - n_sets} "
                        f"netlists (log-uniform {max(500, int(5000 * scale))}-{int(500000 * scale)} "
                        f"cells, depth 10-60), 0-2 Mps windows, gap U[100,5000] ps",
            "instances": len(insts), "gate_events": batch_ev, "batched_runs": len(groups),
            "plain_events_per_s": round(plain_ev / plain_s, 1), "plain_s": round(plain_s, 4),
            "batched_events_per_s": round(batch_ev / batch_s, 1), "batched_s": round(batch_s, 4),
            "speedup": round(plain_s / batch_s, 3), "saif_bit_exact_vs_plain": same,
            "events_same_as_plain": batch_ev == plain_ev,
            "form": "captured hipGraph of all levels per batched run" if replayed
                    else "level-synchronous launches",
            "autotune": {k: tune[k] for k in ("seg_events", "level_blocks", "prefetch", "best_ms")}}


def serving_mix(ctx, n_jobs, scale):
    """Serving loop on a seeded Poisson job stream: netlists log-uniform in
    [10k, 2M] x `scale` cells, windows U[1, 20] Mps x `scale`, 30% of jobs
    reuse a netlist already submitted; the arrival rate is set from the mean
    time of a few jobs run alone, for about 75% device utilisation.  Reports
    events/s and p50/p99 job latency; every job's SAIF is checked against the
    same job run alone."""
    import random
    import gls
    rng = random.Random(0x5EED0004)
    designs, jobs = [], []
    for i in range(n_jobs):
        if designs and rng.random() < 0.3:
            d = rng.choice(designs)
        else:
            c = max(500, int(10 ** rng.uniform(4, 6.3) * scale))
            d = gls.Design.generate(1, 0x5EED4000 + i, c, rng.randint(10, 60), max(16, int(1024 * c / 200000)))
            designs.append(d)
        t_end = int(rng.uniform(1e6, 2e7) * scale)
        s = gls.Stimulus.generate(d, 0x5EED5000 + i, t_end, 500.0, 0.01)
        jobs.append((d, s, 0, t_end + 1))
    log(f"bench: serving workload generated: {n_jobs} jobs, {len(designs)} netlists")
    alone, t_alone = [], []
    for k, (d, s, on, off) in enumerate(jobs):
        # the second run of each job is timed: its netlist is resident, as
        # for a served job
        for rep in range(2):
            t0 = time.perf_counter()
            r = ctx.simulate(d, s, on, off, pool_bytes=pool_bytes_for(d, s))
            dt = time.perf_counter() - t0
            if rep == 0:
                alone.append(r.saif_text())
            r.close()
        t_alone.append(dt)
        if k % 10 == 9:
            log(f"bench: alone job {k + 1}/{n_jobs} done")
    mean_s = sum(t_alone) / len(t_alone)
    rate = 0.75 / mean_s  # jobs per second
    gaps = [rng.expovariate(rate) for _ in jobs]
    srv = gls.Server(ctx, max_running=4)
    ids = []
    for (d, s, on, off), g in zip(jobs, gaps):
        time.sleep(g)
        ids.append(srv.submit(d, s, on, off))
    log(f"bench: {n_jobs} jobs submitted")
    same = True
    for k, j in enumerate(ids):
        if srv.wait(j) != 0 or srv.saif_text(j) != alone[k]:
            same = False
    st = srv.stats()
    srv.close()
    return {"workload": f"{n_jobs} Poisson jobs at {rate:.2f}/s, cells log-uniform in "
                        f"[{int(1e4 * scale)}, {int(2e6 * scale)}], windows U[{int(1e6 * scale)}, "
                        f"{int(2e7 * scale)}] ps, gap 500 ps, 30% netlist reuse",
            "jobs": int(st["done"]), "failed": int(st["failed"]),
            "events_per_s": round(st["events_per_s"], 1), "p50_ms": round(st["p50_ms"], 3),
            "p99_ms": round(st["p99_ms"], 3), "mean_alone_ms": round(mean_s * 1e3, 3),
            "peak_running": int(st["peak_running"]), "resident_programs": int(st["resident_programs"]),
            "saif_same_alone_and_in_mix": same}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=["c1", "c2", "c3"], default="c1",
                    help="c1: the 200k-cell configuration (default); c2: the 2M-cell, 60-level "
                         "log-normal-width configuration at gap 200 ps on a shortened window; "
                         "c3: the full batching mix only (256 stimulus sets on the 200k-cell "
                         "netlist + 64 netlists of 5k-500k cells), batched vs plain")
    ap.add_argument("--cells", type=int, default=None)
    ap.add_argument("--t-end", type=int, default=None)
    ap.add_argument("--seg-events", type=int, default=0)
    ap.add_argument("--path", choices=["fused", "two-pass"], default="fused")
    ap.add_argument("--level-blocks", type=int, default=0)
    ap.add_argument("--no-prefetch", action="store_true")
    ap.add_argument("--hbm-stacks", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-events", type=int, default=700000)
    ap.add_argument("--graph-steps", type=int, default=2,
                    help="also time the default workload as a captured graph (0: skip)")
    ap.add_argument("--batch-sets", type=int, default=32,
                    help="stimulus sets of the batched-vs-plain measurement (0: skip)")
    ap.add_argument("--batch-scale", type=float, default=0.1)
    ap.add_argument("--serve-jobs", type=int, default=40,
                    help="jobs of the serving measurement (0: skip)")
    ap.add_argument("--serve-scale", type=float, default=0.05)
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run a plain device copy of this many MiB (3 times): the "
                         "reference point that scales FETCH_SIZE/WRITE_SIZE counters to bytes")
    args = ap.parse_args()
    if args.workload == "c3":
        if args.gpus != 1:
            print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus})",
                  file=sys.stderr)
            sys.exit(2)
        import gls
        ctx = gls.Context(0)
        t0 = time.time()
        bt = batched_vs_plain(ctx, 256, 1.0)
        log(f"bench: batched vs plain {bt['speedup']}x ({time.time() - t0:.1f}s)")
        print(json.dumps({
            "metric": METRIC, "value": bt["batched_events_per_s"], "unit": "gate events/s",
            "n_gpus": 1, "steps": 1, "warmup": 1, "ms_per_step": round(bt["batched_s"] * 1e3, 3),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u32",
            "data": "synthetic",
            "config": {"workload": "batching: " + bt["workload"], "parallelism": "single GPU"},
            "batched": bt}), flush=True)
        return
    if args.workload == "c1":
        W = dict(config=1, seed=0x5EED0001, cells=args.cells or 200000, levels=40, pis=1024, gap=500.0,
                 t_end=args.t_end or 20000000)
    else:
        # the full 0-20 Mps window at gap 200 ps needs several TB of waveforms;
        # the netlist keeps its real shape and the window is shortened to fit HBM
        W = dict(config=2, seed=0x5EED0002, cells=args.cells or 2000000, levels=60, pis=4096,
                 gap=200.0, t_end=args.t_end or 500000)
        args.batch_sets = args.serve_jobs = 0
    args.cells, args.t_end = W["cells"], W["t_end"]
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "independent jobs would shard across GPUs with no communication.", file=sys.stderr)
        sys.exit(2)
    if args.copy_probe_mib:
        import torch
        n = args.copy_probe_mib << 20
        a = torch.ones(n, dtype=torch.uint8, device="cuda:0")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": n, "copies": 3}), flush=True)
        return

    import gls
    # CPU baseline first: the oracle runs as a child process before this
    # process touches the GPU
    cpu = None if args.no_cpu_baseline else cpu_baseline(args.cpu_events)
    if cpu:
        log(f"bench: cpu baseline {cpu['value']} gate events/s ({cpu['sample']})")
    t0 = time.time()
    d = gls.Design.generate(W["config"], W["seed"], W["cells"], W["levels"], W["pis"])
    s = gls.Stimulus.generate(d, W["seed"], W["t_end"], W["gap"], 0.01)
    info = d.info()
    log(f"bench: generated {info['n_instances']} cells / {info['n_levels']} levels, "
        f"{s.info()['n_events']} stimulus events in {time.time() - t0:.1f}s")
    ctx = gls.Context(0)
    path = gls.PATH_FUSED if args.path == "fused" else gls.PATH_TWO_PASS
    run = ctx.prepare(d, s, 0, args.t_end + 1, seg_events=args.seg_events, path=path,
                      level_blocks=args.level_blocks, prefetch=not args.no_prefetch,
                      hbm_stacks=args.hbm_stacks)
    log(f"bench: prepared (program + stimulus resident in HBM), "
        f"{run.stats()['level_blocks']} workgroups per level launch")
    for i in range(args.warmup):
        run.execute()
        log(f"bench: warmup {i} done ({run.stats()['kernel_ms']:.1f} ms device, "
            f"{run.stats()['n_events']} gate events)")
    walls, stats = [], []
    for i in range(args.steps):
        w0 = time.perf_counter()
        run.execute()
        walls.append(time.perf_counter() - w0)
        stats.append(run.stats())
        log(f"bench: step {i} wall {walls[-1] * 1e3:.1f} ms, device {stats[-1]['kernel_ms']:.1f} ms, "
            f"k_eval {stats[-1]['eval_ms']:.1f} ms, {stats[-1]['n_events']} gate events, "
            f"{stats[-1]['eval_launches']} level launches")
    st = stats[-1]
    ms_per_step = sum(walls) / len(walls) * 1e3
    value = st["n_events"] / (ms_per_step / 1e3)
    # dominant kernel: the level kernel (k_level fused / k_eval two-pass).
    # Algorithmic bytes = every input word read once per reading gate + every
    # output word written once; time = HIP events around each launch on the
    # job's stream.
    eval_ms = sum(x["eval_ms"] for x in stats) / len(stats)
    eval_bytes = 4 * (st["in_words"] + st["out_words"])
    achieved = eval_bytes / (eval_ms / 1e3) / 1e9
    launches = st["eval_launches"]
    traffic = measured_traffic(args, launches)
    # saif self-check: T0+T1+TX = DURATION for every record
    saif = run.saif_text()
    dur = args.t_end + 1
    ok = all(sum(int(x) for x in line.replace("(", " ").replace(")", " ").split()[1::2]) == dur
             for line in saif.splitlines() if line.strip().startswith("(T0"))
    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "gate events/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {"workload": (f"{info['n_instances']}-cell netlist, {info['n_inputs']} inputs, "
                                f"{info['n_levels']} levels"
                                + (", log-normal level widths" if W["config"] == 2 else "")
                                + f", 0-{args.t_end} ps window, mean toggle gap {W['gap']:g} ps, "
                                "Pareto(1.5) fanout; plain per-level HIP path"),
                   "cells": info["n_instances"], "levels": info["n_levels"], "inputs": info["n_inputs"],
                   "t_end_ps": args.t_end, "mean_gap_ps": W["gap"], "seed": hex(W["seed"]),
                   "path": "plain per-level (" + args.path + ")", "parallelism": "single GPU"},
        "gate_events": st["n_events"],
        "alg_bytes": st["alg_bytes"],
        "device_ms_per_step": round(sum(x["kernel_ms"] for x in stats) / len(stats), 3),
        "alg_hbm_gbs_level_loop": round(st["alg_bytes"] / (ms_per_step / 1e3) / 1e9, 1),
        "roofline": {"bound": "hbm", "kernel": "k_level" if args.path == "fused" else "k_eval",
                     "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic, "launches_per_step": launches,
                     "alg_bytes_per_launch": round(eval_bytes / max(1, launches)),
                     "ms_per_launch": round(eval_ms / max(1, launches), 3),
                     "alg_bytes_per_step": eval_bytes, "ms_per_step": round(eval_ms, 3)},
        "saif_durations_consistent": ok,
        "tiles_per_step": st["tiles"],
        "hbm_path_tile_frac": round(st["tiles_hbm"] / max(1, st["tiles"]), 4),
        "cpu_baseline": cpu,
    }
    # the same workload as one captured graph (no host round trip per level)
    if args.graph_steps:
        run.close()
        g = ctx.prepare(d, s, 0, args.t_end + 1, seg_events=args.seg_events,
                        level_blocks=args.level_blocks, prefetch=not args.no_prefetch, graph=True,
                        hbm_stacks=args.hbm_stacks)
        g.execute()  # eager run, captured
        gw = []
        for i in range(args.graph_steps):
            w0 = time.perf_counter()
            g.execute()
            gw.append(time.perf_counter() - w0)
        gst = g.stats()
        same = g.saif_text() == saif
        out["graph"] = {"ms_per_step": round(sum(gw) / len(gw) * 1e3, 3),
                        "events_per_s": round(gst["n_events"] / (sum(gw) / len(gw)), 1),
                        "alg_hbm_gbs": round(gst["alg_bytes"] / (sum(gw) / len(gw)) / 1e9, 1),
                        "replayed": gst["eval_launches"] == 1, "saif_same_as_plain": same}
        log(f"bench: graph replay {out['graph']['ms_per_step']} ms per step")
        run = g
    run.close()
    if args.batch_sets:
        t0 = time.time()
        out["batched"] = batched_vs_plain(ctx, args.batch_sets, args.batch_scale)
        log(f"bench: batched vs plain {out['batched']['speedup']}x ({time.time() - t0:.1f}s)")
    if args.serve_jobs:
        t0 = time.time()
        out["serving"] = serving_mix(ctx, args.serve_jobs, args.serve_scale)
        log(f"bench: serving p50 {out['serving']['p50_ms']} ms p99 {out['serving']['p99_ms']} ms "
            f"({time.time() - t0:.1f}s)")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
