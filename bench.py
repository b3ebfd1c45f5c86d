"""Benchmark: gate waveform events/s of the plain per-level HIP path.

Headline workload (one "step" = one full simulation of every level over the
whole window, stimulus already resident in HBM): the 2,000,000-cell
configuration -- generic 12-cell library, 4,096 primary inputs, 60 levels
with log-normal widths, Pareto(1.5) fanout capped at 256, SDF delays
U{5..50} ps, Poisson stimulus with mean gap 200 ps over 0..20,000,000 ps,
1% X toggles (seed 0x5EED0002).  Its waveforms (1.15e12 gate events, about
4.6 TB) do not fit one GPU: the run streams them through time chunks.  Synthetic data from the
seeded generator.

Extras on the same JSON line, while the time budget lasts: the 200,000-cell
configuration (plain and one captured graph), the serving mix (Poisson jobs
of 10k-2M cells, 1-20 Mps windows, 30% netlist reuse), the stimulus ingest
of the 2M-cell configuration (its VCD read by the parallel reader, then the
pinned upload), and the batching configuration (256 stimulus sets on a
200k-cell netlist + 64 netlists of 5k-500k cells, batched vs plain with the
same tuned kernel settings).

The CPU baseline (the oracle, a serial restatement of the reference CPU
simulator) simulates the 2,000-cell configuration in full on one pinned
core, concurrently with the GPU work.

Prints one JSON line.  Only one GPU is used; --gpus N > 1 exits.
"""
import argparse
import json
import math
import os
import random
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "gate-level-simulator_amd"))

METRIC = "gate waveform events/sec, SAIF bit-exact; batched vs plain; p99 job latency"
HBM_PEAK_GBS = 8000.0
T0 = time.time()

WORKLOADS = {
    "c2": dict(config=2, seed=0x5EED0002, cells=2000000, levels=60, pis=4096, gap=200.0,
               t_end=20000000, shape="log-normal level widths"),
    "c1": dict(config=1, seed=0x5EED0001, cells=200000, levels=40, pis=1024, gap=500.0,
               t_end=20000000, shape="uniform level widths"),
}


def log(*a):
    print("bench [%.0fs]:" % (time.time() - T0), *a, file=sys.stderr, flush=True)


def host_info():
    model = ""
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in out.splitlines():
            if line.startswith("Model name"):
                model = line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return os.cpu_count() or 0, model


def start_cpu_baseline(tmp):
    """The oracle on the 2,000-cell configuration in full (0..2 Mps, mean gap
    1,000 ps; 18.05 M gate events) as a child process pinned to core 0.
    Started before this process touches the GPU; collected at the end."""
    import gls
    exe = os.path.join(ROOT, "oracle", "_build", "cpu_ref")
    if not os.path.exists(exe):
        return None
    d = gls.Design.generate(0, 0x5EED0000, 2000, 20, 64)
    s = gls.Stimulus.generate(d, 0x5EED0000, 2000000, 1000.0, 0.01)
    paths = [os.path.join(tmp, n) for n in ("c0.vlib", "c0.gv", "c0.sdf", "c0.vcd")]
    d.write(*paths[:3])
    s.write_vcd(paths[3])
    cmd = [exe] + paths + ["0", "2000001", os.path.join(tmp, "c0.saif")]
    pinned = False
    if hasattr(os, "sched_setaffinity") and (os.cpu_count() or 1) > 1:
        try:
            p = subprocess.Popen(["taskset", "-c", "0"] + cmd, stdout=subprocess.PIPE,
                                 stderr=subprocess.DEVNULL, text=True)
            pinned = True
        except OSError:
            p = subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.DEVNULL, text=True)
        # keep this process (and its host threads) off core 0
        os.sched_setaffinity(0, set(range(1, os.cpu_count())))
        os.environ.setdefault("GLS_HOST_THREADS", str(max(1, min(16, os.cpu_count() - 1))))
    else:
        p = subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.DEVNULL, text=True)
    return p, pinned


def finish_cpu_baseline(started):
    if not started:
        return None
    p, pinned = started
    out, _ = p.communicate(timeout=900)
    if p.returncode != 0:
        return None
    r = json.loads(out.strip().splitlines()[-1])
    ncpu, model = host_info()
    return {"value": round(r["events_per_s"], 1), "unit": "gate events/s", "cores": 1, "kind": "port",
            "sample": f"the 2,000-cell configuration in full: 20 levels, 64 inputs, 0..2,000,000 ps, "
                      f"mean gap 1,000 ps, {r['n_events']} gate events in {r['sim_seconds']:.1f} s "
                      f"(oracle/cpu_ref, 1 thread{', pinned to core 0' if pinned else ''}, run "
                      f"concurrently with the GPU work)",
            "host_nproc": ncpu, "host_cpu": model}


def profile_traffic(workload):
    """HBM bytes of the committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary of
    this command's workload (counters are read by a profiler run, not from
    inside this process): the dominant kernel per launch and the whole
    pipeline per step, with the file they come from."""
    name = {"c2": "k_level_2M_cells_hbm_counters.txt",
            "c1": "k_level_200k_cells_hbm_counters.txt"}[workload]
    p = os.path.join(ROOT, "profiles", name)
    vals = {}
    try:
        for line in open(p):
            if ":" in line and line.startswith("value "):
                k, v = line[6:].split(":", 1)
                vals[k.strip()] = float(v.split()[0])
    except OSError:
        return None
    if "k_level_bytes_per_launch" not in vals:
        return None
    return {"k_level_bytes_per_launch": vals["k_level_bytes_per_launch"],
            "pipeline_bytes_per_step": vals.get("pipeline_bytes_per_step"),
            "pipeline_over_alg": vals.get("pipeline_over_alg"),
            "source": os.path.relpath(p, ROOT)}


def saif_consistent(saif, duration):
    return all(sum(int(x) for x in line.replace("(", " ").replace(")", " ").split()[1::2]) == duration
               for line in saif.splitlines() if line.strip().startswith("(T0"))


def plain_steps(ctx, name, steps, warmup, args, graph_steps=0):
    """Generate a workload, prepare it (program + stimulus resident in HBM),
    run `warmup` untimed and `steps` timed simulations."""
    import gls
    W = WORKLOADS[name]
    cells = args.cells or W["cells"]
    t_end = args.t_end or W["t_end"]
    t0 = time.time()
    d = gls.Design.generate(W["config"], W["seed"], cells, W["levels"], W["pis"])
    s = gls.Stimulus.generate(d, W["seed"], t_end, W["gap"], 0.01)
    info = d.info()
    log(f"{name}: generated {info['n_instances']} cells / {info['n_levels']} levels, "
        f"{s.info()['n_events']} stimulus events in {time.time() - t0:.1f}s")
    run = ctx.prepare(d, s, 0, t_end + 1, seg_events=args.seg_events, level_blocks=args.level_blocks,
                      prefetch=not args.no_prefetch, hbm_stacks=args.hbm_stacks)
    log(f"{name}: prepared, {run.stats()['level_blocks']} workgroups per level launch")
    for i in range(warmup):
        run.execute()
        st = run.stats()
        log(f"{name}: warmup {i} {st['kernel_ms']:.1f} ms device, {st['chunks']} chunks")
    walls, stats = [], []
    for i in range(steps):
        w0 = time.perf_counter()
        run.execute()
        walls.append(time.perf_counter() - w0)
        stats.append(run.stats())
        log(f"{name}: step {i} wall {walls[-1] * 1e3:.1f} ms, k_level {stats[-1]['eval_ms']:.1f} ms, "
            f"{stats[-1]['n_events']} gate events, {stats[-1]['chunks']} chunks "
            f"({stats[-1]['chunk_retries']} run again shorter), {stats[-1]['eval_launches']} level launches")
    st = stats[-1]
    ms_per_step = sum(walls) / len(walls) * 1e3
    eval_ms = sum(x["eval_ms"] for x in stats) / len(stats)
    eval_bytes = 4 * (st["in_words"] + st["out_words"])
    launches = st["eval_launches"]
    res = {
        "workload": (f"{info['n_instances']}-cell netlist, {info['n_inputs']} inputs, "
                     f"{info['n_levels']} levels ({W['shape']}), 0-{t_end} ps window, mean toggle gap "
                     f"{W['gap']:g} ps, Pareto(1.5) fanout, seed {hex(W['seed'])}; plain per-level HIP path"),
        "cells": info["n_instances"], "levels": info["n_levels"], "inputs": info["n_inputs"],
        "t_end_ps": t_end, "mean_gap_ps": W["gap"], "seed": hex(W["seed"]),
        "gate_events": st["n_events"], "ms_per_step": round(ms_per_step, 3),
        "events_per_s": round(st["n_events"] / (ms_per_step / 1e3), 1),
        "device_ms_per_step": round(sum(x["kernel_ms"] for x in stats) / len(stats), 3),
        "alg_bytes": st["alg_bytes"], "time_chunks": st["chunks"], "chunk_retries": st["chunk_retries"],
        "pool_bytes": st["pool_bytes"], "out_words": st["out_words"],
        "k_level": {"launches_per_step": launches, "ms_per_step": round(eval_ms, 3),
                    "ms_per_launch": round(eval_ms / max(1, launches), 4),
                    "alg_bytes_per_step": eval_bytes,
                    "alg_bytes_per_launch": round(eval_bytes / max(1, launches)),
                    "achieved_gbs": round(eval_bytes / (eval_ms / 1e3) / 1e9, 1)},
        "tiles_per_step": st["tiles"], "hbm_path_tile_frac": round(st["tiles_hbm"] / max(1, st["tiles"]), 4),
        "saif_durations_consistent": saif_consistent(run.saif_text(), t_end + 1),
    }
    plain_saif = run.saif_text() if graph_steps else None
    run.close()  # its buffers (the event pool) go back to the context for the graph run
    if graph_steps:
        g = ctx.prepare(d, s, 0, t_end + 1, seg_events=args.seg_events, level_blocks=args.level_blocks,
                        prefetch=not args.no_prefetch, graph=True, hbm_stacks=args.hbm_stacks)
        g.execute()  # eager run, captured
        gw = []
        for i in range(graph_steps):
            w0 = time.perf_counter()
            g.execute()
            gw.append(time.perf_counter() - w0)
        gst = g.stats()
        res["graph"] = {"ms_per_step": round(sum(gw) / len(gw) * 1e3, 3),
                        "events_per_s": round(gst["n_events"] / (sum(gw) / len(gw)), 1),
                        "time_chunks": gst["chunks"], "replayed": gst["eval_launches"] == 1,
                        "saif_same_as_plain": g.saif_text() == plain_saif}
        g.close()
    return res


def batched_vs_plain(ctx, n_sets, n_nets, scale, plain_default=False):
    """Batching workload: one netlist from the 200k-cell generator at `scale`
    of its size driven by `n_sets` seeded stimulus sets (0-2 Mps windows,
    mean gap U[100, 5000] ps), plus `n_nets` netlists with log-uniform cell
    counts (5k-500k x scale) and depths U{10..60}, one stimulus set each.
    Batched: every instance in one prepared run (level l of all instances in
    one launch; instance groups that each fit the event pool, time chunks
    inside a group only as HBM requires), the whole group, chunk and level
    sequence captured as one hipGraph, kernel settings autotuned on the batch.
    Plain: every instance prepared and executed alone, one after the other,
    with the same tuned settings (and, with plain_default, separately with
    the default settings).
    Every instance's SAIF is compared with its plain run."""
    import gls
    rng = random.Random(0x5EED0003)
    cells = max(1000, int(200000 * scale))
    main_d = gls.Design.generate(1, 0x5EED0003, cells, 40, max(16, int(1024 * scale)))
    insts = []
    for i in range(n_sets):
        s = gls.Stimulus.generate(main_d, 0x5EED0003 + 1 + i, 2000000, rng.uniform(100, 5000), 0.01)
        insts.append((main_d, s))
    log(f"c3: {n_sets} stimulus sets generated")
    for i in range(n_nets):
        c = max(500, int(round(10 ** rng.uniform(3.699, 5.699) * scale)))
        d = gls.Design.generate(1, 0x5EED1003 + i, c, rng.randint(10, 60), max(16, int(1024 * c / 200000)))
        s = gls.Stimulus.generate(d, 0x5EED2003 + i, 2000000, rng.uniform(100, 5000), 0.01)
        insts.append((d, s))
    log(f"c3: {len(insts)} instances generated")
    n = len(insts)
    # the autotuner runs the batch once per candidate setting: it tunes on a
    # representative eighth of the instances (every eighth stimulus set and
    # netlist), and the full batch runs with what it picked
    sub = insts[0:n_sets:8] + insts[n_sets::8]
    t = ctx.prepare_batch([x[0] for x in sub], [x[1] for x in sub], [0] * len(sub),
                          [2000001] * len(sub))
    log(f"c3: tuning on {len(sub)} instances")
    tune = t.autotune()
    t.close()
    knobs = {"seg_events": tune["seg_events"], "level_blocks": tune["level_blocks"],
             "prefetch": bool(tune["prefetch"])}
    log(f"c3: autotune picked {knobs}")
    b = ctx.prepare_batch([x[0] for x in insts], [x[1] for x in insts], [0] * n, [2000001] * n,
                          graph=True, **knobs)
    b.execute()  # eager run, captured as one graph
    t0 = time.perf_counter()
    b.execute()  # graph replay: every chunk and level of every instance, no host round trip
    batch_s = time.perf_counter() - t0
    bst = b.stats()
    batch_saif = [b.instance_saif_text(k) for k in range(n)]
    b.close()
    log(f"c3: batched {batch_s * 1e3:.1f} ms, {bst['n_events']} gate events, {bst['chunks']} chunks")

    def plain(opts):
        total, ev, saifs = 0.0, 0, []
        for k, (d, s) in enumerate(insts):
            r = ctx.prepare(d, s, 0, 2000001, **opts)
            r.execute()  # the netlist image and pool are set up by the first run
            t0 = time.perf_counter()
            r.execute()
            total += time.perf_counter() - t0
            ev += r.stats()["n_events"]
            saifs.append(r.saif_text())
            r.close()
        return total, ev, saifs

    tuned_s, tuned_ev, tuned_saif = plain(knobs)
    log(f"c3: plain (tuned settings) {tuned_s * 1e3:.1f} ms")
    default = {"seg_events": 32, "level_blocks": 0, "prefetch": True}
    if knobs == default or tune["level_blocks"] == 0:
        dflt_s = tuned_s
    elif plain_default:
        dflt_s, _, _ = plain({})
        log(f"c3: plain (default settings) {dflt_s * 1e3:.1f} ms")
    else:
        dflt_s = None
    return {"workload": f"{n_sets} stimulus sets on a {cells}-cell netlist + {n_nets} netlists "
                        f"(log-uniform {int(5000 * scale)}-{int(500000 * scale)} cells, depth 10-60), "
                        f"0-2 Mps windows, gap U[100,5000] ps",
            "instances": n, "gate_events": bst["n_events"], "time_chunks": bst["chunks"],
            "batched_s": round(batch_s, 4), "batched_events_per_s": round(bst["n_events"] / batch_s, 1),
            "plain_tuned_s": round(tuned_s, 4),
            "plain_tuned_events_per_s": round(tuned_ev / tuned_s, 1),
            "plain_default_s": round(dflt_s, 4) if dflt_s else None,
            "speedup": round(tuned_s / batch_s, 3),
            "speedup_vs_default_plain": round(dflt_s / batch_s, 3) if dflt_s else None,
            "saif_bit_exact_vs_plain": batch_saif == tuned_saif,
            "events_same_as_plain": bst["n_events"] == tuned_ev,
            "form": "one captured hipGraph of every instance group's chunks and levels"
                    if bst["eval_launches"] == 1 else "level-synchronous launches",
            "autotune": {**{k: tune[k] for k in ("seg_events", "level_blocks", "prefetch", "best_ms")},
                         "tuned_on_instances": len(sub)}}


def serving_mix(ctx, n_jobs, scale, max_batch=4):
    """Serving loop on a seeded Poisson job stream: netlists log-uniform in
    [10k, 2M] x `scale` cells (the same generator), windows U[1, 20] Mps x
    `scale`, 30% of jobs reuse a netlist already submitted.  One server (4
    jobs at once, batches of up to `max_batch` queued jobs admitted
    together).  Every job is first run alone on it, one job at a time (so
    with the same per-job pool limit and cached device buffers as in the
    mix); the arrival rate is set from the mean alone time for about 75%
    device utilisation.  Reports the mix's gate events/s (its events over
    the time from its first submission to its last completion) and p50/p99
    job latency, and compares every job's SAIF in the mix with its alone
    run."""
    import gls
    rng = random.Random(0x5EED0004)
    designs, jobs = [], []
    for i in range(n_jobs):
        if designs and rng.random() < 0.3:
            d = rng.choice(designs)
        else:
            c = max(500, int(10 ** rng.uniform(4, 6.301) * scale))
            d = gls.Design.generate(1, 0x5EED4000 + i, c, rng.randint(10, 60),
                                    max(16, int(1024 * c / 200000)))
            designs.append(d)
        t_end = int(rng.uniform(1e6, 2e7) * scale)
        s = gls.Stimulus.generate(d, 0x5EED5000 + i, t_end, 500.0, 0.01)
        jobs.append((d, s, 0, t_end + 1))
    log(f"c4: {n_jobs} jobs generated, {len(designs)} netlists")
    srv = gls.Server(ctx, max_running=4, max_batch=max_batch)
    # alone: one job at a time (each job run twice, the second time with its
    # netlist resident and the buffers cached, as in the mix)
    alone, t_alone = {}, []
    for k in range(n_jobs):
        d, s, on, off = jobs[k]
        t0 = time.perf_counter()
        j = srv.submit(d, s, on, off)
        srv.wait(j)
        alone[k] = srv.saif_text(j)
        t1 = time.perf_counter()
        j = srv.submit(d, s, on, off)
        srv.wait(j)
        t_alone.append(time.perf_counter() - t1)
        log(f"c4: alone {k}: {d.info()['n_instances']} cells, 0-{off - 1} ps: first {1e3 * (t1 - t0):.0f} ms, "
            f"resident {1e3 * t_alone[-1]:.0f} ms")
    mean_s = sum(t_alone) / len(t_alone)
    rate = 0.75 / mean_s
    gaps = [rng.expovariate(rate) for _ in jobs]
    ids = []
    t_first = time.perf_counter()
    for (d, s, on, off), g in zip(jobs, gaps):
        time.sleep(g)
        ids.append(srv.submit(d, s, on, off))
    ok = True
    for k, j in enumerate(ids):
        ok = srv.wait(j) == 0 and ok
        log(f"c4: mix job {k} done")
    t_last = time.perf_counter()
    infos = [srv.info(j) for j in ids]
    same = all(srv.saif_text(ids[k]) == alone[k] for k in alone)
    st = srv.stats()
    srv.close()
    lat = sorted(i["latency_ms"] for i in infos)
    pct = lambda q: lat[min(len(lat) - 1, int(math.ceil(q * len(lat))) - 1)]
    events = sum(i["n_events"] for i in infos)
    return {"max_batch": max_batch,
            "workload": f"{n_jobs} Poisson jobs at {rate:.2f}/s, cells log-uniform in "
                        f"[{int(1e4 * scale)}, {int(2e6 * scale)}], windows U[{int(1e6 * scale)}, "
                        f"{int(2e7 * scale)}] ps, gap 500 ps, 30% netlist reuse",
            "jobs": len(ids), "failed": sum(1 for i in infos if i["status"] != 0), "all_ok": ok,
            "gate_events": events, "mix_s": round(t_last - t_first, 3),
            "events_per_s": round(events / (t_last - t_first), 1),
            "p50_ms": round(pct(0.50), 3), "p99_ms": round(pct(0.99), 3),
            "max_ms": round(lat[-1], 3), "mean_alone_ms": round(mean_s * 1e3, 3),
            "peak_running": int(st["peak_running"]), "resident_programs": int(st["resident_programs"]),
            "device_buffers_reused": int(st["arena_hits"]), "device_buffers_allocated": int(st["arena_misses"]),
            "saif_same_alone_and_in_mix": same, "jobs_checked_alone": len(alone)}


def vcd_ingest(ctx, tmp):
    """Stimulus ingest of the 2M-cell configuration: its VCD written to a
    file (4,096 inputs, 0-20 Mps, gap 200 ps), read back by the parallel
    reader (line rate), and uploaded through the pinned staging buffers
    (the prepare of a netlist already resident).  Falls back to the
    200k-cell stimulus when the scratch disk is short of room."""
    import shutil
    import gls
    name = "c2"
    if shutil.disk_usage(tmp).free < (12 << 30):
        name = "c1"
    W = WORKLOADS[name]
    d = gls.Design.generate(W["config"], W["seed"], W["cells"], W["levels"], W["pis"])
    s = gls.Stimulus.generate(d, W["seed"], W["t_end"], W["gap"], 0.01)
    path = os.path.join(tmp, "stim.vcd")
    t0 = time.perf_counter()
    s.write_vcd(path)
    t_write = time.perf_counter() - t0
    size = os.path.getsize(path)
    s.close()
    t0 = time.perf_counter()
    s2 = gls.Stimulus.from_vcd(path)
    t_parse = time.perf_counter() - t0
    os.remove(path)
    r = ctx.prepare(d, s2, 0, W["t_end"] + 1)  # program image built and made resident
    r.close()
    t0 = time.perf_counter()
    r = ctx.prepare(d, s2, 0, W["t_end"] + 1)
    t_prep = time.perf_counter() - t0
    r.close()
    n = s2.info()["n_events"]
    log(f"ingest: {size / 1e9:.2f} GB VCD read in {t_parse:.2f} s, prepare {t_prep:.2f} s")
    return {"workload": f"{W['cells']}-cell configuration's stimulus, {W['pis']} inputs, "
                        f"0-{W['t_end']} ps, gap {W['gap']} ps",
            "vcd_bytes": size, "stimulus_events": n, "write_s": round(t_write, 3),
            "parse_s": round(t_parse, 3), "parse_gb_per_s": round(size / t_parse / 1e9, 3),
            "prepare_s": round(t_prep, 3),
            "upload_gb_per_s": round(8 * n / t_prep / 1e9, 3)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=["c2", "c1"], default="c2",
                    help="headline: c2 (2M cells, full 0-20 Mps window, default) or c1 (200k cells)")
    ap.add_argument("--cells", type=int, default=None)
    ap.add_argument("--t-end", type=int, default=None)
    ap.add_argument("--seg-events", type=int, default=0)
    ap.add_argument("--level-blocks", type=int, default=0)
    ap.add_argument("--no-prefetch", action="store_true")
    ap.add_argument("--hbm-stacks", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--extras", default="c1,c4,ingest,c3",
                    help="comma list of extra measurements after the headline (c1, c3, c4, "
                         "ingest; '' none)")
    ap.add_argument("--budget-s", type=float, default=600.0,
                    help="extras start only while the elapsed time leaves room for them")
    ap.add_argument("--c3-scale", type=float, default=1.0)
    ap.add_argument("--c3-plain-default", action="store_true",
                    help="batching mix: also time the plain runs with the default kernel settings")
    ap.add_argument("--c4-jobs", type=int, default=24)
    ap.add_argument("--c4-scale", type=float, default=1.0)
    ap.add_argument("--c4-batch", type=int, default=4,
                    help="serving mix: queued jobs admitted together run as one batch of up to this many")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run a plain device copy of this many MiB (3 times): the "
                         "reference point that scales FETCH_SIZE/WRITE_SIZE counters to bytes")
    args = ap.parse_args()
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
    tmp = tempfile.mkdtemp(prefix="glsbench")
    # the CPU baseline starts before this process touches the GPU
    cpu_run = None if args.no_cpu_baseline else start_cpu_baseline(tmp)
    ctx = gls.Context(0)
    head = plain_steps(ctx, args.workload, args.steps, args.warmup, args)
    kl = head["k_level"]
    out = {
        "metric": METRIC,
        "value": head["events_per_s"],
        "unit": "gate events/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {"workload": head["workload"], "cells": head["cells"], "levels": head["levels"],
                   "inputs": head["inputs"], "t_end_ps": head["t_end_ps"],
                   "mean_gap_ps": head["mean_gap_ps"], "seed": head["seed"],
                   "path": "plain per-level (fused), time-chunked", "parallelism": "single GPU"},
        "gate_events": head["gate_events"],
        "alg_bytes": head["alg_bytes"],
        "device_ms_per_step": head["device_ms_per_step"],
        "alg_hbm_gbs_step": round(head["alg_bytes"] / (head["ms_per_step"] / 1e3) / 1e9, 1),
        "time_chunks": head["time_chunks"],
        "chunk_retries": head["chunk_retries"],
        "roofline": {"bound": "hbm", "kernel": "k_level", "achieved": kl["achieved_gbs"],
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(kl["achieved_gbs"] / HBM_PEAK_GBS, 4),
                     "traffic": None, "launches_per_step": kl["launches_per_step"],
                     "alg_bytes_per_launch": kl["alg_bytes_per_launch"],
                     "ms_per_launch": kl["ms_per_launch"],
                     "alg_bytes_per_step": kl["alg_bytes_per_step"], "ms_per_step": kl["ms_per_step"]},
        "saif_durations_consistent": head["saif_durations_consistent"],
        "tiles_per_step": head["tiles_per_step"],
        "hbm_path_tile_frac": head["hbm_path_tile_frac"],
    }
    tr = profile_traffic(args.workload) if not args.cells and not args.t_end else None
    if tr:
        out["roofline"]["traffic"] = round(tr["k_level_bytes_per_launch"])
        out["roofline"]["traffic_source"] = tr["source"]
        out["pipeline_counter_bytes_per_step"] = tr["pipeline_bytes_per_step"]
        out["pipeline_counter_over_alg"] = tr["pipeline_over_alg"]
    extras = [x for x in args.extras.split(",") if x]
    # rough costs of each extra (s), to stay inside the budget
    cost = {"c1": 35.0, "c3": 280.0 * args.c3_scale + 20, "c4": 2.0 * args.c4_jobs * args.c4_scale + 10,
            "ingest": 60.0}
    skipped = []
    for x in extras:
        if time.time() - T0 + cost.get(x, 0) > args.budget_s:
            skipped.append(x)
            log(f"{x}: skipped (time budget)")
            continue
        t0 = time.time()
        if x == "c1" and args.workload != "c1":
            out["c1"] = plain_steps(ctx, "c1", 2, 1, argparse.Namespace(**{**vars(args), "cells": None,
                                                                        "t_end": None}), graph_steps=1)
        elif x == "c3":
            out["batched"] = batched_vs_plain(ctx, max(1, int(256 * min(1.0, args.c3_scale))),
                                              max(1, int(64 * min(1.0, args.c3_scale))), args.c3_scale,
                                              args.c3_plain_default)
        elif x == "c4":
            out["serving"] = serving_mix(ctx, args.c4_jobs, args.c4_scale, args.c4_batch)
        elif x == "ingest":
            out["ingest"] = vcd_ingest(ctx, tmp)
        log(f"{x}: done in {time.time() - t0:.1f}s")
    if skipped:
        out["extras_skipped_for_time"] = skipped
    out["cpu_baseline"] = finish_cpu_baseline(cpu_run)
    if out["cpu_baseline"]:
        log(f"cpu baseline {out['cpu_baseline']['value']} gate events/s")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
