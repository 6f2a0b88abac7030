#!/usr/bin/env python3
"""Benchmark of the device N-body step on one MI355X.

Default workload: direct-sum FP32 gravity, N = 1,048,576 particles (an
exponential disc of scale length 3 and scale height 0.3 on circular orbits
plus a Hernquist bulge holding 20% of the mass), the reference's force law
(law R, a = G m r_hat / (r^2 + S)) with S = eps^2 = 0.05^2, G = 1, dt = 5e-4,
kick-then-drift steps with per-target compensated accumulation.

Prints ONE JSON line.  `value` is whole-job interactions/s over the timed
steps (N^2 pair interactions per step, inputs resident in HBM).  `roofline`
prices the force kernel at 20 FLOP per interaction against the FP32 vector
peak, from HIP events around every force-kernel launch on the stream it runs
on.  `cpu_baseline` times the CPU restatement of the reference's
BruteForceCPU::Exec on a row sample of the same input on this host's cores.

    python bench.py [--gpus 1] [--steps 20] [--warmup 2] [--law R|P] [--n N]
                    [--no-cpu-baseline] [--no-energy] [--accuracy-rows 1024]
"""
import argparse
import importlib.util
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.join(ROOT, "procedural-universe_amd")

FP32_PEAK_TFLOPS = 157.3   # MI355X FP32 vector peak (AMD public spec)
# HBM-side bytes per launch of the fused direct kernel at N = 2^20 with the
# default shape, from rocprofv3 FETCH_SIZE and WRITE_SIZE scaled by a 256 MiB
# copy (profiles/direct_force_n1m_law_r_traffic.md).
DIRECT_1M_TRAFFIC_BYTES = 2.1835e9
# The same launch's SQ counters (profiles/direct_force_n1m_law_r_counters.md):
# VALU instructions per pair, VALU instructions issued per SIMD per cycle and
# rocprof's VALUBusy, at the 2.265 GHz the launch ran at.
DIRECT_1M_VALU = {"valu_instructions_per_pair": 13.30, "rsq_per_pair": 1.004, "valu_issue_per_simd_per_clock": 0.326,
                  "valu_lane_ops_per_clock_per_cu": 83.5, "valu_busy_pct": 140, "lds_wait_pct_of_wave_time": 1.0,
                  "clock_ghz": 2.265, "source": "rocprofv3 --pmc, profiles/direct_force_n1m_law_r_counters.md"}
# The same for the Barnes-Hut traversal step launch at N = 2^22, theta 0.5
# (profiles/bh_traverse_n4m_theta05_traffic.md) and the 512-galaxy batch step
# (profiles/batch_step_512_galaxies_traffic.md); they hold for the default
# workloads only.
BH_4M_TRAFFIC_BYTES = 3.2272e+09
BATCH_512_TRAFFIC_BYTES = 1.6415e9
FLOP_PER_INTERACTION = 20  # Nyland, Harris & Prins (GPU Gems 3, ch. 31) convention
METRIC = "interactions/s & % FP32 VALU peak at N=1M; batched p50/p99 step latency"


def load_package():
    name = "procedural_universe_amd"
    spec = importlib.util.spec_from_file_location(
        name, os.path.join(PKG_DIR, "__init__.py"), submodule_search_locations=[PKG_DIR]
    )
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


def load_oracle():
    """CPU restatement of the reference path: used only as the timed CPU baseline
    and as the accuracy checker, never as the measured GPU path."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from helpers import Oracle  # noqa: E402

    return Oracle()


def host_cores():
    # The GPU hosts share their CPUs: at most 16 threads per one-GPU job.
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    return max(1, min(16, n))


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def bench_batch(pu, args):
    """512 independent galaxies (log-uniform 1,024..65,536 particles, Plummer or
    exponential disc) advanced together by the persistent batch kernel in one
    hipGraph launch, against each galaxy alone with plain launches."""
    count, steps = args.galaxies, args.steps
    # the measured traffic belongs to the default workload only
    default_batch = count == 512 and args.seed == 0x5EED0004 and args.eps == 0.01 and args.law == "R"
    t0 = time.time()
    gal = pu.batch_galaxies(count, args.seed, 1024, 65536)
    t_ic = time.time() - t0
    total = sum(len(m) for (_, _, m) in gal)
    cap = sum((len(m) + 255) // 256 * 256 for (_, _, m) in gal)
    soft = args.eps * args.eps
    cfg = pu.default_config(law=pu.LAW_R if args.law == "R" else pu.LAW_P, softening=soft)
    b = pu.Batch(cfg, cap)
    for p, v, m in gal:
        b.add(p, v, m)
    inf = b.info()
    per_step = float(inf.interactions_per_step)
    tuning = b.tune(args.dt)
    if args.warmup:
        b.step(args.dt, args.warmup)
    b.synchronize()
    t = time.perf_counter()
    b.step(args.dt, steps)
    b.synchronize()
    elapsed = time.perf_counter() - t
    tm = b.timing()
    value = per_step * steps / elapsed
    # Baselines on a fresh copy: every galaxy alone, one launch per step.
    seq_steps = args.baseline_steps
    s = pu.Batch(cfg, cap)
    for p, v, m in gal:
        s.add(p, v, m)
    seq = {}
    for split_small in (False, True):
        s.synchronize()
        t = time.perf_counter()
        s.step_sequential(args.dt, seq_steps, split_small=split_small)
        s.synchronize()
        seq[split_small] = per_step * seq_steps / (time.perf_counter() - t)
    s.close()
    b.close()
    seq_value = seq[False]
    return {
        "metric": "batched aggregate interactions/s (512 galaxies, one persistent hipGraph launch)",
        "value": value,
        "unit": "interactions/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed * 1e3 / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"{count} galaxies, N log-uniform in [1024, 65536] (Plummer or exponential disc, 50/50), "
            f"direct-sum FP32, eps={args.eps}, dt={args.dt}",
            "galaxies": count,
            "particles": total,
            "interactions_per_step": per_step,
            "tuning": tuning,
            "graph_device_ms": tm.step_ms,
            "ic_seconds": t_ic,
        },
        "baseline": {
            "kind": "sequential per-galaxy plain launches (same kernel and arena, one workgroup per 256 targets, "
            "synchronised per galaxy)",
            "steps": seq_steps,
            "value": seq_value,
            "unit": "interactions/s",
            "speedup": value / seq_value,
            "tuned_single_system": {
                "kind": "sequential per-galaxy launches of the single-system default (sources split over more "
                "workgroups for small systems, plus a reduction launch)",
                "value": seq[True],
                "speedup": value / seq[True],
            },
        },
        "roofline": {
            "bound": "valu",
            "achieved": FLOP_PER_INTERACTION * value / 1e12,
            "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": FLOP_PER_INTERACTION * value / 1e12 / FP32_PEAK_TFLOPS,
            "traffic": BATCH_512_TRAFFIC_BYTES if default_batch else None,
            "traffic_source": "rocprofv3 FETCH_SIZE x2.0 + WRITE_SIZE x1.0 (256 MiB copy calibration), "
            "profiles/batch_step_512_galaxies_traffic.md" if default_batch else None,
        },
    }


def cpu_path_config(pu, oracle, cores, n=4096, steps=100, eps=0.01, dt=1e-3, seed=0x5EED0001):
    """The reference's CPU-path configuration timed on both sides: N = 4096
    equal-mass Plummer sphere (G = M = a = 1), law R with S = eps^2, FP32
    state, 100 kick-drift steps.  The GPU runs the library's direct sum; the
    CPU runs the restated BruteForceCPU::Update (FP64 velocities and sums) on
    `cores` threads over the same input.  Reports both rates, the energy drift
    of each and how far apart their final positions are."""
    soft = eps * eps
    pos, vel, mass = pu.ic_plummer(n, seed)
    with pu.NBodySim(pu.default_config(law=pu.LAW_R, softening=soft)) as sim:
        sim.init_soa(pos, vel, mass)
        a0 = sim.accelerations()
        e0 = sim.diagnostics()["total"]
        sim.step(dt, 1)  # warm-up launch, then the state is reset
        sim.init_soa(pos, vel, mass)
        sim.synchronize()
        t = time.perf_counter()
        sim.step(dt, steps)
        sim.synchronize()
        t_plain = time.perf_counter() - t
        p_gpu, v_gpu, _ = sim.read_soa()
        e1 = sim.diagnostics()["total"]
        # The same 100-step call replayed from a captured graph (one launch per
        # call instead of one or two per step): the first call captures it, the
        # second is timed; the trajectory simply continues.
        sim.set_graph(True)
        sim.step(dt, steps)
        sim.synchronize()
        t = time.perf_counter()
        sim.step(dt, steps)
        sim.synchronize()
        t_graph = time.perf_counter() - t
    t_gpu = min(t_plain, t_graph)
    t = time.perf_counter()
    p_cpu, v_cpu = oracle.ref_step(pos, vel.astype(np.float64), mass, dt, steps, G=1.0, S=soft, scale=1.0,
                                   workers=cores)
    t_cpu = time.perf_counter() - t
    ce0 = oracle.fp64_energy(0, pos, vel, mass, 1.0, soft)["total"]
    ce1 = oracle.fp64_energy(0, p_cpu, v_cpu.astype(np.float32), mass, 1.0, soft)["total"]
    a64 = oracle.fp64_accel(0, pos, mass, 1.0, soft, workers=cores)
    aerr = np.linalg.norm(a0 - a64, axis=1) / np.linalg.norm(a64, axis=1)
    perr = np.linalg.norm(p_gpu - p_cpu, axis=1) / np.linalg.norm(p_cpu, axis=1)
    inter = float(n) * n * steps
    return {
        "workload": f"N={n} Plummer sphere (G=M=a=1), law R, eps={eps}, dt={dt}, {steps} steps, FP32 state",
        "gpu_interactions_per_s": inter / t_gpu,
        "gpu_seconds": t_gpu,
        "gpu_seconds_plain_launches": t_plain,
        "gpu_seconds_graph_replay": t_graph,
        "cpu_interactions_per_s": inter / t_cpu,
        "cpu_seconds": t_cpu,
        "cpu_cores": cores,
        "cpu_kind": "port: restated BruteForceCPU::Update (FP32 geometry, FP64 velocities and sums)",
        "speedup": t_cpu / t_gpu,
        "accel_vs_fp64": {"rms_rel": float(np.sqrt(np.mean(aerr ** 2))), "max_rel": float(aerr.max())},
        "final_position_gpu_vs_cpu_max_rel": float(perr.max()),
        "energy_drift_gpu": float((e1 - e0) / abs(e0)),
        "energy_drift_cpu": float((ce1 - ce0) / abs(ce0)),
    }


def batch_step_latency(pu, steps=50, seed=0x5EED0004, eps=0.01, dt=1e-3, count=512):
    """Per-step latency of the batched workload: 512 galaxies (N log-uniform in
    [1024, 65536], Plummer or exponential disc), autotuned, then `steps` single
    steps, each timed on the batch's stream by the HIP events around its graph
    launch (nbody_batch_get_timing).  Returns p50/p99 of those step times and
    the aggregate interactions/s they give."""
    gal = pu.batch_galaxies(count, seed, 1024, 65536)
    cap = sum((len(m) + 255) // 256 * 256 for (_, _, m) in gal)
    cfg = pu.default_config(law=pu.LAW_R, softening=eps * eps)
    with pu.Batch(cfg, cap) as b:
        for p, v, m in gal:
            b.add(p, v, m)
        per_step = float(b.info().interactions_per_step)
        tuning = b.tune(dt)
        b.step(dt, 2)  # both parities' graphs captured before timing
        b.synchronize()
        ms = []
        for _ in range(steps):
            b.step(dt, 1)
            ms.append(b.timing().step_ms)
    ms = np.array(ms)
    return {
        "workload": f"{count} galaxies, N log-uniform in [1024, 65536] (Plummer or exponential disc, 50/50), "
        f"direct-sum FP32, eps={eps}, dt={dt}, one graph launch per step",
        "steps": steps,
        "p50_ms": float(np.percentile(ms, 50)),
        "p99_ms": float(np.percentile(ms, 99)),
        "mean_ms": float(ms.mean()),
        "max_ms": float(ms.max()),
        "interactions_per_step": per_step,
        "interactions_per_s": per_step / (ms.mean() * 1e-3),
        "timing": "HIP events around each step's graph launch on the batch's stream",
        "tuning": tuning,
    }


def bench_bh(pu, args):
    """Barnes-Hut, theta = 0.5, monopoles: two exponential-disc galaxies of
    N/2 particles on a parabolic encounter (separation 40, pericentre 5, second
    disc tilted 30 degrees)."""
    n = args.n
    soft = args.eps * args.eps
    default_bh = n == 1 << 22 and args.theta == 0.5 and args.seed == 0x5EED0003 and args.eps == 0.05
    t0 = time.time()
    pos, vel, mass = pu.ic_disc_encounter(n, args.seed, 3.0, 0.3, 40.0, 5.0, 30.0)
    t_ic = time.time() - t0
    cfg = pu.default_config(method=pu.METHOD_BARNES_HUT, law=pu.LAW_R, softening=soft, theta=args.theta)
    sim = pu.NBodySim(cfg)
    sim.init_soa(pos, vel, mass)
    sim.set_timing(True)
    acc0 = sim.accelerations()
    per_eval = sim.bh_stats()["interactions"]
    keys, order = sim.bh_tree_order()
    rng = np.random.default_rng(2024)
    rows = np.sort(rng.choice(n, size=min(args.subsample, n), replace=False))
    a_dir = sim.field(pos[rows])
    err = np.linalg.norm(acc0[rows].astype(np.float64) - a_dir, axis=1) / np.linalg.norm(a_dir, axis=1)
    if args.warmup:
        sim.step(args.dt, args.warmup)
    sim.synchronize()
    t = time.perf_counter()
    sim.step(args.dt, args.steps)
    sim.synchronize()
    elapsed = time.perf_counter() - t
    tm = sim.timing()
    stats = sim.bh_stats()
    interactions = float(stats["interactions"])
    value = interactions / elapsed
    force_ms = tm.force_ms / max(1, tm.force_launches)
    tree_ms = tm.tree_ms / max(1, tm.force_launches)
    result = {
        "metric": "Barnes-Hut interactions/s (counted cell + particle interactions), theta=0.5, N=4,194,304",
        "value": value,
        "unit": "interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"two exponential discs of {n // 2} particles each, parabolic encounter (separation 40, "
            f"pericentre 5, tilt 30 deg), Barnes-Hut theta={args.theta}, monopoles, 63-bit Morton keys, "
            f"FP64 cell mass/COM, eps={args.eps}, dt={args.dt}",
            "n": n,
            "theta": args.theta,
            "interactions_per_particle": interactions / (n * args.steps),
            "direct_equivalent_interactions_per_s": float(n) * float(n) * args.steps / elapsed,
            "tree_ms_per_step": tree_ms,
            "traversal_ms_per_step": force_ms,
            "stack_overflow": stats["stack_overflow"],
            "walk": {
                "list_entries_per_wave": 64.0 * stats["list_entries"] / (n * args.steps),
                "cells_opened_per_wave": 64.0 * stats["cells_opened"] / (n * args.steps),
                "lane_efficiency": interactions / (64.0 * max(1, stats["list_entries"])),
            },
            "ic_seconds": t_ic,
        },
        "accuracy_vs_direct": {
            "reference": "device direct sum (FP32 + compensated sums) at the sampled particles, same state",
            "rows": int(rows.size),
            "rms_rel": float(np.sqrt(np.mean(err ** 2))),
            "max_rel": float(err.max()),
        },
        "roofline": {
            "bound": "valu",
            "achieved": FLOP_PER_INTERACTION * interactions / (args.steps * force_ms * 1e-3) / 1e12,
            "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": FLOP_PER_INTERACTION * interactions / (args.steps * force_ms * 1e-3) / 1e12 / FP32_PEAK_TFLOPS,
            "traffic": BH_4M_TRAFFIC_BYTES if default_bh else None,
            "traffic_source": "rocprofv3 FETCH_SIZE x2.0 + WRITE_SIZE x1.0 (256 MiB copy calibration), "
            "profiles/bh_traverse_n4m_theta05_traffic.md" if default_bh else None,
            "kernel": "traverse_kernel",
        },
    }
    if not args.no_cpu_baseline:
        # CPU leg: the keys and the sorted order against the CPU computation on
        # the full N, and one step of the restated reference octree timed.
        oracle = load_oracle()
        lo, size = oracle.dynamic_root(pos)
        ref_keys, outside = oracle.morton_keys(pos, lo, size)
        ref_order = oracle.sort_order(ref_keys)
        cores = host_cores()
        t = time.perf_counter()
        oracle.bh_forces(pos, mass, args.theta, lo, size, G=1.0, S=soft, workers=cores)
        ct = time.perf_counter() - t
        result["keys_bit_exact"] = bool(np.array_equal(order, ref_order) and np.array_equal(keys, ref_keys[ref_order]))
        result["cpu_baseline"] = {
            "value": per_eval / ct,
            "unit": "interactions/s",
            "cores": cores,
            "kind": "port",
            "sample": f"one full force evaluation of the restated reference octree (Octree::Add/CalculateMass/"
            f"CalculateForce) on the same {n} particles ({ct:.1f} s, serial build), interactions counted on the device",
        }
    sim.close()
    return result


def bh_summary(pu):
    """The Barnes-Hut configuration in short (N = 4,194,304 two-disc encounter,
    theta 0.5, eps 0.05, 5 timed steps, no CPU leg), for the default line."""
    a = argparse.Namespace(n=1 << 22, seed=0x5EED0003, theta=0.5, eps=0.05, dt=5e-4, steps=5, warmup=1,
                           subsample=65536, no_cpu_baseline=True)
    r = bench_bh(pu, a)
    c = r["config"]
    return {
        "workload": c["workload"],
        "interactions_per_s": r["value"],
        "ms_per_step": r["ms_per_step"],
        "tree_ms_per_step": c["tree_ms_per_step"],
        "traversal_ms_per_step": c["traversal_ms_per_step"],
        "interactions_per_particle": c["interactions_per_particle"],
        "lane_efficiency": c["walk"]["lane_efficiency"],
        "traversal_frac_fp32_peak": r["roofline"]["frac"],
        "accuracy_vs_direct": r["accuracy_vs_direct"],
        "stack_overflow": c["stack_overflow"],
    }


def serve_summary(pu):
    """The served workload in short (the same Poisson stream and mix for 3 s of
    arrivals instead of 60, drained), for the default line: aggregate
    interactions/s and completion latency."""
    a = argparse.Namespace(seed=0x5EED0005, rate=200.0, duration=3.0, max_seconds=120.0, memory_gib=96.0, max_bh=16,
                           bh_ahead=1, progress="", seq_every=0, eps=0.02, dt=1e-3)
    r = bench_serve(pu, a)
    ld = r["load"]
    return {
        "workload": r["config"]["workload"],
        "interactions_per_s": r["value"],
        "latency_s": r["latency_s"],
        "requests": ld["offered_requests"],
        "completed": ld["completed"],
        "wall_seconds": ld["wall_seconds"],
        "note": "the full 60-s stream is `bench.py --workload serve`",
    }


def bench_serve(pu, args):
    """Served workload: Poisson arrivals (rate 200/s for `duration` s), sizes
    log-uniform in [2048, 524288], steps uniform in [10, 200]; below 32768
    particles direct sum, above Barnes-Hut theta = 0.6; eps = 0.02."""
    rng = np.random.default_rng(args.seed)
    arrivals = []
    t = 0.0
    while True:
        t += rng.exponential(1.0 / args.rate)
        if t > args.duration:
            break
        arrivals.append(t)
    m = len(arrivals)
    sizes = np.rint(2048 * (524288 / 2048) ** rng.random(m)).astype(np.int64)
    steps = rng.integers(10, 201, m)
    cfg = pu.server_config(law=pu.LAW_R, softening=args.eps ** 2, theta=0.6, bh_threshold=32768,
                           memory_budget=int(args.memory_gib * (1 << 30)), direct_capacity=64 << 20,
                           max_active_bh=args.max_bh, dt=args.dt, bh_steps_ahead=args.bh_ahead,
                           bh_no_graph=1 if getattr(args, "bh_no_graph", False) else 0)
    srv = pu.Server(cfg)
    offered_direct = float(sum(int(n) * int(n) * int(s) for n, s in zip(sizes, steps) if n < 32768))
    t0 = srv.clock()
    i = 0
    results = []
    timeline = []
    progress = open(args.progress, "a") if args.progress else None
    last_report = -1e9
    while True:
        now = srv.clock() - t0
        while i < m and arrivals[i] <= now:
            srv.submit_plummer(int(sizes[i]), int(steps[i]), args.seed * 1000003 + i, arrival=t0 + arrivals[i])
            i += 1
        st = srv.stats()
        if i == m and st["pending"] == 0 and st["active_direct"] + st["active_bh"] == 0:
            break
        if now > args.max_seconds:
            break
        if st["pending"] == 0 and st["active_direct"] + st["active_bh"] == 0:
            time.sleep(min(0.005, max(0.0, arrivals[i] - now)))
            continue
        srv.tick()
        results += srv.poll()
        if len(timeline) == 0 or now - timeline[-1][0] >= 1.0:
            timeline.append((now, st["pending"], st["active_direct"], st["active_bh"], st["memory_used"]))
        if now - last_report >= 2.0:
            last_report = now
            line = (f"serve t={now:.1f}s submitted={i}/{m} pending={st['pending']} active_direct={st['active_direct']} "
                    f"active_bh={st['active_bh']} completed={len(results)} ticks={st['ticks']}")
            print(line, file=sys.stderr, flush=True)
            if progress:
                progress.write(line + "\n")
                progress.flush()
    results += srv.poll()
    st = srv.stats()
    wall = srv.clock() - t0
    srv.close()
    seq = serve_sequential(pu, args, sizes, steps, progress) if args.seq_every > 0 else None
    lat = np.array([r["completed"] - r["arrival"] for r in results])
    direct = [r for r in results if r["method"] == pu.METHOD_DIRECT]
    bh = [r for r in results if r["method"] == pu.METHOD_BARNES_HUT]
    inter = float(sum(r["interactions"] for r in results))
    value = inter / wall
    pct = (lambda a, q: float(np.percentile(a, q)) if len(a) else None)
    return {
        "metric": "served aggregate interactions/s with p50/p99 completion latency (continuous batching)",
        "value": value,
        "unit": "interactions/s",
        "n_gpus": 1,
        "steps": int(st["ticks"]),
        "warmup": 0,
        "ms_per_step": 1e3 * st["busy_seconds"] / max(1, st["ticks"]),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"Poisson arrivals {args.rate}/s for {args.duration} s, N log-uniform [2048, 524288] "
            f"(Plummer spheres generated on the device), steps uniform [10, 200], direct below 32768 else "
            f"Barnes-Hut theta=0.6, eps={args.eps}, dt={args.dt}",
            "memory_budget_gib": args.memory_gib,
            "max_active_bh": args.max_bh,
            "bh_steps_ahead": args.bh_ahead,
        },
        "latency_s": {
            "p50": pct(lat, 50), "p99": pct(lat, 99),
            "p50_direct": pct([r["completed"] - r["arrival"] for r in direct], 50),
            "p99_direct": pct([r["completed"] - r["arrival"] for r in direct], 99),
            "p50_bh": pct([r["completed"] - r["arrival"] for r in bh], 50),
            "p99_bh": pct([r["completed"] - r["arrival"] for r in bh], 99),
        },
        "sequential_baseline": seq,
        "vs_sequential": (inter / wall) / seq["value"] if seq else None,
        "load": {
            "offered_requests": m,
            "submitted": int(st["submitted"]),
            "completed": len(results),
            "completed_direct": len(direct),
            "completed_bh": len(bh),
            "offered_direct_interactions": offered_direct,
            "completed_interactions": inter,
            "wall_seconds": wall,
            "memory_peak_gib": st["memory_peak"] / 2 ** 30,
            "timeline_s_pending_activeD_activeBH_mem": timeline[:: max(1, len(timeline) // 20)],
        },
    }


def serve_sequential(pu, args, sizes, steps, progress):
    """Every `seq_every`-th request of the served list run alone, one after
    another: one direct and one Barnes-Hut handle (reused, so no allocation per
    request), a device-generated Plummer sphere, all steps enqueued, then a
    synchronisation.  Interactions are counted as the server counts them (N^2
    per direct step, counted cell and particle interactions for Barnes-Hut)."""
    sel = np.arange(0, len(sizes), args.seq_every)
    soft = args.eps ** 2
    cfg_d = pu.default_config(law=pu.LAW_R, softening=soft)
    cfg_b = pu.default_config(method=pu.METHOD_BARNES_HUT, law=pu.LAW_R, softening=soft, theta=0.6)
    inter = 0.0
    with pu.NBodySim(cfg_d) as sd, pu.NBodySim(cfg_b) as sb:
        sb.set_timing(True)
        t0 = time.perf_counter()
        last = t0
        for k, i in enumerate(sel):
            n, ns = int(sizes[i]), int(steps[i])
            sim = sd if n < 32768 else sb
            sim.init_plummer(n, args.seed * 1000003 + int(i))
            sim.step(args.dt, ns)
            sim.synchronize()
            inter += float(n) * n * ns if sim is sd else float(sim.timing().interactions)
            now = time.perf_counter()
            if now - last >= 5.0:
                last = now
                line = f"serve sequential {k + 1}/{len(sel)} requests, {now - t0:.1f}s"
                print(line, file=sys.stderr, flush=True)
                if progress:
                    progress.write(line + "\n")
                    progress.flush()
        secs = time.perf_counter() - t0
    return {
        "value": inter / secs,
        "unit": "interactions/s",
        "requests": int(len(sel)),
        "seconds": secs,
        "kind": f"every {args.seq_every}th request of the same list alone and in turn (single-system handles, "
        "plain launches, device-generated initial conditions, one synchronisation per request)",
    }


def bench_copy(args):
    """Calibration for the HBM-traffic counters: three device-to-device copies
    of a known size through the HIP runtime's copy kernel."""
    import ctypes

    hip = ctypes.CDLL("libamdhip64.so")
    nbytes = 256 << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    if hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) or hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)):
        raise RuntimeError("hipMalloc failed")
    hip.hipMemset(a, 1, ctypes.c_size_t(nbytes))
    for _ in range(3):
        if hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3):  # hipMemcpyDeviceToDevice
            raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    return {"workload": "copy calibration", "bytes_per_copy": nbytes, "copies": 3}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--law", choices=["R", "P"], default="R")
    ap.add_argument("--n", type=int, default=1 << 20)
    ap.add_argument("--eps", type=float, default=0.05)
    ap.add_argument("--dt", type=float, default=5e-4)
    ap.add_argument("--seed", type=int, default=0x5EED0002)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-energy", action="store_true")
    ap.add_argument("--accuracy-rows", type=int, default=1024)
    ap.add_argument("--cpu-rows", type=int, default=0, help="rows of the CPU baseline sample (0 = auto)")
    ap.add_argument("--workload", choices=["direct", "batch", "bh", "serve", "copy"], default="direct")
    ap.add_argument("--rate", type=float, default=200.0)
    ap.add_argument("--duration", type=float, default=60.0)
    ap.add_argument("--max-seconds", type=float, default=300.0)
    ap.add_argument("--memory-gib", type=float, default=96.0)
    ap.add_argument("--max-bh", type=int, default=16, help="serve: concurrent Barnes-Hut requests")
    ap.add_argument("--bh-ahead", type=int, default=1, help="serve: Barnes-Hut steps enqueued per request at a time")
    ap.add_argument("--bh-no-graph", action="store_true",
                    help="serve: plain launches for Barnes-Hut requests instead of graph replay (for kernel traces)")
    ap.add_argument("--progress", default="", help="append serve progress lines to this file")
    ap.add_argument("--seq-every", type=int, default=40,
                    help="serve: run every k-th request alone as the sequential baseline (0 = skip)")
    ap.add_argument("--theta", type=float, default=0.5)
    ap.add_argument("--subsample", type=int, default=65536)
    ap.add_argument("--galaxies", type=int, default=512)
    ap.add_argument("--baseline-steps", type=int, default=10)
    ap.add_argument("--no-batched", action="store_true", help="direct: skip the batched step-latency leg")
    ap.add_argument("--no-bh", action="store_true", help="direct: skip the Barnes-Hut summary leg")
    ap.add_argument("--no-served", action="store_true", help="direct: skip the short served-stream leg")
    ap.add_argument("--hw-queues", type=int, default=16, choices=range(1, 33), metavar="1..32",
                    help="serve, direct: GPU_MAX_HW_QUEUES for this process, in place of the environment's "
                    "(the server's overlapping streams)")
    args = ap.parse_args()

    if args.workload == "copy":
        print(json.dumps(bench_copy(args)))
        return

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(
            "bench.py: the library measures one GPU only; --gpus must be 1 "
            "(multi-GPU sharding is a design, not a measured path)",
            file=sys.stderr,
        )
        sys.exit(2)

    if args.workload in ("serve", "direct"):
        # The server (the serve workload, and the default line's short served
        # leg) overlaps one stream per Barnes-Hut request with the batch's
        # stream.  With HIP's default of 4 hardware queues the direct
        # requests wait behind Barnes-Hut work (10-s stream: direct p50 1.3 s
        # against 0.5 s with 16, the same throughput and p99).  Read when the
        # HIP runtime starts, so it is set before the first device call.
        os.environ["GPU_MAX_HW_QUEUES"] = str(args.hw_queues)
    pu = load_package()
    if pu.device_count() < 1:
        print("bench.py: no HIP device visible", file=sys.stderr)
        sys.exit(3)

    if args.workload == "batch":
        if "--eps" not in sys.argv:
            args.eps = 0.01
        if "--dt" not in sys.argv:
            args.dt = 1e-3
        if "--steps" not in sys.argv:
            args.steps = 50
        if "--seed" not in sys.argv:
            args.seed = 0x5EED0004
        print(json.dumps(bench_batch(pu, args)))
        return

    if args.workload == "serve":
        if "--eps" not in sys.argv:
            args.eps = 0.02
        if "--dt" not in sys.argv:
            args.dt = 1e-3
        if "--seed" not in sys.argv:
            args.seed = 0x5EED0005
        print(json.dumps(bench_serve(pu, args)))
        return

    if args.workload == "bh":
        if "--n" not in sys.argv:
            args.n = 1 << 22
        if "--steps" not in sys.argv:
            args.steps = 10
        if "--seed" not in sys.argv:
            args.seed = 0x5EED0003
        print(json.dumps(bench_bh(pu, args)))
        return

    n = args.n
    law = pu.LAW_R if args.law == "R" else pu.LAW_P
    soft = args.eps * args.eps
    t_ic = time.time()
    pos, vel, mass = pu.ic_disc_bulge(n, args.seed, Rd=3.0, z0=0.3, bulge_fraction=0.2, ab=0.6)
    t_ic = time.time() - t_ic

    cfg = pu.default_config(law=law, softening=soft, G=1.0, scale=1.0)
    sim = pu.NBodySim(cfg)
    sim.init_soa(pos, vel, mass)
    sim.set_timing(True)

    acc0 = sim.accelerations()  # first force evaluation, kept for the accuracy check
    cores = host_cores()

    e0 = None if args.no_energy else sim.diagnostics()

    if args.warmup > 0:
        sim.step(args.dt, args.warmup)
    sim.synchronize()
    t0 = time.perf_counter()
    sim.step(args.dt, args.steps)
    sim.synchronize()
    elapsed = time.perf_counter() - t0
    timing = sim.timing()

    e1 = None if args.no_energy else sim.diagnostics()

    interactions = float(n) * float(n) * args.steps
    value = interactions / elapsed
    ms_per_step = elapsed * 1e3 / args.steps
    force_ms = timing.force_ms / max(1, timing.force_launches)
    achieved_tflops = FLOP_PER_INTERACTION * float(n) * float(n) / (force_ms * 1e-3) / 1e12

    result = {
        "metric": METRIC,
        "value": value,
        "unit": "interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "direct-sum FP32, N=1,048,576 exponential disc (Rd=3, z0=0.3) + 20% Hernquist bulge (a=0.6)"
            if n == 1 << 20
            else f"direct-sum FP32, N={n} disc+bulge",
            "n": n,
            "law": "R (reference: G m rhat/(r^2+S))" if law == pu.LAW_R else "P (Plummer)",
            "eps": args.eps,
            "dt": args.dt,
            "G": 1.0,
            "seed": args.seed,
            "kernel_shape": {"i_per_lane": timing.i_per_lane, "block": timing.block, "split": timing.split},
            "pct_fp32_peak": 100.0 * achieved_tflops / FP32_PEAK_TFLOPS,
        },
        "roofline": {
            "bound": "valu",
            "achieved": achieved_tflops,
            "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved_tflops / FP32_PEAK_TFLOPS,
            "traffic": DIRECT_1M_TRAFFIC_BYTES
            if (n == 1 << 20 and timing.i_per_lane == 1 and timing.block == 256 and timing.split == 1)
            else None,
            "traffic_source": "rocprofv3 FETCH_SIZE x2.0 + WRITE_SIZE x1.0 (256 MiB copy calibration), "
            "profiles/direct_force_n1m_law_r_traffic.md",
            "algorithmic_bytes": 80.0 * n,
            "kernel": "direct_force_kernel (fused kick-drift)",
            "kernel_ms": force_ms,
            "flop_per_interaction": FLOP_PER_INTERACTION,
            "valu_counters": DIRECT_1M_VALU
            if (n == 1 << 20 and timing.i_per_lane == 1 and timing.block == 256 and timing.split == 1)
            else None,
        },
    }
    if e0 is not None:
        result["energy"] = {
            "E0": e0["total"],
            "E1": e1["total"],
            "drift_rel": abs(e1["total"] - e0["total"]) / abs(e0["total"]),
            "steps": args.warmup + args.steps,
            "note": "kick-drift velocities are half a step from positions (reference K-D form)",
        }

    sim.close()
    if not args.no_batched and n == 1 << 20:
        # the metric's second half: batched per-step latency (512 galaxies)
        result["batched"] = batch_step_latency(pu)
    if not args.no_bh and n == 1 << 20:
        result["barnes_hut"] = bh_summary(pu)
    if not args.no_served and n == 1 << 20:
        result["served_short"] = serve_summary(pu)

    if not args.no_cpu_baseline:
        # CPU-baseline leg: the CPU restatement is the checker of the first
        # force evaluation (FP64 direct sum on a row sample) and the timed
        # reference CPU path (BruteForceCPU::Exec restated, full j-loops for a
        # subset of targets, scaled to interactions/s).
        oracle = load_oracle()
        rng = np.random.default_rng(12345)
        rows = np.sort(rng.choice(n, size=min(args.accuracy_rows, n), replace=False)).astype(np.int64)
        t_acc = time.time()
        a64 = oracle.fp64_accel(law, pos, mass, 1.0, soft, rows=rows, workers=cores)
        t_acc = time.time() - t_acc
        err = np.linalg.norm(acc0[rows].astype(np.float64) - a64, axis=1) / np.linalg.norm(a64, axis=1)
        result["accuracy"] = {
            "reference": "FP64 direct sum (CPU oracle), same law and inputs, first force evaluation",
            "rows": int(rows.size),
            "rms_rel": float(np.sqrt(np.mean(err ** 2))),
            "max_rel": float(err.max()),
            "oracle_seconds": t_acc,
        }
        nrows = args.cpu_rows or max(cores * 3072, 4096)  # about 10 s of CPU work
        crow = np.arange(nrows, dtype=np.int64) * (n // nrows)
        t = time.perf_counter()
        oracle.ref_forces(pos, mass, G=1.0, S=soft, rows=crow, workers=cores)
        ct = time.perf_counter() - t
        if n == 1 << 20:
            result["cpu_path_config"] = cpu_path_config(pu, oracle, cores)
        result["cpu_baseline"] = {
            "value": nrows * float(n) / ct,
            "unit": "interactions/s",
            "cores": cores,
            "kind": "port",
            "sample": f"{nrows} target rows x {n} sources of the same input ({ct:.1f} s), "
            f"C++ restatement of BruteForceCPU::Exec (FP32 geometry, FP64 accumulate) on {cpu_model()}",
        }
    print(json.dumps(result))


if __name__ == "__main__":
    main()
