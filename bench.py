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
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(
            "bench.py: this build measures one GPU only; --gpus must be 1 "
            "(multi-GPU sharding is a design, not a measured path)",
            file=sys.stderr,
        )
        sys.exit(2)

    pu = load_package()
    if pu.device_count() < 1:
        print("bench.py: no HIP device visible", file=sys.stderr)
        sys.exit(3)

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
            "traffic": None,
            "kernel": "direct_force_kernel (fused kick-drift)",
            "kernel_ms": force_ms,
            "flop_per_interaction": FLOP_PER_INTERACTION,
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
        nrows = args.cpu_rows or max(cores * 512, 1024)
        crow = np.arange(nrows, dtype=np.int64) * (n // nrows)
        t = time.perf_counter()
        oracle.ref_forces(pos, mass, G=1.0, S=soft, rows=crow, workers=cores)
        ct = time.perf_counter() - t
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
