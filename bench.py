"""Benchmark: particle-steps/s of the force-and-advance step on one MI355X.

Workload (the single-GPU configuration the metric is quoted on): a DPD fluid
of N = 1,048,576 particles at number density 3 in a periodic cube of side
70.44 with 70^3 cells, alpha = 25, gamma = 4.5, sigma = 3 with the
counter-based pair noise, cutoff 1, dt = 0.01, FP32 state and FP64
reductions.  A step is one engine_step: non-bonded DPD forces, advance, cell
migration and the cell-list rebuild, all on the GPU with the state resident
in HBM.

Printed: one JSON line with the throughput, the roofline of the dominant
kernel (the fused force/advance kernel, timed with HIP events on the
engine's stream) and the CPU baseline (this repository's restatement of the
reference's CPU step, multi-threaded, timed on a bounded sample of the same
workload on the host cores).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# Algorithmic bytes per particle-step for Newtonian / DPD particles:
# read x, cell, type, v (32 B), write and read f (24 B), write x, cell, v
# (28 B) = 84 B.  The fused kernel keeps f in registers, so it moves less.
BYTES_PER_PARTICLE_STEP = 84


def build_workload(args):
    import mechanica_amd as m
    from mechanica_amd import workloads as W
    if args.workload == "dpd":
        w = W.dpd_fluid(N=args.n, sigma=3.0, seed=2)
    elif args.workload == "lj":
        w = W.argon_lj(N=args.n, seed=1)
    else:
        raise SystemExit("unknown workload " + args.workload)
    u = W.build_universe(m, w, seed=1234)
    if args.workload == "dpd":
        u.set_dpd_noise(1)
    return m, w, u


def cpu_baseline(w, seconds, threads):
    """Times the oracle (CPU restatement of the reference step) with
    `threads` runners on the same workload, for about `seconds`."""
    import mechanica_amd as m
    from oracle import oracle as O
    o = O.from_workload(w, m, threads=threads, seed=1234)
    if w["name"] == "dpd_fluid":
        o.set_dpd_noise(1)
    o.step(1)  # warm-up
    N = len(w["x"])
    t0 = time.perf_counter()
    o.step(1)
    one = time.perf_counter() - t0
    steps = max(1, min(50, int(seconds / max(one, 1e-3)) - 1))
    t0 = time.perf_counter()
    o.step(steps)
    dt = time.perf_counter() - t0
    total_steps = steps + 1
    total_t = dt + one
    return {
        "value": N * total_steps / total_t,
        "unit": "particle-steps/s",
        "cores": threads,
        "kind": "port",
        "sample": "%s N=%d, %d timed steps after 1 warm-up step, %d threads, %.1f s" % (
            w["name"], N, total_steps, threads, total_t),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="dpd")
    ap.add_argument("--n", type=int, default=1048576)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    m, w, u = build_workload(args)
    N = len(w["x"])
    print("workload %s N=%d built" % (w["name"], N), file=sys.stderr, flush=True)

    u.step(args.warmup)
    u.sync()
    t0 = time.perf_counter()
    u.step(args.steps)
    u.sync()
    t1 = time.perf_counter()
    wall = t1 - t0
    value = N * args.steps / wall
    print("timed %d steps in %.3f s" % (args.steps, wall), file=sys.stderr, flush=True)

    # Kernel durations with HIP events on the engine's stream.
    u.set_timing(True)
    ksteps = min(args.steps, 100)
    u.step(ksteps)
    force_ms, step_ms = u.timing()
    u.set_timing(False)
    bytes_alg = BYTES_PER_PARTICLE_STEP * N
    achieved = bytes_alg / (force_ms * 1e-3) / 1e9

    threads = min(16, os.cpu_count() or 1)
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(w, args.cpu_seconds, threads)

    line = {
        "metric": "particle-steps/s (1M LJ/DPD)",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": ("dpd_fluid N=%d density=3 alpha=25 gamma=4.5 sigma=3 (counter RNG) cutoff=1 dt=0.01 "
                         "cells=%dx%dx%d" % (N, *w["cells"])) if w["name"] == "dpd_fluid" else
                        ("argon_lj N=%d density=10 LJ(0.275,1.0) tol=1e-3 cutoff=1 dt=0.01" % N),
            "particles": N,
            "cells": w["cells"],
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "kernel": "k_step (fused non-bonded force + advance)",
            "kernel_ms": force_ms,
            "step_ms_events": step_ms,
            "bytes_per_launch": bytes_alg,
        },
        "cpu_baseline": cpu,
    }
    if cpu:
        line["speedup_vs_cpu"] = value / cpu["value"]
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
