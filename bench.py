"""Benchmark: particle-steps/s of the force-and-advance step on one MI355X.

Default workload (the single-GPU configuration the metric is quoted on): a
DPD fluid of N = 1,048,576 particles at number density 3 in a periodic cube
of side 70.44 with 70^3 cells, alpha = 25, gamma = 4.5, sigma = 3 with the
counter-based pair noise, cutoff 1, dt = 0.01, FP32 state and FP64
reductions.  A step is one engine_step: non-bonded DPD forces, advance, cell
migration and the cell-list rebuild, all on the GPU with the state resident
in HBM.

Other workloads (--workload): lj (argon LJ at density 10, 1M particles),
lattice (64^3 bonded Morse lattice, overdamped), molecules (64^3 four-bead
chains with bonds, angles, dihedrals and exclusions in an LJ fluid, 1M
particles), channel (examples/dpd1.py at 1M particles: DPD between no-slip
walls under a body force), species (the 1M DPD fluid carrying three species
with five fluxes), ensemble (256 cell-sorting
replicas in one engine, compared with one universe of the same total size)
and adjoint (65,536-particle Morse trajectory of 50 steps, forward and
backward, compared with the plain step on the same system).

Printed: one JSON line with the throughput (the median of five timed runs
of --steps steps each), the roofline of the dominant kernel (timed with HIP
events on the stream it runs on: the HBM fraction, and beside it the FP32
rate of the algorithmic flops and the VALU issue fraction from the measured
instruction count) and the CPU baseline (this repository's restatement of
the reference's CPU step, multi-threaded on every core the job may use,
timed on a bounded sample of the same workload).  The default run also
measures 1M-particle argon LJ, the other half of the metric, and reports it
under "lj" with its own roofline and CPU baseline.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
SIMDS = 1024          # 256 CUs x 4 SIMDs
CLOCK_GHZ = 2.4       # peak engine clock
# Algorithmic FP32 flops (add, mul, sqrt, div count 1) per unique in-cutoff
# pair and per particle advance, from the reference's operation sequence:
# table pair 45, DPD pair 39, Newtonian advance 21, overdamped 15.
FLOPS_TABLE_PAIR = 45
FLOPS_DPD_PAIR = 39
FLOPS_ADVANCE = {"newtonian": 21, "overdamped": 15}
# Algorithmic bytes per particle-step: read x, cell, type (20 B) and v
# (12 B) for Newtonian or DPD particles, write and read f (24 B), write x
# and cell (16 B) and v (12 B): 84 B; overdamped particles have no v: 60 B.
# Bonds add 12 B each (two indices and a potential id).
BYTES_NEWTONIAN = 84
BYTES_OVERDAMPED = 60
BYTES_PER_BOND = 12
# Angles, dihedrals and exclusions reach the force kernel as one 16-byte
# entry per role in the particle's term list (plus a 4-byte list offset per
# particle); the term kernel that writes them is profiled on its own.
BYTES_PER_TERM_ENTRY = 16
BYTES_TERM_OFFSET = 4

# HBM traffic per force-kernel launch measured with rocprofv3 --pmc
# FETCH_SIZE and WRITE_SIZE on the 1M DPD and LJ workloads at the brick the
# autotuner has picked for them ([5,5,3] list 12 and [6,4,1] list 16), scaled by
# the copy-kernel calibration (bench.py --copy-calibration under the same
# counters); the summaries and their commands are kept in
# profiles/k_brick_dpd1m_counters.md and profiles/k_brick_lj1m_counters.md
# (bytes per launch).
MEASURED_TRAFFIC = {("dpd", 1048576): 130.4e6, ("lj", 1048576): 139.1e6}
# Wave64 VALU instructions per force-kernel launch (rocprofv3 --pmc
# SQ_INSTS_VALU on the same workloads, bricks and build;
# profiles/k_brick_lanes_dpd1m_lj1m.md).  A
# wave64 FP32 instruction holds a 16-lane SIMD for 4 cycles.
MEASURED_VALU = {("dpd", 1048576): 141.5e6, ("lj", 1048576): 455.1e6}
MEASURED_BRICK = {("dpd", 1048576): [[5, 5, 3], 12], ("lj", 1048576): [[6, 4, 1], 16]}

METRIC = "particle-steps/s (1M LJ/DPD)"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def timed_steps(u, steps, warmup, graphs=True, runs=5):
    """Wall time of `steps` steps: the median of `runs` timed runs (each
    bracketed by a synchronisation), and the list of all of them."""
    u.set_graphs(graphs)
    u.step(warmup)
    u.sync()
    walls = []
    for _ in range(runs):
        t0 = time.perf_counter()
        u.step(steps)
        u.sync()
        walls.append(time.perf_counter() - t0)
    return sorted(walls)[len(walls) // 2], walls


def kernel_ms(u, steps):
    """Mean force-kernel and step time (ms) from HIP events on the engine's
    stream."""
    u.set_timing(True)
    u.step(min(steps, 100))
    f, s = u.timing()
    u.set_timing(False)
    return f, s


def roofline(bytes_alg, ms, kernel, traffic=None, flops=None, valu=None):
    """HBM roofline of the kernel; flops (algorithmic FP32 per launch) and
    valu (measured VALU instructions per launch) add the FP32 rate against
    the vector peak and the VALU issue fraction."""
    achieved = bytes_alg / (ms * 1e-3) / 1e9
    out = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel": kernel, "kernel_ms": ms,
           "bytes_per_launch": bytes_alg}
    if flops is not None:
        tf = flops / (ms * 1e-3) / 1e12
        out["fp32"] = {"flops_per_launch": flops, "achieved": tf, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                       "frac": tf / FP32_PEAK_TFLOPS}
    if valu is not None:
        busy = valu * 4 / (SIMDS * CLOCK_GHZ * 1e9 * ms * 1e-3)
        out["valu"] = {"insts_per_launch": valu, "issue_frac": busy,
                       "note": "wave64 VALU instructions x 4 cycles / (1024 SIMDs x 2.4 GHz x kernel time); "
                               "instruction count measured with rocprofv3 on this workload"}
    return out


def host_threads():
    """Threads for the CPU baseline: every core the job may use.  The GPU
    hosts grant a job a share of the machine through OMP_NUM_THREADS; the
    affinity mask can show more cores than that share."""
    avail = len(os.sched_getaffinity(0))
    share = os.environ.get("OMP_NUM_THREADS")
    return min(avail, int(share)) if share and share.isdigit() and int(share) > 0 else avail, avail


def cpu_baseline_oracle(w, seconds, threads, dpd_noise=False, seed=1234):
    """Times the oracle (CPU restatement of the reference step) with
    `threads` runners on workload w for about `seconds`."""
    import mechanica_amd as m
    from oracle import oracle as O
    o = O.from_workload(w, m, threads=threads, seed=seed)
    if dpd_noise:
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
    total_steps, total_t = steps + 1, dt + one
    return {"value": N * total_steps / total_t, "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "sample": "%s N=%d, %d timed steps after 1 warm-up step, %d threads (of %d cores in the affinity "
                      "mask), %.1f s" % (w["name"], N, total_steps, threads, len(os.sched_getaffinity(0)),
                                         total_t)}


def line(args, value, wall, dtype, config, roof, cpu, metric=METRIC, walls=None, **extra):
    out = {"metric": metric, "value": value, "unit": "particle-steps/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": 1e3 * wall / args.steps, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": dtype, "data": "synthetic", "config": config,
           "roofline": roof, "cpu_baseline": cpu}
    if walls:
        out["runs_ms_per_step"] = [1e3 * t / args.steps for t in walls]
    if cpu:
        out["speedup_vs_cpu"] = value / cpu["value"]
    out.update(extra)
    return out


# ------------------------------------------------------------------ workloads

def bench_single(args, threads, workload=None):
    import mechanica_amd as m
    from mechanica_amd import workloads as W
    workload = workload or args.workload
    if workload == "dpd":
        w = W.dpd_fluid(N=args.n, sigma=3.0, seed=2)
        w["dpd_noise"] = True
        cfg = "dpd_fluid N=%d density=3 alpha=25 gamma=4.5 sigma=3 (counter RNG) cutoff=1 dt=0.01 cells=%dx%dx%d"
        per, bonds = BYTES_NEWTONIAN, 0
    elif workload == "lj":
        w = W.argon_lj(N=args.n, seed=1)
        cfg = "argon_lj N=%d density=10 LJ 12-6 on [0.275,1] tol=1e-3 cutoff=1 dt=0.01 cells=%dx%dx%d"
        per, bonds = BYTES_NEWTONIAN, 0
    elif workload == "molecules":
        side = max(2, int(round((args.n // 4) ** (1.0 / 3.0))))
        w = W.molecules(n=side, seed=8)
        cfg = ("molecules N=%d: four-bead chains on a lattice of spacing 1.6, per molecule 3 harmonic bonds "
               "(k=100 r0=0.3), 2 harmonic angles (k=20 theta0=110deg), 1 harmonic dihedral (k=1 n=3 delta=0), "
               "5 exclusions; LJ 12-6 sigma=0.25 eps=0.5 on [0.1,1] cutoff=1, dt=0.002 cells=%dx%dx%d")
        per, bonds = BYTES_NEWTONIAN + BYTES_TERM_OFFSET, len(w["bonds"]["i"])
        per_terms = BYTES_PER_TERM_ENTRY * (3 * len(w["angles"]["i"]) + 4 * len(w["dihedrals"]["i"])
                                            + 2 * len(w["exclusions"]["i"]))
    elif workload == "channel":
        w = W.channel(N=args.n, seed=9)
        w["dpd_noise"] = True
        cfg = ("channel N=%d (examples/dpd1.py at scale): DPD alpha=1 gamma=1 sigma=1.5 (counter RNG) cutoff=1, "
               "mass 10, no-slip walls in z, constant force (0.01,0,0), dt=0.1 cells=%dx%dx%d")
        per, bonds = BYTES_NEWTONIAN, 0
    elif workload == "species":
        w = W.species_fluid(N=args.n, seed=10)
        w["dpd_noise"] = True
        cfg = ("species N=%d: DPD fluid at density 3 (sigma=3, counter RNG) in two types with three species, "
               "5 fluxes (Fick within and between types, secrete, uptake, decay), dt=0.01 cells=%dx%dx%d")
        per, bonds = BYTES_NEWTONIAN, 0
    else:  # lattice
        w = W.lattice(n=64, seed=3)
        cfg = ("lattice N=%d (64^3) Morse(d=1,a=6,r0=1) cutoff=2 + 774144 harmonic bonds k=100 r0=1 "
               "dissociation 5, overdamped dt=0.005 cells=%dx%dx%d")
        per, bonds = BYTES_OVERDAMPED, len(w["bonds"]["i"])
    if args.integrator == "rk4":
        w["integrator"] = 1
        cfg += " integrator=RK4 (four force stages per step)"
    u = W.build_universe(m, w, seed=1234)
    N = len(w["x"])
    log("workload %s N=%d built" % (w["name"], N))
    tune = None
    if args.brick:
        u.set_brick(*[int(v) for v in args.brick.split(",")])
    elif not args.no_autotune:
        u.step(args.warmup)  # tune on a settled state, not the initial one
        tune = u.autotune(steps_per=5)
    wall, walls = timed_steps(u, args.steps, args.warmup)
    value = N * args.steps / wall
    log("timed 5 x %d steps, median %.3f s" % (args.steps, wall))
    force_ms, step_ms = kernel_ms(u, args.steps)
    bytes_alg = per * N + BYTES_PER_BOND * bonds + (per_terms if workload == "molecules" else 0)
    # Algorithmic FP32 flops: unique in-cutoff pairs of the current
    # configuration (half the neighbour-count sum) and the advance.
    pairs = int(u.neighbour_counts().astype("int64").sum()) // 2
    dyn = "overdamped" if w["types"][0]["dynamics"] == 1 else "newtonian"
    pair_flops = FLOPS_DPD_PAIR if w["pots"][0][0] == "dpd" else FLOPS_TABLE_PAIR
    flops = pairs * pair_flops + N * FLOPS_ADVANCE[dyn] + bonds * (FLOPS_TABLE_PAIR + 3)
    roof = roofline(bytes_alg, force_ms, "k_brick (LDS brick force + advance) with its global-memory fallback",
                    MEASURED_TRAFFIC.get((workload, args.n)), flops=flops,
                    valu=MEASURED_VALU.get((workload, args.n)))
    roof["step_ms_events"] = step_ms
    roof["unique_pairs"] = pairs
    if (workload, args.n) in MEASURED_BRICK:
        # The counters were read at one layout; the autotuner may pick another.
        roof["counters_brick"] = MEASURED_BRICK[(workload, args.n)]
    cpu = None if args.no_cpu_baseline else cpu_baseline_oracle(w, args.cpu_seconds, threads,
                                                                dpd_noise=bool(w.get("dpd_noise")))
    config = {"workload": cfg % ((N,) + tuple(w["cells"])), "particles": N, "cells": w["cells"],
              "brick": u.brick(), "waves_per_simd": u.waves()}
    extra = {"autotune": tune} if tune else {}
    return line(args, value, wall, "f32", config, roof, cpu, walls=walls, **extra)


def bench_ensemble(args, threads):
    import mechanica_amd as m
    from mechanica_amd import workloads as W
    e = W.ensemble(R=args.replicas)
    sizes = [len(r["x"]) for r in e["replicas"]]
    N = sum(sizes)
    u = W.build_ensemble(m, e, seed=1234)
    log("ensemble of %d replicas, %d particles built" % (len(sizes), N))
    tune = None if args.no_autotune else u.autotune(steps_per=2)
    wall, walls = timed_steps(u, args.steps, args.warmup)
    value = N * args.steps / wall
    force_ms, step_ms = kernel_ms(u, args.steps)
    ebrick = u.brick()
    del u
    # The plain kernel: one cell-sorting universe of the same total size.
    w = W.cell_sort(N=N, seed=5)
    p = W.build_universe(m, w, seed=1234)
    ptune = None if args.no_autotune else p.autotune(steps_per=2)
    pwall, _ = timed_steps(p, args.steps, args.warmup)
    pforce, _ = kernel_ms(p, args.steps)
    plain = {"value": N * args.steps / pwall, "workload": "cell_sort N=%d cells=%dx%dx%d (one universe)" % (
        (N,) + tuple(w["cells"])), "force_kernel_ms": pforce, "brick": p.brick()}
    del p
    roof = roofline(BYTES_OVERDAMPED * N, force_ms, "k_brick over all replicas (one launch per step)")
    roof["step_ms_events"] = step_ms
    cpu = None
    if not args.no_cpu_baseline:
        # A bounded sample: the oracle on replicas taken in order until
        # about the requested time is spent, rate over the sample.
        import numpy as np
        from oracle import oracle as O
        done_n, done_t, used = 0, 0.0, 0
        for r in e["replicas"]:
            wr = dict(r)
            wr.update({k: e[k] for k in ("cutoff", "dt", "types", "pots", "random")})
            wr.update({"name": "cell_sort", "bonds": None, "v": np.zeros_like(r["x"])})
            o = O.from_workload(wr, m, threads=threads, seed=1234)
            t0 = time.perf_counter()
            o.step(5)
            done_t += time.perf_counter() - t0
            done_n += 5 * len(wr["x"])
            used += 1
            if done_t > args.cpu_seconds:
                break
        cpu = {"value": done_n / done_t, "unit": "particle-steps/s", "cores": threads, "kind": "port",
               "sample": "first %d replicas, 5 steps each, %d threads, %.1f s" % (used, threads, done_t)}
    config = {"workload": "ensemble of %d two-type cell-sorting replicas, N_i log-uniform in [500, 20000], "
                          "total %d, density 0.8, Morse d=3/1 a=6 r0=1 cutoff 2.5, overdamped Langevin kT=0.1, "
                          "dt=0.01, hipGraph-captured steps" % (len(sizes), N), "particles": N,
              "replicas": len(sizes)}
    config["brick"] = ebrick
    return line(args, value, wall, "f32", config, roof, cpu, walls=walls,
                metric="batched particle-steps/s vs plain kernel", plain_kernel=plain,
                batched_over_plain=value / plain["value"], autotune=tune and tune["brick"])


def bench_adjoint(args, threads):
    import numpy as np
    import torch
    import mechanica_amd as m
    import mechanica_amd.adjoint as A
    from mechanica_amd import workloads as W
    w = W.adjoint_morse(N=65536, L=40.0, seed=6)
    N, T, every = len(w["x"]), 50, 10
    kw = dict(mass=w["mass"], radius=w["radius"], smin=w["smin"], smax=w["smax"])
    s = A.MorseSystem(w["type"], w["dim"], w["cutoff"], w["dt"], **kw)
    target = torch.as_tensor(W.rdf_target(w)).cuda()
    x0 = torch.as_tensor(w["x"]).cuda()
    th0 = torch.as_tensor(w["theta"])

    def one():
        x = x0.clone().requires_grad_(True)
        th = th0.clone().requires_grad_(True)
        loss = A.rdf_loss(s, x, th, target, T, every)
        loss.backward()
        return loss

    reps = max(1, args.steps // T)
    for _ in range(max(1, args.warmup // T)):
        one()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        loss = one()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    value = N * T * reps / wall
    # Forward alone.
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        s.run(x0, th0, T, every)
    torch.cuda.synchronize()
    fwd = N * T * reps / (time.perf_counter() - t0)
    # The plain engine step on the same system (tabulated Morse).
    u = m.Universe(dim=w["dim"], cells=[16, 16, 16], cutoff=2.5, dt=0.01)
    for _ in range(2):
        u.add_type(1.0, 1.0, 1)
    keep = []
    for (a, b), row in zip(((0, 0), (0, 1), (1, 1)), w["theta"]):
        p = m.Potential.morse(d=row[0], a=row[1], r0=row[2], min=1e-4, max=2.5, tol=1e-3)
        u.bind(p, a, b)
        keep.append(p)
    u.load(w["x"], None, w["type"])
    if not args.no_autotune:
        u.autotune(steps_per=2)
    pwall, _ = timed_steps(u, args.steps, args.warmup)
    plain = {"value": N * args.steps / pwall, "workload": "same system through the engine step (tabulated "
             "Morse tol 1e-3, overdamped)", "brick": u.brick()}
    step_ms = 1e3 * wall / (reps * T)
    roof = roofline(BYTES_OVERDAMPED * N * 3, step_ms,
                    "adjoint step: forward + recomputed forward + k_backward per trajectory step")
    cpu = None
    if not args.no_cpu_baseline:
        wc = {"name": "adjoint_morse", "dim": w["dim"], "cells": [16, 16, 16], "cutoff": 2.5, "dt": 0.01,
              "types": [{"mass": 1.0, "radius": 1.0, "dynamics": 1}] * 2, "x": w["x"],
              "v": np.zeros_like(w["x"]), "type": w["type"], "bonds": None,
              "pots": [("morse", (row[0], row[1], row[2], 1e-4, 2.5, 1e-3), a, b)
                       for (a, b), row in zip(((0, 0), (0, 1), (1, 1)), w["theta"])]}
        cpu = cpu_baseline_oracle(wc, args.cpu_seconds, threads)
        cpu["sample"] += " (forward step only: the reference has no adjoint)"
    config = {"workload": "adjoint: N=65536 two-type Morse, box 40, uniform start, 50 overdamped steps dt=0.01, "
                          "loss = MSE of the soft 64-bin g(r) on [0,2.5] vs a seeded target, checkpoints every 10, "
                          "gradients w.r.t. x0 and (d,a,r0) per type pair", "particles": N, "steps_per_trajectory": T,
              "trajectories": reps}
    return line(args, value, wall, "f32 forward, f64 adjoint accumulation", config, roof, cpu,
                metric="adjoint particle-steps/s (forward+backward) vs plain kernel", plain_kernel=plain,
                forward_only=fwd, adjoint_over_plain=value / plain["value"], loss=float(loss))


def copy_calibration(args):
    """A plain device copy of known size for scaling FETCH_SIZE and
    WRITE_SIZE: 256 MiB read and 256 MiB written per launch."""
    import torch
    n = 64 * 1024 * 1024
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)  # the runtime's copy kernel
        torch.mul(a, 2.0, out=b)  # an elementwise kernel: reads a, writes b
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n, "launches": 10}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="dpd",
                    choices=["dpd", "lj", "lattice", "molecules", "channel", "species", "ensemble", "adjoint"])
    ap.add_argument("--n", type=int, default=1048576)
    ap.add_argument("--integrator", default="euler", choices=["euler", "rk4"],
                    help="forward Euler (the reference default) or Runge-Kutta 4")
    ap.add_argument("--replicas", type=int, default=256)
    ap.add_argument("--brick", default=None, help="bx,by,bz,cap: fixed force-kernel layout (implies --no-autotune)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default brick layout instead of timing the candidates before warm-up")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-lj", action="store_true", help="default workload only: skip the 1M LJ line")
    ap.add_argument("--copy-calibration", action="store_true")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        copy_calibration(args)
        return
    threads, _ = host_threads()
    if args.workload in ("dpd", "lj", "lattice", "molecules", "channel", "species"):
        out = bench_single(args, threads)
        if args.workload == "dpd" and not args.no_lj:
            # The metric's other half: 1M argon LJ with its own roofline and
            # CPU baseline.
            lj = bench_single(args, threads, workload="lj")
            out["lj"] = {k: lj[k] for k in ("value", "unit", "ms_per_step", "runs_ms_per_step", "config",
                                            "roofline", "cpu_baseline", "speedup_vs_cpu") if k in lj}
    elif args.workload == "ensemble":
        out = bench_ensemble(args, threads)
    else:
        out = bench_adjoint(args, threads)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
