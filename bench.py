#!/usr/bin/env python3
"""Benchmark: particle-turns/s of FP64 multi-particle ring tracking on one MI355X.

Workload (default ``dba``): the 20-cell DBA ring + RF (501 elements per turn), 1e6
particles (Gaussian, seeded), 1,000 turns -- one "step" is one full ringpass! of that
beam, started from the same device-resident initial distribution (a device-to-device copy
inside the step).  The timed region holds exactly K steps between two stream
synchronisations; inputs are resident in HBM before it starts.

The JSON line carries:
  value          scheduled particle-turns/s over the timed region (N * T * K / t)
  roofline       algorithmic FP64 flops of one kernel launch (flops per particle-turn x
                 live particle-turns of the launch) / the launch's average duration,
                 measured with HIP events on the library's stream, against the MI355X
                 FP64 peak of 78.6 TFLOP/s
  cpu_baseline   the oracle (the C++ restatement of the reference's CPU path, same loop
                 order) on this host's cores, on a bounded sample of the same workload

Usage: python bench.py [--gpus 1] [--steps 2] [--warmup 1] [--workload dba|fodo]
                       [--kernel fast|exact] [--turns-per-launch 100]
"""

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "trackpad.jl_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import numpy as np  # noqa: E402

FP64_PEAK = 78.6e12
HBM_PEAK = 8.0e12


def host_cores() -> int:
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


def workload(name, n_override, t_override):
    from trackpad_hip import Beam
    from trackpad_hip import workloads as W

    beam = Beam(W.ENERGY)
    if name == "dba":
        lat = W.dba_ring(beam=beam)
        n, T = 1_000_000, 1000
        mk = lambda n_: W.dba_particles(n_, seed=2)
        aperture = None
    elif name == "fodo":
        lat = W.fodo_lattice()
        n, T = 10_000, 1000
        mk = lambda n_: W.fodo_particles(n_, seed=1)
        aperture = 0.02
    else:
        raise SystemExit(f"unknown workload {name!r}")
    if n_override:
        n = n_override
    if t_override:
        T = t_override
    return beam, lat, n, T, mk, aperture


def cpu_baseline(lat, beam, mk, aperture, n_sample, t_sample, threads):
    from oracle import pyoracle as O

    c0 = mk(n_sample)
    t0 = time.perf_counter()
    _, lf, _, _ = O.ringpass(c0, lat, beam, np.zeros(n_sample, np.int64), t_sample, aperture=aperture,
                             nthreads=threads)
    dt = time.perf_counter() - t0
    return n_sample * t_sample / dt, dt


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="dba", choices=["dba", "fodo"])
    ap.add_argument("--kernel", default="fast", choices=["fast", "exact"])
    ap.add_argument("--particles", type=int, default=0, help="override N (default: the workload's)")
    ap.add_argument("--turns", type=int, default=0, help="override turns per step (default: the workload's)")
    ap.add_argument("--turns-per-launch", type=int, default=100)
    ap.add_argument("--block", type=int, default=256)
    ap.add_argument("--ppl", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample", default="16384x20", help="particles x turns of the CPU baseline sample")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU only (the tracking path shards by particles; "
              "multi-GPU runs are not part of it). Use --gpus 1.", file=sys.stderr)
        sys.exit(2)

    from trackpad_hip import HIPContext, HIPParticles, compile_table, track_
    from trackpad_hip.flops import flops_per_turn

    beam, lat, n, T, mk, aperture = workload(args.workload, args.particles, args.turns)
    fast = args.kernel == "fast"
    ctx = HIPContext(0)
    ctx.set_launch(args.block, args.ppl, args.turns_per_launch)
    tab = ctx.table(compile_table(lat, beam, aperture=aperture))
    c0 = mk(n)
    init = HIPParticles.from_coords(c0, ctx=ctx)
    work = HIPParticles(n, ctx)
    print(f"[bench] workload={args.workload} N={n} T={T} elements={len(lat)} kernel={args.kernel} "
          f"block={args.block} ppl={args.ppl} turns/launch={args.turns_per_launch}", file=sys.stderr, flush=True)

    def step():
        work.copy_from(init)
        track_(work, tab, nturns=T, fast=fast, timed=True)
        return ctx.timing()

    for i in range(args.warmup):
        step()
        print(f"[bench] warmup step {i + 1}/{args.warmup} done", file=sys.stderr, flush=True)
    ctx.sync()
    kernel_ms = 0.0
    launches = 0
    t0 = time.perf_counter()
    for i in range(args.steps):
        tm = step()
        kernel_ms += tm["kernel_ms"]
        launches += tm["launches"]
        print(f"[bench] step {i + 1}/{args.steps}: kernel {tm['kernel_ms']:.1f} ms in {tm['launches']} launches",
              file=sys.stderr, flush=True)
    ctx.sync()
    t1 = time.perf_counter()
    wall = t1 - t0
    _, lf, lt, le = work.download()

    fpt = flops_per_turn(lat)
    n_seq = len(lat)
    lost = lt >= 0
    live_turns = np.where(lost, lt + (le + 1) / n_seq, T).sum()  # particle-turns actually tracked
    sched = float(n) * T * args.steps
    value = sched / wall
    live = float(live_turns) * args.steps / wall
    avg_launch_s = kernel_ms / 1e3 / max(launches, 1)
    live_per_launch = float(live_turns) * args.steps / max(launches, 1)
    achieved = fpt * live_per_launch / avg_launch_s

    out = {
        "metric": "particle-turns/s FP64 (1 GPU, % FP64 peak); batched adjoint/forward cost ratio",
        "value": value,
        "unit": "particle-turns/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Gaussian beam)",
        "config": {"workload": f"{args.workload}_{n}x{T}", "particles": n, "turns": T, "elements_per_turn": n_seq,
                   "kernel": args.kernel, "block": args.block, "particles_per_lane": args.ppl,
                   "turns_per_launch": args.turns_per_launch, "flops_per_particle_turn": fpt},
        "live_particle_turns_per_s": live,
        "lost_particles": int(lost.sum()),
        "roofline": {"bound": "mfma", "pipe": "vector FP64 VALU (no MFMA: not a contraction); FP64 peak 78.6 TFLOP/s",
                     "achieved": achieved / 1e12, "peak": FP64_PEAK / 1e12, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK, "traffic": None,
                     "avg_launch_ms": avg_launch_s * 1e3, "launches": launches},
    }
    if not args.no_cpu_baseline:
        ns, ts = (int(v) for v in args.cpu_sample.lower().split("x"))
        cores = host_cores()
        v_all, dt_all = cpu_baseline(lat, beam, mk, aperture, ns, ts, cores)
        n1 = max(256, ns // 16)
        v_1, dt_1 = cpu_baseline(lat, beam, mk, aperture, n1, ts, 1)
        out["cpu_baseline"] = {"value": v_all, "unit": "particle-turns/s", "cores": cores, "kind": "port",
                               "sample": f"{ns} particles x {ts} turns of the same workload ({dt_all:.1f} s), "
                                         "oracle C++ restatement, OpenMP over particles"}
        out["cpu_baseline_1thread"] = {"value": v_1, "unit": "particle-turns/s", "cores": 1, "kind": "port",
                                       "sample": f"{n1} particles x {ts} turns ({dt_1:.1f} s)"}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
