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
  streamed       after the timed region: the streamed-state variant (one turn per launch,
                 the beam read from and written to HBM every turn) -- its HBM GB/s against
                 the 8 TB/s peak and its FP64 rate (--streamed-turns, 0 skips it)
  parity_kernel  after the timed region: the parity build (no FMA contraction, IEEE
                 division and square root, the reference's operation order) on the same
                 beam, autotuned on its own, one launch of --parity-turns turns
  batched, batched_adjoint, vs_per_variant_launches
                 the metric's second half, measured after the DBA legs: the batched
                 workload below, one forward graph launch and one adjoint graph launch

Workload ``batched``: 256 variants of the DBA ring (strengths x (1 + N(0, 1%)), seeded),
N_v log-uniform in [1k, 64k], T_v uniform in [100, 2000]; one step is one graph launch of
the persistent batched kernel over all variants.  The line adds the batched adjoint (one
graph launch: checkpointed forward, reverse pass, per-variant reduction) and its cost
relative to the forward, and both against plain per-variant launches on the same GPU at
the variants' full turn counts: --compare-variants variants spread over the beam sizes,
timed both as one batch and as plain launches, and the 256-variant per-variant time
extrapolated from them (time per turn interpolated in log N).

Workload ``da``: the dynamic-aperture scan -- the 2,000-magnet ring (80 DBA cells + RF, every
magnet with order-10 normal and skew errors) on the 1024 x 1024 x 3 grid (3,145,728
particles); one step is --turns turns (default 20: the full 10,000-turn scan is ~45 minutes
of GPU time, and the rate is independent of the turn count).

Workload ``da_lossless``: the same ring on 1e6 particles of the DBA beam's distribution at
half its size (none lost), 5 turns per step: the kernel's rate on the order-10 lattice
without the scan's losses.

Workload ``diff``: the differentiation configuration -- dL/d(120 k1, 80 k2, x0) of
L = mean(x^2 + y^2) after 100 turns of the DBA ring over 4,096 particles; one step is one
tp_adjoint call (forward with checkpoints + reverse pass + reduction).  ``value`` is
gradient particle-turns/s over the device time of the two passes; the line adds their cost
relative to plain tracking of the same beam.

Usage: python bench.py [--gpus 1] [--steps 2] [--warmup 1] [--workload dba|fodo|da|da_lossless|diff|batched]
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

# HBM bytes per particle of one tracking-kernel launch, measured with rocprofv3 --pmc on the
# 1e6-particle DBA launch of 100 turns, per register budget of the kernel (waves/SIMD; 0 = the
# compiler's, 3 waves): FETCH_SIZE x 46,875 / 23,456.4 (the counter reads 23,456.4 KiB for a
# plain 48 MB device copy in the same run) + WRITE_SIZE x 1.00 (exact on the same copy), per
# 1e6 particles.  Budget 0: 25,588.4 + 47,459.3 KiB (the part above the 46,875 KiB of
# coordinates is the tile queue's atomics), independent of the turns per launch since the
# state stays in registers in between; budgets 5 and 6 add the part of their register spills
# (scratch) that leaves the L2 over those 100 turns: 5 waves ~0.9 B per particle-turn, a
# negligible share of the kernel's time at 8 TB/s; at 6 waves the spills reach the hot paths
# (6.3 GB per launch), which is why the autotuner does not pick that budget for this ring.
_FETCH_SCALE = 46875.0 / 23456.4
MEASURED_BYTES_PER_PARTICLE_LAUNCH = {("dba", 0): (25588.4 * _FETCH_SCALE + 47459.3) * 1024 / 1e6,
                                      ("dba", 5): (38009.6 * _FETCH_SCALE + 112941.9) * 1024 / 1e6,
                                      ("dba", 6): (49959.4 * _FETCH_SCALE + 6135626.5) * 1024 / 1e6}

# FP64 flops the fast DBA kernel executes per particle-turn, from rocprofv3 PMC on the same
# launch (2 SQ_INSTS_VALU_FMA_F64 + SQ_INSTS_VALU_MUL_F64 + SQ_INSTS_VALU_ADD_F64 per 64-particle
# tile-turn): fewer than the reference's 313,587 because the fast build does not repeat work
# (a pure-dipole bend body as one rotation, drifts as a two-FMA series, the two half-drifts
# where Yoshida steps meet as one drift, the edges' fringe term reused across the ring's
# identical bends).
MEASURED_EXECUTED_FLOPS_PER_PARTICLE_TURN = {("dba", "fast"): 142_178}


def visible_cores() -> int:
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def host_cores() -> int:
    """Threads of the CPU baseline: every core this process may use, or the share that
    OMP_NUM_THREADS grants it (set on shared hosts, where the visible cores belong to
    other jobs too)."""
    n = visible_cores()
    omp = os.environ.get("OMP_NUM_THREADS", "")
    if omp.isdigit() and int(omp) > 0:
        n = min(n, int(omp))
    return max(1, n)


def cores_note(cores: int) -> str:
    vis = visible_cores()
    if vis == cores:
        return f"{cores} threads = all visible host cores"
    return f"{cores} threads = this process's CPU share (OMP_NUM_THREADS); the host shows {vis} cores"


# The FP64 work of every kernel here runs on the vector ALUs (VALU): the maps are chains of
# small nonlinear updates per particle, not contractions, so there is no MFMA work.
BOUND = "vector_fp64"
PIPE = "vector FP64 VALU (no MFMA: not a contraction); FP64 peak 78.6 TFLOP/s"


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
    elif name == "da":
        lat = W.da_ring(beam=beam)
        n, T = 3 * 1024 * 1024, 20

        def mk(n_):
            g = W.da_grid()
            # a sample of the grid spread over all three momenta and the whole (x, y) square
            return g if n_ >= len(g) else g[np.linspace(0, len(g) - 1, n_).astype(np.int64)]

        aperture = None
    elif name == "da_lossless":
        # the dynamic-aperture ring on a beam it keeps (the DBA distribution at half size):
        # the kernel's own rate on the order-10 lattice, without the scan's losses
        lat = W.da_ring(beam=beam)
        n, T = 1_000_000, 5
        mk = lambda n_: W.dba_particles(n_, seed=3) * 0.5
        aperture = None
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


def live_turns_(lt, le, n_seq, T):
    """Particle-turns actually tracked: a particle lost at position e of turn t ran t turns
    and e + 1 elements of the next."""
    return float(np.where(lt >= 0, lt + (le + 1) / n_seq, T).sum())


def bench_diff(args, fast):
    from trackpad_hip import Beam, HIPContext, HIPParticles, compile_table, ringpass_grad, track_
    from trackpad_hip import workloads as W
    from trackpad_hip.flops import flops_per_turn

    beam = Beam(W.ENERGY)
    lat = W.dba_ring(beam=beam)
    n = args.particles or 4096
    T = args.turns or 100
    c0 = W.dba_particles(n, seed=4)
    ctx = HIPContext(0)
    ctx.set_waves(args.fwd_waves, args.back_waves)
    tab = ctx.table(compile_table(lat, beam))
    init = HIPParticles.from_coords(c0, ctx=ctx)
    work = HIPParticles(n, ctx)
    print(f"[bench] diff: DBA ring, N={n} T={T} ckpt_stride={args.ckpt_stride} kernel={args.kernel}",
          file=sys.stderr, flush=True)

    def step():
        g = ringpass_grad(c0, lat, beam, T, ckpt_stride=args.ckpt_stride, fast=fast, timed=True, ctx=ctx)
        work.copy_from(init)
        track_(work, tab, nturns=T, fast=fast, timed=True)
        return g, ctx.timing()["kernel_ms"]

    for i in range(args.warmup):
        step()
        print(f"[bench] warmup step {i + 1}/{args.warmup} done", file=sys.stderr, flush=True)
    fwd = bwd = plain = 0.0
    t0 = time.perf_counter()
    for i in range(args.steps):
        g, pl = step()
        fwd += g.forward_ms
        bwd += g.backward_ms
        plain += pl
        print(f"[bench] step {i + 1}/{args.steps}: forward {g.forward_ms:.1f} ms, reverse {g.backward_ms:.1f} ms, "
              f"plain tracking {pl:.1f} ms", file=sys.stderr, flush=True)
    wall = time.perf_counter() - t0
    K = args.steps
    grad_s = (fwd + bwd) / 1e3
    fpt = flops_per_turn(lat)
    out = {
        "metric": "particle-turns/s FP64 (1 GPU, % FP64 peak); batched adjoint/forward cost ratio",
        "value": float(n) * T * K / grad_s,
        "unit": "gradient particle-turns/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": (fwd + bwd) / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Gaussian beam)",
        "config": {"workload": f"diff_dba_{n}x{T}", "particles": n, "turns": T, "elements_per_turn": len(lat),
                   "strengths": int((g.dk != 0).sum()), "kernel": args.kernel, "ckpt_stride": args.ckpt_stride,
                   "flops_per_particle_turn": fpt},
        "adjoint": {"forward_ms": fwd / K, "reverse_ms": bwd / K, "plain_tracking_ms": plain / K,
                    "gradient_over_forward": (fwd + bwd) / plain, "wall_ms_per_step": wall * 1e3 / K,
                    "loss": g.loss, "survivors": g.n_alive},
        "roofline": {"bound": BOUND, "pipe": "vector FP64 VALU; 4,096 particles are 64 waves on 1,024 SIMDs: "
                     "occupancy-bound, not at the roofline",
                     "achieved": fpt * n * T / (plain / K / 1e3) / 1e12, "peak": FP64_PEAK / 1e12, "unit": "TFLOP/s",
                     "frac": fpt * n * T / (plain / K / 1e3) / FP64_PEAK, "traffic": None,
                     "avg_launch_ms": plain / K, "launches": K},
    }
    print(json.dumps(out), flush=True)


class BatchedSweep:
    """The batched workload: 256 DBA variants (strengths x (1 + N(0, 1 %)), seeded), N_v
    log-uniform in [1k, 64k], T_v uniform in [100, 2000], bound to one tp_batch."""

    def __init__(self, args, fast):
        from trackpad_hip import Beam, HIPBatch, HIPContext, HIPParticles, compile_table
        from trackpad_hip import workloads as W
        from trackpad_hip.flops import flops_per_turn

        self.args, self.fast = args, fast
        self.beam = Beam(W.ENERGY)
        self.nvar = args.variants
        self.rings, self.ns, ts = W.variant_rings(nvar=self.nvar, seed=11)
        if args.turn_scale != 1.0:
            ts = np.maximum(1, np.round(ts * args.turn_scale)).astype(np.int64)
        self.ts = ts
        self.ctx = ctx = HIPContext(0)
        ctx.set_waves(args.fwd_waves, args.back_waves)
        self.tabs = [ctx.table(compile_table(r, self.beam)) for r in self.rings]
        c0 = [W.dba_particles(int(n), seed=1000 + v) for v, n in enumerate(self.ns)]
        self.parts = [HIPParticles.from_coords(c, ctx=ctx) for c in c0]
        self.init = [HIPParticles.from_coords(c, ctx=ctx) for c in c0]
        self.batch = HIPBatch(self.tabs, self.parts, ctx=ctx)
        self.block, self.fwd_waves, self.back_waves = args.block, args.fwd_waves, args.back_waves
        self.batch.set_launch(self.block, self.fwd_waves, self.back_waves)
        self.K = args.batched_ckpt_stride
        self.tune = None
        if not args.no_autotune:
            probe = np.maximum(1, self.ts // 100).astype(np.int32)
            self.tune = self.batch.autotune(probe, ckpt_stride=self.K, adjoint=True, fast=fast, cache=args.tune_cache)
            self.block, self.fwd_waves, self.back_waves = (self.tune["block"], self.tune["fwd_waves"],
                                                           self.tune["back_waves"])
            print(f"[bench] batched autotune: {self.tune}", file=sys.stderr, flush=True)
        self.n_seq = len(self.rings[0])
        self.fpt = flops_per_turn(self.rings[0])
        self.sched = float(np.sum(self.ns * self.ts))
        print(f"[bench] batched: {self.nvar} variants, sum N = {int(self.ns.sum())}, T in [{self.ts.min()}, "
              f"{self.ts.max()}], sum N*T = {self.sched:.3e}, kernel={args.kernel}", file=sys.stderr, flush=True)
        # load the kernels (a 1-turn pass of every variant), then restore the beams
        self.batch.track(np.ones(self.nvar, np.int32), fast=fast)
        self.reset()

    def reset(self):
        for p, q in zip(self.parts, self.init):
            p.copy_from(q)
        self.ctx.sync()

    def forward(self, steps, warmup):
        """Timed graph launches of the batched forward; returns (wall s, kernel ms) summed
        over the timed steps."""
        b, ts = self.batch, self.ts.astype(np.int32)
        for i in range(warmup):
            self.reset()
            b.track(ts, fast=self.fast)
            print(f"[bench] batched warmup step {i + 1}/{warmup} done", file=sys.stderr, flush=True)
        wall = kernel_ms = 0.0
        for i in range(steps):
            self.reset()
            t0 = time.perf_counter()
            b.track(ts, fast=self.fast, timed=True)
            wall += time.perf_counter() - t0
            kernel_ms += self.ctx.timing()["kernel_ms"]
            print(f"[bench] batched forward step {i + 1}/{steps}: {wall * 1e3 / (i + 1):.0f} ms", file=sys.stderr,
                  flush=True)
        live = 0.0
        for p, t in zip(self.parts, self.ts):
            _, _, lt, le = p.download()
            live += live_turns_(lt, le, self.n_seq, int(t))
        self.live = live
        return wall, kernel_ms

    def adjoint(self):
        """One graph launch of the batched adjoint; returns (gradients, graph ms, wall s)."""
        self.reset()
        t0 = time.perf_counter()
        g = self.batch.adjoint(self.ts.astype(np.int32), ckpt_stride=self.K, fast=self.fast, timed=True)
        wall = time.perf_counter() - t0
        ms = self.ctx.timing()["kernel_ms"]
        self.info_a = self.batch.info()
        print(f"[bench] batched adjoint: {ms:.0f} ms (graph), {wall * 1e3:.0f} ms wall, checkpoints "
              f"{self.info_a['ckpt_bytes'] / 2**30:.1f} GiB", file=sys.stderr, flush=True)
        return g, ms, wall

    def compare(self, nsub, fwd_graph_ms, adj_graph_ms):
        """Plain per-variant launches (tp_track / tp_adjoint one variant after the other on the
        library's stream) against the batch, at the variants' own turn counts, on nsub variants
        spread evenly over the sorted beam sizes: the same nsub variants as one batch and as
        nsub plain launches, both measured; and the 256-variant per-variant time extrapolated
        from the subset's time per turn as a function of N (interpolated in log N), against the
        measured 256-variant graphs."""
        import ctypes as C
        import trackpad_hip._lib as L
        from trackpad_hip import HIPBatch

        ctx, fast, K = self.ctx, self.fast, self.K
        order = np.argsort(self.ns, kind="stable")
        sub = order[np.unique(np.round(np.linspace(0, self.nvar - 1, min(nsub, self.nvar))).astype(np.int64))]
        sb = HIPBatch([self.tabs[v] for v in sub], [self.parts[v] for v in sub], ctx=ctx)
        sb.set_launch(self.block, self.fwd_waves, self.back_waves)
        tsub = self.ts[sub].astype(np.int32)
        flags = L.TP_TRACK_FAST if fast else 0
        self.reset()
        sb.track(tsub, fast=fast)  # plan + capture outside the timed region
        self.reset()
        sb.track(tsub, fast=fast, timed=True)
        bf = ctx.timing()["kernel_ms"]
        self.reset()
        per_f = []
        for v in sub:
            t0 = time.perf_counter()
            L.check(ctx._lib.tp_track(ctx.handle, self.tabs[v].handle, self.parts[v].handle, 0, int(self.ts[v]), flags))
            per_f.append((time.perf_counter() - t0) * 1e3)
        print(f"[bench] compare forward ({len(sub)} variants, full turns): batched {bf:.0f} ms, per-variant "
              f"{sum(per_f):.0f} ms", file=sys.stderr, flush=True)
        self.reset()
        sb.adjoint(tsub, ckpt_stride=K, fast=fast)  # plan + capture
        self.reset()
        sb.adjoint(tsub, ckpt_stride=K, fast=fast, timed=True)
        ba = ctx.timing()["kernel_ms"]
        self.reset()
        loss, na = C.c_double(), C.c_int64()
        dk = np.zeros(self.n_seq)
        dx = np.zeros(6 * int(self.ns.max()))
        pd = C.POINTER(C.c_double)
        per_a = []
        for v in sub:
            t0 = time.perf_counter()
            L.check(ctx._lib.tp_adjoint(ctx.handle, self.tabs[v].handle, self.parts[v].handle, 0, int(self.ts[v]), K,
                                        flags, C.byref(loss), C.byref(na), dk.ctypes.data_as(pd),
                                        dx.ctypes.data_as(pd)))
            per_a.append((time.perf_counter() - t0) * 1e3)
        print(f"[bench] compare adjoint ({len(sub)} variants, full turns): batched {ba:.0f} ms, per-variant "
              f"{sum(per_a):.0f} ms", file=sys.stderr, flush=True)
        sb.close()
        self.reset()
        # extrapolation to every variant: time per turn of a plain launch as a function of N
        lx = np.log(self.ns[sub].astype(float))
        est_f = float(np.sum(self.ts * np.interp(np.log(self.ns.astype(float)), lx,
                                                 np.array(per_f) / self.ts[sub])))
        est_a = float(np.sum(self.ts * np.interp(np.log(self.ns.astype(float)), lx,
                                                 np.array(per_a) / self.ts[sub])))
        return {"variants_measured": int(len(sub)), "turns": "each variant's own (full) turn count",
                "particles_measured": int(self.ns[sub].sum()),
                "forward_batched_ms": bf, "forward_per_variant_ms": float(sum(per_f)),
                "forward_speedup": float(sum(per_f)) / bf,
                "adjoint_batched_ms": ba, "adjoint_per_variant_ms": float(sum(per_a)),
                "adjoint_speedup": float(sum(per_a)) / ba,
                "all_variants_per_variant_forward_ms_est": est_f,
                "all_variants_forward_speedup_est": est_f / fwd_graph_ms,
                "all_variants_per_variant_adjoint_ms_est": est_a,
                "all_variants_adjoint_speedup_est": est_a / adj_graph_ms,
                "note": "subset: the same variants as one batch and as plain launches, both timed; all_variants: "
                        "per-variant time per turn interpolated in log N from the subset, times each variant's "
                        "turns, against the measured 256-variant graphs"}


def batched_results(args, fast, steps, warmup):
    """Batched forward (steps graph launches), batched adjoint (one graph launch) and the
    comparison with per-variant launches; returns (sweep, dict of results, forward wall s)."""
    sw = BatchedSweep(args, fast)
    wall, kernel_ms = sw.forward(steps, warmup)
    avg_s = kernel_ms / 1e3 / steps
    achieved = sw.fpt * sw.live / avg_s
    g, adj_ms, adj_wall = sw.adjoint()
    res = {
        "forward": {"particle_turns_per_s": sw.sched * steps / wall, "live_particle_turns_per_s": sw.live / avg_s,
                    "graph_ms": avg_s * 1e3, "fp64_TFLOPs": achieved / 1e12, "fp64_frac": achieved / FP64_PEAK,
                    "work_items": sw.batch.info()["n_items"], "grid_blocks": sw.batch.info()["grid_fwd"]},
        "batched_adjoint": {"graph_ms": adj_ms, "wall_ms": adj_wall * 1e3, "forward_graph_ms": avg_s * 1e3,
                            "adjoint_over_forward": adj_ms / (avg_s * 1e3), "ckpt_stride": sw.K,
                            "ckpt_bytes": sw.info_a["ckpt_bytes"], "grid_back_blocks": sw.info_a["grid_back"],
                            "mean_loss": float(np.mean(g.loss)), "survivors": int(g.n_alive.sum())},
        "config": {"variants": sw.nvar, "particles_total": int(sw.ns.sum()), "turns_min": int(sw.ts.min()),
                   "turns_max": int(sw.ts.max()), "particle_turns": sw.sched, "elements_per_turn": sw.n_seq,
                   "block": sw.block, "fwd_waves": sw.fwd_waves, "back_waves": sw.back_waves,
                   "autotuned": sw.tune is not None, "turn_scale": args.turn_scale},
    }
    if args.compare_variants > 0:
        res["vs_per_variant_launches"] = sw.compare(args.compare_variants, avg_s * 1e3, adj_ms)
    return sw, res, wall, achieved, avg_s


def bench_batched(args, fast):
    from trackpad_hip import workloads as W

    sw, res, wall, achieved, avg_s = batched_results(args, fast, args.steps, args.warmup)
    out = {
        "metric": "particle-turns/s FP64 (1 GPU, % FP64 peak); batched adjoint/forward cost ratio",
        "value": sw.sched * args.steps / wall,
        "unit": "particle-turns/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Gaussian beams, seeded strength perturbations)",
        "config": dict({"workload": f"batched_dba_{sw.nvar}_variants", "kernel": args.kernel,
                        "flops_per_particle_turn": sw.fpt}, **res["config"]),
        "live_particle_turns_per_s": res["forward"]["live_particle_turns_per_s"],
        "roofline": {"bound": BOUND, "pipe": PIPE, "achieved": achieved / 1e12, "peak": FP64_PEAK / 1e12,
                     "unit": "TFLOP/s", "frac": achieved / FP64_PEAK, "traffic": None, "avg_launch_ms": avg_s * 1e3,
                     "launches": args.steps},
        "batched_adjoint": res["batched_adjoint"],
    }
    if "vs_per_variant_launches" in res:
        out["vs_per_variant_launches"] = res["vs_per_variant_launches"]
    if not args.no_cpu_baseline:
        ns_, ts_ = (int(v) for v in args.cpu_sample.lower().split("x"))
        cores = host_cores()
        mk = lambda n_: W.dba_particles(n_, seed=1000)
        v_all, dt_all = cpu_baseline(sw.rings[0], sw.beam, mk, None, ns_, ts_, cores)
        out["cpu_baseline"] = {"value": v_all, "unit": "particle-turns/s", "cores": cores, "kind": "port",
                               "sample": f"variant 0, {ns_} particles x {ts_} turns ({dt_all:.1f} s), oracle C++ "
                                         f"restatement, OpenMP over particles, {cores_note(cores)}"}
    print(json.dumps(out), flush=True)


def parity_leg(ctx, tab, init, work, args, fpt, n):
    """The parity build (no FMA contraction, IEEE division and square root: the reference's
    operations in its order) on the same beam after the timed region: autotuned on its own,
    one launch of --parity-turns turns, timed with HIP events."""
    from trackpad_hip import track_

    T = args.parity_turns
    tune = ctx.autotune(tab, init, probe_turns=2, fast=False, cache=args.tune_cache)
    ctx.set_launch(tune["block"], tune["ppl"], 0)
    work.copy_from(init)
    track_(work, tab, nturns=1, fast=False)  # untimed: first launch at this shape
    work.copy_from(init)
    ctx.sync()
    track_(work, tab, nturns=T, fast=False, timed=True)
    tm = ctx.timing()
    s = tm["kernel_ms"] / 1e3
    rate = float(n) * T / s
    print(f"[bench] parity build: {rate:.3e} particle-turns/s ({T} turns, {tm['kernel_ms']:.0f} ms)", file=sys.stderr,
          flush=True)
    return {"particle_turns_per_s": rate, "fp64_TFLOPs": fpt * rate / 1e12, "frac": fpt * rate / FP64_PEAK,
            "avg_launch_ms": tm["kernel_ms"] / max(tm["launches"], 1), "launches": tm["launches"], "turns": T,
            "block": tune["block"], "particles_per_lane": tune["ppl"], "fwd_waves": tune["fwd_waves"],
            "note": "parity build: -ffp-contract=off, IEEE division and square root, the reference's operation "
                    "order (bit-exact with the oracle where the maps use only + - * / sqrt)"}


def streamed(ctx, tab, init, work, args, fast, fpt, n, n_seq):
    """The streamed-state variant, measured after the timed region: one turn per launch, so
    every launch reads the beam from HBM and writes it back (48 B in, 48 B out and 8 B of
    loss record per particle: 104 B algorithmic).  Reports the launches' HBM rate against
    the 8 TB/s peak next to their FP64 rate."""
    from trackpad_hip import track_

    T = args.streamed_turns
    ctx.set_launch(args.block, args.ppl, 1)
    work.copy_from(init)
    track_(work, tab, nturns=1, fast=fast)  # untimed: first launch at this shape
    work.copy_from(init)
    ctx.sync()
    track_(work, tab, nturns=T, fast=fast, timed=True)
    tm = ctx.timing()
    ctx.set_launch(args.block, args.ppl, args.turns_per_launch)
    _, _, lt, le = work.download()
    live = float(live_turns_(lt, le, n_seq, T))
    launches = max(tm["launches"], 1)
    avg_s = tm["kernel_ms"] / 1e3 / launches
    alive_per_launch = live / launches
    algo = 104.0 * alive_per_launch
    bpp = MEASURED_BYTES_PER_PARTICLE_LAUNCH.get((args.workload, args.fwd_waves))
    flops = fpt * alive_per_launch / avg_s
    return {"turns_per_launch": 1, "turns": T, "launches": tm["launches"], "avg_launch_ms": avg_s * 1e3,
            "bytes_per_launch": algo, "achieved_GBps": algo / avg_s / 1e9, "peak_GBps": HBM_PEAK / 1e9,
            "hbm_frac": algo / avg_s / HBM_PEAK,
            "traffic": None if bpp is None else bpp * n,
            "fp64_TFLOPs": flops / 1e12, "fp64_frac": flops / FP64_PEAK,
            "note": "104 B per live particle per launch (48 in + 48 out + 8 loss record); rates over the "
                    "launches' HIP-event time; compute-bound even at one turn per launch"}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None, help="timed steps (default 2; batched 1)")
    ap.add_argument("--warmup", type=int, default=None, help="untimed steps (default 1; batched 0)")
    ap.add_argument("--workload", default="dba", choices=["dba", "fodo", "da", "da_lossless", "diff", "batched"])
    ap.add_argument("--variants", type=int, default=256, help="batched: number of lattice variants")
    ap.add_argument("--turn-scale", type=float, default=1.0, help="batched: scale every T_v (1 = the workload)")
    ap.add_argument("--compare-variants", type=int, default=16,
                    help="batched: variants (spread over the beam sizes) timed as plain per-variant launches at "
                         "their full turn counts (0 = skip the comparison)")
    ap.add_argument("--ckpt-stride", type=int, default=1, help="diff: turns between checkpoints")
    ap.add_argument("--batched-ckpt-stride", type=int, default=10, help="batched adjoint: turns between checkpoints")
    ap.add_argument("--no-batched", action="store_true",
                    help="dba: skip the batched legs (256-variant forward, adjoint, per-variant comparison)")
    ap.add_argument("--parity-turns", type=int, default=50,
                    help="dba: turns of the parity-build measurement after the timed region (0 = skip)")
    ap.add_argument("--kernel", default="fast", choices=["fast", "exact"])
    ap.add_argument("--particles", type=int, default=0, help="override N (default: the workload's)")
    ap.add_argument("--turns", type=int, default=0, help="override turns per step (default: the workload's)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="use --block/--ppl/--turns-per-launch/--*-waves as given instead of the autotuner's choice")
    ap.add_argument("--tune-cache", default=None, help="autotuner cache file (default: none, always time)")
    ap.add_argument("--turns-per-launch", type=int, default=100)
    ap.add_argument("--block", type=int, default=256)
    ap.add_argument("--ppl", type=int, default=1)
    ap.add_argument("--fwd-waves", type=int, default=0, help="register budget of the tracking kernels (0, 5, 6)")
    ap.add_argument("--back-waves", type=int, default=0, help="register budget of the reverse kernels (0, 3, 4)")
    ap.add_argument("--streamed-turns", type=int, default=10,
                    help="turns of the streamed-state measurement (1 turn per launch; 0 = skip)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample", default=None,
                    help="particles x turns of the CPU baseline sample (default 65536x20, about 13 s on 16 cores; da 2048x5)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU only (the tracking path shards by particles; "
              "multi-GPU runs are not part of it). Use --gpus 1.", file=sys.stderr)
        sys.exit(2)
    batched = args.workload == "batched"
    if args.steps is None:
        args.steps = 1 if batched else 2
    if args.warmup is None:
        args.warmup = 0 if batched else 1
    if args.cpu_sample is None:
        args.cpu_sample = "2048x5" if args.workload == "da" else "65536x20"
    if batched:
        bench_batched(args, args.kernel == "fast")
        return
    if args.workload == "diff":
        bench_diff(args, args.kernel == "fast")
        return

    from trackpad_hip import HIPContext, HIPParticles, compile_table, track_
    from trackpad_hip.flops import flops_per_turn

    beam, lat, n, T, mk, aperture = workload(args.workload, args.particles, args.turns)
    fast = args.kernel == "fast"
    ctx = HIPContext(0)
    ctx.set_launch(args.block, args.ppl, args.turns_per_launch)
    ctx.set_waves(args.fwd_waves, args.back_waves)
    tab = ctx.table(compile_table(lat, beam, aperture=aperture))
    c0 = mk(n)
    init = HIPParticles.from_coords(c0, ctx=ctx)
    work = HIPParticles(n, ctx)
    tune = None
    if not args.no_autotune:
        # launch shape from the autotuner: every candidate timed on a copy of the beam
        tune = ctx.autotune(tab, init, probe_turns=2, fast=fast, cache=args.tune_cache)
        args.block, args.ppl, args.turns_per_launch = tune["block"], tune["ppl"], tune["turns_per_launch"]
        args.fwd_waves = tune["fwd_waves"]
        print(f"[bench] autotune: {tune}", file=sys.stderr, flush=True)
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
    live_turns = live_turns_(lt, le, n_seq, T)
    sched = float(n) * T * args.steps
    value = sched / wall
    live = float(live_turns) * args.steps / wall
    avg_launch_s = kernel_ms / 1e3 / max(launches, 1)
    live_per_launch = float(live_turns) * args.steps / max(launches, 1)
    achieved = fpt * live_per_launch / avg_launch_s
    bpp = MEASURED_BYTES_PER_PARTICLE_LAUNCH.get((args.workload, args.fwd_waves))
    traffic = None if bpp is None else bpp * n  # HBM bytes per launch (rocprof PMC, calibrated)

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
                   "turns_per_launch": args.turns_per_launch, "fwd_waves": args.fwd_waves,
                   "autotuned": tune is not None, "flops_per_particle_turn": fpt},
        "live_particle_turns_per_s": live,
        "lost_particles": int(lost.sum()),
        "roofline": {"bound": BOUND, "pipe": PIPE,
                     "achieved": achieved / 1e12, "peak": FP64_PEAK / 1e12, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK, "traffic": traffic,
                     "avg_launch_ms": avg_launch_s * 1e3, "launches": launches,
                     "flops": "algorithmic: the reference's FP64 operations per particle-turn, counted as written"},
    }
    xf = MEASURED_EXECUTED_FLOPS_PER_PARTICLE_TURN.get((args.workload, args.kernel))
    if xf is not None:
        ex = xf * live_per_launch / avg_launch_s
        out["roofline"]["executed"] = {
            "flops_per_particle_turn": xf, "TFLOPs": ex / 1e12, "frac": ex / FP64_PEAK,
            "note": "FP64 flops the kernel executes (rocprofv3 PMC per tile-turn, a constant of this kernel) "
                    "at this run's rate: fewer than the reference's because the fast build does not repeat work"}
    if args.streamed_turns > 0:
        out["streamed"] = streamed(ctx, tab, init, work, args, fast, fpt, n, n_seq)
    if not args.no_cpu_baseline:
        ns, ts = (int(v) for v in args.cpu_sample.lower().split("x"))
        cores = host_cores()
        v_all, dt_all = cpu_baseline(lat, beam, mk, aperture, ns, ts, cores)
        n1 = max(256, ns // 16)
        v_1, dt_1 = cpu_baseline(lat, beam, mk, aperture, n1, ts, 1)
        out["cpu_baseline"] = {"value": v_all, "unit": "particle-turns/s", "cores": cores, "kind": "port",
                               "sample": f"{ns} particles x {ts} turns of the same workload ({dt_all:.1f} s), "
                                         f"oracle C++ restatement, OpenMP over particles, {cores_note(cores)}"}
        out["cpu_baseline_1thread"] = {"value": v_1, "unit": "particle-turns/s", "cores": 1, "kind": "port",
                                       "sample": f"{n1} particles x {ts} turns ({dt_1:.1f} s)"}
    if args.workload == "dba" and args.parity_turns > 0:
        out["parity_kernel"] = parity_leg(ctx, tab, init, work, args, fpt, n)
    if args.workload == "dba" and not args.no_batched:
        # the metric's second half: the batched sweep, forward and adjoint, one graph launch each
        del init, work
        _, res, _, _, _ = batched_results(args, fast, 1, 0)
        out["batched"] = dict(res["forward"], config=res["config"])
        out["batched_adjoint"] = res["batched_adjoint"]
        if "vs_per_variant_launches" in res:
            out["vs_per_variant_launches"] = res["vs_per_variant_launches"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
