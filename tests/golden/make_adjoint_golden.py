"""Writes tests/golden/adjoint_dba_T2000.json: exact directional derivatives of the DBA
ring's loss after 2,000 turns, for the batched adjoint test at the batched workload's
longest turn count (tests/test_gpu_fast_tolerance.py).

For two of the batched workload's ring variants (W.variant_rings(2, seed=10)), 8 particles
each, the loss S = sum over survivors of (x^2 + y^2) after T = 2,000 turns is evaluated by
the oracle in quad precision (__float128) at (k + eps dk, x0 + eps dx0) for eps = +-h and
+-h/2 with h = 2^-23; the directions dk (every quad k1 and sextupole k2) and dx0 are dyadic,
so every perturbed input is exact in double.  Central differences D(h), D(h/2) carry a
truncation error c h^2 that is not small at this turn count (4e-7 relative at h/2: the
phase advance over 2,000 turns makes S strongly nonlinear in k); the fixture stores the
Richardson extrapolation (4 D(h/2) - D(h)) / 3, whose remaining error is O(h^4), together
with D(h) and D(h/2) so that a reader sees the extrapolation step.

Runtime: about 5 minutes on 8 cores.  Run from the repository root:
    python tests/golden/make_adjoint_golden.py
"""

import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
for p in (ROOT, os.path.join(ROOT, "trackpad.jl_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

import mpmath as mp  # noqa: E402
import numpy as np  # noqa: E402

from oracle import pyoracle as O  # noqa: E402
from trackpad_hip import Beam, set_strengths, strength_positions  # noqa: E402
from trackpad_hip import workloads as W  # noqa: E402

T = 2000
N = 8
H = 2.0 ** -23
OUT = os.path.join(HERE, "adjoint_dba_T2000.json")


def directions(lat, n, seed):
    """Dyadic directions: strengths in multiples of 2^-10, coordinates in multiples of
    2^-10 times a per-plane power of two near the beam's spread."""
    rng = np.random.default_rng(seed)
    pos = list(strength_positions(lat))
    dk = np.round(rng.normal(0.0, 1.0, len(pos)) * 2 ** 10) / 2 ** 10
    unit = np.array([2.0 ** -14, 2.0 ** -16, 2.0 ** -17, 2.0 ** -19, 2.0 ** -13, 2.0 ** -10])
    dx0 = np.round(rng.normal(0.0, 1.0, (n, 6)) * 2 ** 10) / 2 ** 10 * unit
    return pos, dk, dx0


def main():
    mp.mp.dps = 40
    beam = Beam(W.ENERGY)
    rings, _, _ = W.variant_rings(nvar=2, seed=10)
    out = {"turns": T, "particles": N, "h": H, "variant_seed": 10, "variants": []}
    for v, lat in enumerate(rings):
        c0 = W.dba_particles(N, seed=80 + v)
        pos, dkp, dx0 = directions(lat, N, 500 + v)
        k0 = np.array([lat[p].k1 if hasattr(lat[p], "k1") else lat[p].k2 for p in pos])

        def s_quad(eps):
            ks, c = k0 + eps * dkp, c0 + eps * dx0
            assert np.all(ks - k0 == eps * dkp) and np.all(c - c0 == eps * dx0), "inexact perturbation"
            (hi, lo), na = O.loss_quad(c, set_strengths(lat, pos, list(ks)), beam, T, nthreads=8)
            assert na == N, "a particle was lost"
            return mp.mpf(hi) + mp.mpf(lo)

        t0 = time.time()
        d1 = (s_quad(H) - s_quad(-H)) / (2 * H)
        d2 = (s_quad(H / 2) - s_quad(-H / 2)) / H
        rich = (4 * d2 - d1) / 3
        s0 = s_quad(0.0)
        print(f"variant {v}: D(h) {mp.nstr(d1, 17)}  D(h/2) {mp.nstr(d2, 17)}  extrapolated {mp.nstr(rich, 17)} "
              f"({time.time() - t0:.0f} s)", flush=True)
        out["variants"].append({
            "particle_seed": 80 + v, "direction_seed": 500 + v, "positions": [int(p) for p in pos],
            "dk": dkp.tolist(), "dx0": dx0.tolist(), "loss_sum": mp.nstr(s0, 30),
            "dS_fd_h": mp.nstr(d1, 30), "dS_fd_h2": mp.nstr(d2, 30), "dS": mp.nstr(rich, 30)})
    with open(OUT, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
