"""Lattices and seeded particle distributions of the benchmark workloads.

The workloads are fixed by their parameters here (lattice layout, strengths, seeds), so
that the oracle, the tests and the benchmark all track the same inputs:

* ``fodo``     -- 8 x [QF(0.5, +1.2, n=4), D(2.0), QD(0.5, -1.2, n=4), D(2.0)] per turn,
                  10,000 particles, 1,000 turns, R = 0.02 m aperture, E = 3 GeV.
* ``dba``      -- 20 DBA cells of 25 elements (2 exact sector bends L = 1.5 m, angle 2pi/40,
                  with linear edge/fringe maps; 6 thick quads |k1| in [0.5, 2.4]; 4 thick
                  sextupoles, k2 = -32 (SD) and +27 (SF); 13 drifts) plus one RF cavity per turn
                  (V = 1 MV, h = 400, E = 3 GeV), 1e6 particles, 1,000 turns.
* ``da``       -- 80 DBA cells (bend angle 2pi/160), every magnet with random normal and
                  skew multipole errors up to order 10, a 1024 x 1024 (x, y) grid over
                  +-15 mm times delta in {-2 %, 0, +2 %}, 10,000 turns.
* ``variants`` -- 256 DBA rings with strengths perturbed by N(0, 1 %), uneven particle and
                  turn counts, for the batched path.

Decisions the configurations leave open (recorded so they are reproducible):
the RF frequency is set from the harmonic number as f = h beta c / C (the reference map
uses ``freq`` and ignores ``h``, src/tracking.jl:510), the cavity has L = 0, lag = 0,
philag = 0 and ``energy = 3e9`` (the reference default 0 would disable the kick,
src/elements/rf.jl:18); the beam energy is passed to ``Beam`` as kinetic energy as the
reference does (src/lattice.jl:51); the bend fringe uses fint = 0.5, gap = 0.03 m with the
Brown edge method; px, py are matched to the periodic beta functions at the ring start
(a symmetry point, alpha = 0).
"""

from __future__ import annotations

import math
from typing import List, Optional, Tuple

import numpy as np

from .elements import Drift, ExactSBendLinearEdge, Quadrupole, RFCavity, Sextupole
from .lattice import Beam, Lattice, total_length

__all__ = ["fodo_lattice", "fodo_particles", "dba_cell", "dba_ring", "dba_particles", "da_ring",
           "da_grid", "gaussian_particles", "DBA_K", "ENERGY", "variant_rings"]

ENERGY = 3.0e9
DBA_K = {"kf1": 2.1, "kd1": -2.4, "kf2": 0.5, "ksd": -32.0, "ksf": 27.0}


def fodo_lattice(ncell: int = 8, steps: int = 4) -> Lattice:
    els = []
    for _ in range(ncell):
        els += [Quadrupole(0.5, 1.2, num_int_steps=steps, name="QF"), Drift(2.0, name="D"),
                Quadrupole(0.5, -1.2, num_int_steps=steps, name="QD"), Drift(2.0, name="D")]
    return Lattice(els, name="FODO")


def gaussian_particles(n: int, seed: int, sx: float, spx: float, sy: float, spy: float,
                       sd: float) -> np.ndarray:
    """N x 6 coordinates: x, px, y, py ~ N(0, s), z = 0, delta ~ N(0, sd)."""
    rng = np.random.default_rng(seed)
    c = np.zeros((n, 6))
    c[:, 0] = rng.normal(0.0, sx, n)
    c[:, 1] = rng.normal(0.0, spx, n)
    c[:, 2] = rng.normal(0.0, sy, n)
    c[:, 3] = rng.normal(0.0, spy, n)
    c[:, 5] = rng.normal(0.0, sd, n)
    return c


def fodo_particles(n: int = 10_000, seed: int = 1) -> np.ndarray:
    return gaussian_particles(n, seed, 1e-4, 1e-5, 1e-4, 1e-5, 1e-3)


def dba_cell(ncell: int = 20, k: Optional[dict] = None, scale: Optional[np.ndarray] = None,
             steps: int = 10) -> List:
    """One symmetric DBA cell of 25 elements; ``scale`` (6 + 4 entries) multiplies each
    quad k1 and sextupole k2 (the batched variants)."""
    k = dict(DBA_K if k is None else k)
    th = 2.0 * math.pi / (2 * ncell)
    s = np.ones(10) if scale is None else np.asarray(scale, dtype=float)

    def B(nm):
        return ExactSBendLinearEdge(1.5, th, 0.0, 0.0, fint1=0.5, fint2=0.5, gap=0.03,
                                    num_int_steps=steps, name=nm)

    D = Drift
    q = [k["kf1"] * s[0], k["kd1"] * s[1], k["kf2"] * s[2], k["kf2"] * s[3], k["kd1"] * s[4], k["kf1"] * s[5]]
    x = [k["ksd"] * s[6], k["ksf"] * s[7], k["ksf"] * s[8], k["ksd"] * s[9]]
    return [D(2.0, name="DL"), Quadrupole(0.3, q[0], num_int_steps=steps, name="QF1"), D(0.25, name="D1"),
            Quadrupole(0.3, q[1], num_int_steps=steps, name="QD1"), D(0.4, name="D2"), B("B1"), D(0.3, name="D3"),
            Sextupole(0.2, x[0], num_int_steps=steps, name="SD"), D(0.2, name="D4"),
            Quadrupole(0.4, q[2], num_int_steps=steps, name="QF2"), D(0.2, name="D5"),
            Sextupole(0.2, x[1], num_int_steps=steps, name="SF"), D(0.1, name="DM"),
            Sextupole(0.2, x[2], num_int_steps=steps, name="SF"), D(0.2, name="D5"),
            Quadrupole(0.4, q[3], num_int_steps=steps, name="QF2"), D(0.2, name="D4"),
            Sextupole(0.2, x[3], num_int_steps=steps, name="SD"), D(0.3, name="D3"), B("B2"),
            D(0.4, name="D2"), Quadrupole(0.3, q[4], num_int_steps=steps, name="QD1"), D(0.25, name="D1"),
            Quadrupole(0.3, q[5], num_int_steps=steps, name="QF1"), D(2.0, name="DL")]


def _rf(lat_len: float, beam: Beam, h: float = 400.0, volt: float = 1.0e6) -> RFCavity:
    freq = h * beam.beta * 2.99792458e8 / lat_len
    return RFCavity(0.0, volt, freq, 0.0, h=h, philag=0.0, energy=ENERGY, name="RF")


def dba_ring(ncell: int = 20, beam: Optional[Beam] = None, scales: Optional[np.ndarray] = None,
             steps: int = 10) -> Lattice:
    """20 DBA cells + one RF cavity (501 elements)."""
    beam = beam or Beam(ENERGY)
    els = []
    for c in range(ncell):
        els += dba_cell(ncell, scale=None if scales is None else scales[c], steps=steps)
    L = total_length(Lattice(els))
    return Lattice(els + [_rf(L, beam)], name="DBA")


def dba_particles(n: int = 1_000_000, seed: int = 2, beta_xy: Tuple[float, float] = (None, None)) -> np.ndarray:
    """Gaussian DBA beam: sigma_x = 50 um, sigma_y = 10 um, sigma_delta = 1e-3; px, py
    matched as sigma / beta with the periodic beta functions at the ring start."""
    bx, by = beta_xy
    if bx is None:
        bx, by = DBA_BETA
    return gaussian_particles(n, seed, 50e-6, 50e-6 / bx, 10e-6, 10e-6 / by, 1e-3)


# Periodic beta functions at the start of the 20-cell DBA ring (symmetry point), from the
# oracle's one-turn Jacobian at the origin (tests/test_workloads.py recomputes them).
DBA_BETA = (2.922440729213387, 6.123935623898734)


def da_ring(ncell: int = 80, seed: int = 7, err: float = 1e-4, r_ref: float = 0.01,
            beam: Optional[Beam] = None, steps: int = 10) -> Lattice:
    """2,001-element DA ring: 80 DBA cells (bend angle 2pi/160) + RF, every magnet with
    normal and skew errors b_n, a_n = N(0, err) x (main field at r_ref) / r_ref^n for
    n = 0..10.  The main-order normal error is folded into the magnet's k (the reference
    overwrites that polynomial slot with k, src/tracking.jl:395, 446)."""
    beam = beam or Beam(ENERGY)
    rng = np.random.default_rng(seed)
    els = []
    for _ in range(ncell):
        for e in dba_cell(ncell, steps=steps):
            if isinstance(e, Quadrupole):
                main, order = e.k1 * r_ref, 1
            elif isinstance(e, Sextupole):
                main, order = e.k2 / 2 * r_ref ** 2, 2
            elif isinstance(e, ExactSBendLinearEdge):
                main, order = e.angle / e.L, 0
            else:
                els.append(e)
                continue
            n = np.arange(11)
            bn = rng.normal(0.0, err, 11) * abs(main) / r_ref ** n
            an = rng.normal(0.0, err, 11) * abs(main) / r_ref ** n
            if isinstance(e, Quadrupole):
                e.k1 = e.k1 + bn[1]
                bn[1] = 0.0
            elif isinstance(e, Sextupole):
                e.k2 = e.k2 + 2 * bn[2]
                bn[2] = 0.0
            e.polynom_b = bn
            e.polynom_a = an
            e.max_order = 10
            els.append(e)
    L = total_length(Lattice(els))
    return Lattice(els + [_rf(L, beam)], name="DA")


def da_grid(nx: int = 1024, ny: int = 1024, half: float = 0.015,
            deltas=(-0.02, 0.0, 0.02)) -> np.ndarray:
    """(len(deltas) * nx * ny) x 6 grid, delta-major then y then x; px = py = z = 0."""
    xs = np.linspace(-half, half, nx)
    ys = np.linspace(-half, half, ny)
    out = np.zeros((len(deltas) * nx * ny, 6))
    X, Y = np.meshgrid(xs, ys)  # Y rows, X columns
    k = 0
    for d in deltas:
        out[k:k + nx * ny, 0] = X.reshape(-1)
        out[k:k + nx * ny, 2] = Y.reshape(-1)
        out[k:k + nx * ny, 5] = d
        k += nx * ny
    return out


def variant_rings(nvar: int = 256, seed: int = 11, rel: float = 0.01, ncell: int = 20,
                  nmin: int = 1000, nmax: int = 64000, tmin: int = 100, tmax: int = 2000):
    """(rings, n_particles, n_turns) of the batched sweep: strengths x (1 + N(0, rel)),
    N log-uniform in [nmin, nmax], T uniform in [tmin, tmax]."""
    rng = np.random.default_rng(seed)
    rings = []
    for _ in range(nvar):
        sc = 1.0 + rng.normal(0.0, rel, (ncell, 10))
        rings.append(dba_ring(ncell, scales=sc))
    ns = np.exp(rng.uniform(math.log(nmin), math.log(nmax), nvar)).astype(np.int64)
    ts = rng.integers(tmin, tmax + 1, nvar).astype(np.int64)
    return rings, ns, ts
