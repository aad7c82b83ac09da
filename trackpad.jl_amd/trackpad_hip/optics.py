"""Linear optics on the GPU: one-turn maps, tunes, chromaticity and closed orbits.

Mirrors the reference's optics entry points (src/optics.jl) with the same names, arguments
and formulas, but every map derivative is the exact Jacobian of the tracked map from
``tp_jacobian`` (six reverse passes with unit seeds through the adjoint kernels) instead of
the reference's central finite differences with step ``h``; ``h`` is accepted for signature
parity and has no effect.

    one_turn_map(lat, beam; reference, h)               src/optics.jl:110-121
    fastfindm66 / findm66(lat, dp[, order]; E0, m0, orb) src/optics.jl:191-217
    fastfindm66_refpts / findm66_refpts                 src/optics.jl:224-260
    gettune(lat, beam; reference, h)                    src/optics.jl:262-283
    getchrom(lat, beam; dp, reference, h)               src/optics.jl:300-312
    find_closed_orbit_6d / _4d                          src/optics.jl:335-389

``jacobians`` is the batched form: the T-turn Jacobian of every particle of a beam.
"""

from __future__ import annotations

import ctypes as C
import math
from typing import Optional, Sequence, Tuple

import numpy as np

from . import _lib as L
from .lattice import M_ELECTRON, Beam, Lattice
from .table import compile_table
from .tracking import HIPContext, HIPParticles, default_context

__all__ = ["jacobians", "one_turn_map", "fastfindm66", "findm66", "fastfindm66_refpts", "findm66_refpts",
           "gettune", "getchrom", "find_closed_orbit_6d", "find_closed_orbit_4d", "tune_from_map"]

_PD = C.POINTER(C.c_double)


def jacobians(coords, lat: Lattice, beam: Beam, nturns: int = 1, *, turn: int = 0, ckpt_stride: int = 1,
              fast: bool = False, ctx: Optional[HIPContext] = None, aperture: Optional[float] = None,
              time: float = 0.0, dt_turn: float = 0.0):
    """(J, final, lost_flags): J[q] = d(coordinates after nturns) / d(initial coordinates) of
    particle q (N x 6 x 6; zero for lost particles), with the tracked beam and loss flags.
    Time-varying elements are resolved per turn, as in tracking."""
    ctx = ctx or default_context()
    tab = ctx.table(compile_table(lat, beam, aperture=aperture, turn=turn, nturns=nturns, time=time,
                                  dt_turn=dt_turn))
    c0 = np.asarray(coords, dtype=np.float64)
    if c0.ndim == 1:
        c0 = c0.reshape(1, 6)
    p = HIPParticles.from_coords(c0, ctx=ctx)
    jac = np.zeros((max(p.n, 1), 6, 6))
    L.check(ctx._lib.tp_jacobian(ctx.handle, tab.handle, p.handle, int(turn), int(nturns), int(ckpt_stride),
                                 L.TP_TRACK_FAST if fast else 0, jac.ctypes.data_as(_PD)))
    fin, lf, _, _ = p.download()
    return jac[: p.n], np.ascontiguousarray(fin), lf


def _map_at(lat: Lattice, beam: Beam, reference, ctx=None) -> Tuple[np.ndarray, np.ndarray]:
    J, fin, lf = jacobians(np.asarray(reference, dtype=np.float64).reshape(1, 6), lat, beam, 1, ctx=ctx)
    if lf[0]:
        raise ValueError("the reference particle is lost in one turn")
    return J[0], fin[0]


def one_turn_map(lat: Lattice, beam: Beam, *, reference=None, h: float = 1e-8,
                 ctx: Optional[HIPContext] = None) -> np.ndarray:
    """6 x 6 one-turn Jacobian about ``reference`` (src/optics.jl:110-121)."""
    del h
    ref = np.zeros(6) if reference is None else np.asarray(reference, dtype=np.float64)
    return _map_at(lat, beam, ref, ctx)[0]


def _reference6(dp: float, orb) -> np.ndarray:
    orb = np.asarray(orb, dtype=np.float64)
    if orb.shape != (6,):
        raise ValueError("orb must have length 6")
    dpeff = orb[5] if (dp == 0 and orb[5] != 0) else dp  # src/optics.jl:74-84
    return np.array([orb[0], orb[1], orb[2], orb[3], orb[4], dpeff])


def fastfindm66(lat, dp: float = 0.0, *, E0: float = 3.0e9, m0: float = M_ELECTRON, orb=None, h: float = 3e-8,
                ctx: Optional[HIPContext] = None) -> np.ndarray:
    """JuTrack-compatible 6 x 6 map (src/optics.jl:191-201)."""
    del h
    lat = lat if isinstance(lat, Lattice) else Lattice(list(lat))
    beam = Beam(E0, mass=m0)
    return one_turn_map(lat, beam, reference=_reference6(dp, np.zeros(6) if orb is None else orb), ctx=ctx)


def findm66(lat, dp: float, order: int = 0, **kw) -> np.ndarray:
    """JuTrack-compatible map API (src/optics.jl:208-217); any order gives the exact linear map."""
    del order
    return fastfindm66(lat, dp, **kw)


def fastfindm66_refpts(lat, dp: float, refpts: Sequence[int], *, E0: float = 3.0e9, m0: float = M_ELECTRON,
                       orb=None, h: float = 3e-8, ctx: Optional[HIPContext] = None) -> np.ndarray:
    """6 x 6 x len(refpts) maps of the segments ending at each 1-based reference point, each
    about the same reference orbit (src/optics.jl:224-244)."""
    del h
    lat = lat if isinstance(lat, Lattice) else Lattice(list(lat))
    n = len(lat)
    prev = 0
    for rp in refpts:
        if rp < 1 or rp > n:
            raise ValueError(f"refpts entries must be in [1, {n}]")
        if rp < prev:
            raise ValueError("refpts must be non-decreasing")
        prev = rp
    beam = Beam(E0, mass=m0)
    ref = _reference6(dp, np.zeros(6) if orb is None else orb)
    maps = np.zeros((6, 6, len(refpts)))
    prev = 0
    for i, rp in enumerate(refpts):
        seg = Lattice(list(lat.elements[prev:rp]))
        maps[:, :, i] = one_turn_map(seg, beam, reference=ref, ctx=ctx)
        prev = rp
    return maps


def findm66_refpts(lat, dp: float, order: int, refpts: Sequence[int], **kw) -> np.ndarray:
    del order
    return fastfindm66_refpts(lat, dp, refpts, **kw)


def tune_from_map(M: np.ndarray) -> Tuple[float, float]:
    """Fractional tunes from the uncoupled 2 x 2 blocks (src/optics.jl:262-271)."""
    M = np.asarray(M)
    cx = (M[0, 0] + M[1, 1]) / 2
    cy = (M[2, 2] + M[3, 3]) / 2
    sx = math.copysign(1.0, M[0, 1]) * math.sqrt(abs(-M[0, 1] * M[1, 0] - (M[0, 0] - M[1, 1]) ** 2 / 4))
    sy = math.copysign(1.0, M[2, 3]) * math.sqrt(abs(-M[2, 3] * M[3, 2] - (M[2, 2] - M[3, 3]) ** 2 / 4))
    if M[0, 1] == 0:
        sx = 0.0
    if M[2, 3] == 0:
        sy = 0.0
    qx = (math.atan2(sx, cx) / (2 * math.pi)) % 1.0
    qy = (math.atan2(sy, cy) / (2 * math.pi)) % 1.0
    return qx, qy


def gettune(lat: Lattice, beam: Beam, *, reference=None, h: float = 1e-8,
            ctx: Optional[HIPContext] = None) -> Tuple[float, float]:
    """(Qx, Qy) from the one-turn map about ``reference`` (src/optics.jl:278-283)."""
    ref = np.zeros(6) if reference is None else np.asarray(reference, dtype=np.float64)
    M = fastfindm66(lat, ref[5], E0=beam.energy, m0=beam.mass, orb=ref, h=h, ctx=ctx)
    return tune_from_map(M)


def getchrom(lat: Lattice, beam: Beam, *, dp: float = 0.0, reference=None, h: float = 1e-8,
             ctx: Optional[HIPContext] = None) -> Tuple[float, float]:
    """Chromaticity as the reference forms it: tune change over a momentum step of 1e-8
    (src/optics.jl:300-312), with exact one-turn maps."""
    ref = np.zeros(6) if reference is None else np.asarray(reference, dtype=np.float64)
    r0 = np.array([ref[0], ref[1], ref[2], ref[3], ref[4], dp])
    M0 = fastfindm66(lat, dp, E0=beam.energy, m0=beam.mass, orb=r0, h=h, ctx=ctx)
    qx0, qy0 = tune_from_map(M0)
    dpp = 1e-8
    r1 = np.array([ref[0], ref[1], ref[2], ref[3], ref[4], dp + dpp])
    M1 = fastfindm66(lat, dp + dpp, E0=beam.energy, m0=beam.mass, orb=r1, h=h, ctx=ctx)
    qx1, qy1 = tune_from_map(M1)
    return (qx1 - qx0) / dpp, (qy1 - qy0) / dpp


def find_closed_orbit_6d(lat: Lattice, beam: Beam, *, x0=None, tol: float = 1e-10, maxiter: int = 20,
                         h: float = 1e-6, reg: float = 1e-12, ctx: Optional[HIPContext] = None) -> np.ndarray:
    """6-D closed orbit by Newton iterations, x += (I - J + reg I)^-1 (M(x) - x)
    (src/optics.jl:335-358), J the exact one-turn Jacobian at x."""
    del h
    x = np.zeros(6) if x0 is None else np.array(x0, dtype=np.float64)
    eye = np.eye(6)
    for _ in range(maxiter):
        J, xo = _map_at(lat, beam, x, ctx)
        d = xo - x
        if np.linalg.norm(d) < tol:
            return x
        x = x + np.linalg.solve(eye - J + reg * eye, d)
    return x


def find_closed_orbit_4d(lat: Lattice, beam: Beam, *, dp: float = 0.0, x0=None, tol: float = 1e-10,
                         maxiter: int = 20, h: float = 1e-6, reg: float = 1e-12,
                         ctx: Optional[HIPContext] = None) -> np.ndarray:
    """4-D closed orbit (x, px, y, py) at fixed momentum offset ``dp``, z = 0 at entry
    (src/optics.jl:365-389)."""
    del h
    x = np.zeros(4) if x0 is None else np.array(x0, dtype=np.float64)
    eye = np.eye(4)
    for _ in range(maxiter):
        J, xo = _map_at(lat, beam, np.array([x[0], x[1], x[2], x[3], 0.0, dp]), ctx)
        d = xo[:4] - x
        if np.linalg.norm(d) < tol:
            return x
        x = x + np.linalg.solve(eye - J[:4, :4] + reg * eye, d)
    return x
