"""Lattice -> device element table compiler (the writer for include/trackpad_table.h).

``compile_table(lat, beam)`` replaces the reference's per-call recomputation of element
constants inside ``pass!`` (e.g. SL, L1..K2 at src/tracking.jl:324-328, the kick-angle
corrections at :398-414, the edge tangents at :170 and the exact-body cos/sin at :875-876)
by one pass on the host.  Every constant is formed with the same IEEE operations in the
same order as the reference (Python floats are IEEE doubles and never contract), so the
device sees bit-identical values.  Identical elements share one record.
"""

from __future__ import annotations

import math
import struct
from typing import Dict, List, Optional, Tuple

import numpy as np

from .elements import (
    MAXORD, AbstractElement, AccelCavity, C_LIGHT, Corrector, CrabCavity, Drift, DriftSC, ExactSBend,
    ExactSBendLinearEdge, InvLorentzBoost, LBend, LongitudinalRFMap, LongitudinalRLCWake, LongitudinalWake,
    LorentzBoost, Marker, Octupole, OctupoleSC, Quadrupole, QuadrupoleSC, RFCavity, SBend, SBendSC, Sextupole,
    SextupoleSC, Solenoid, SpaceCharge, StrongGaussianBeam, StrongThinGaussianBeam, ThinMultipole,
    TimeVaryingElement, Translation, Wiggler, YRotation,
)
from .lattice import Beam, Lattice, beti

__all__ = [
    "compile_table", "TableInfo", "TP_TABLE_MAGIC", "TP_TABLE_VERSION", "TAGS",
    "DRIFT1", "DRIFT2", "KICK1", "KICK2",
]

TP_TABLE_MAGIC = 0x54504854
TP_TABLE_VERSION = 1

TAGS = {"MARKER": 0, "DRIFT": 1, "MPOLE": 2, "THINMP": 3, "SBEND": 4, "ESBEND": 5,
        "ESBEND_LE": 6, "RFCAV": 7, "SOLENOID": 8, "CORRECTOR": 9, "TRANSLATION": 10, "YROTATION": 11,
        "CRAB": 12, "ACCEL": 13, "LONGRF": 14, "BOOST": 15, "INVBOOST": 16, "GAUSSBB": 17, "RLCWAKE": 18,
        "TABWAKE": 19, "LBEND": 20, "NAN": 21, "WIGGLER": 22}
EPS = 2.220446049250313e-16  # eps(Float64)

F_T1, F_T2, F_R1, F_R2 = 0x1, 0x2, 0x4, 0x8
F_FIN_SHIFT, F_FOUT_SHIFT = 4, 8
F_DRIFTLIMIT, F_RF_DRIFT, F_RF_KICK = 0x1000, 0x2000, 0x4000
H_EXACT_HAMILTONIAN = 0x1

# Yoshida-4 coefficients: the reference's decimal literals (src/tracking.jl:29-32).
DRIFT1 = 0.6756035959798286638
DRIFT2 = -0.1756035959798286639
KICK1 = 1.351207191959657328
KICK2 = -1.702414383919314656
TWO_PI = 2.0 * math.pi  # == Julia Float64(2pi)


class TableInfo:
    """What the writer produced: positions, records and the parameter index of each
    lattice position (used by the adjoint to name gradients)."""

    def __init__(self):
        self.n_seq = 0
        self.n_rec = 0
        self.pool_len = 0
        self.seq: List[int] = []
        self.gorder: List[int] = []     # main polynomial index per position (-1 none)
        self.gscale: List[float] = []   # d b_gorder / d k per position (1, 1/2, 1/6)


def _poly_arrays(e) -> Tuple[np.ndarray, np.ndarray]:
    a = np.zeros(MAXORD + 1)
    b = np.zeros(MAXORD + 1)
    a[: len(e.polynom_a)] = e.polynom_a
    b[: len(e.polynom_b)] = e.polynom_b
    return a, b


def _check_order(e):
    m = int(e.max_order)
    if m < 0 or m > MAXORD:
        raise ValueError(f"{type(e).__name__} {e.name!r}: max_order {m} outside [0, {MAXORD}]")
    return m


def _interleave(a: np.ndarray, b: np.ndarray, m: int) -> List[float]:
    out: List[float] = []
    for i in range(m + 1):
        out.append(float(b[i]))
        out.append(float(a[i]))
    return out


def _kick_corrected(e, a: np.ndarray, b: np.ndarray):
    """Kick-angle correction of Quadrupole / SBend / ExactSBend (src/tracking.jl:398-414)."""
    if e.L > 0.0:
        b[0] = b[0] - math.sin(float(e.kick_angle[0])) / e.L
        a[0] = a[0] + math.sin(float(e.kick_angle[1])) / e.L


def _yoshida(L: float, n: int) -> Tuple[float, float, float, float]:
    """SL, L1, L2, K1, K2 of symplectic4_pass (src/tracking.jl:324-328)."""
    SL = L / n if n != 0 else (math.inf if L > 0 else (-math.inf if L < 0 else math.nan))
    return SL * DRIFT1, SL * DRIFT2, SL * KICK1, SL * KICK2


def _edge_block(irho: float, e: float, fint: float, gap: float, method: int) -> List[float]:
    """[fx, fy_const, e, fringecorr] of edge_fringe_entrance/exit (src/tracking.jl:160-221)."""
    if fint == 0.0 or gap == 0.0 or method == 0:
        fc = 0.0
    else:
        s = math.sin(e)
        c = math.cos(e)
        fc = irho * gap * fint * (1.0 + s * s) / c
    t = math.tan(e)
    fx = irho * t
    return [fx, irho * t, e, fc]


def _fringe_method(v: int, what: str) -> int:
    v = int(v)
    if v < 0 or v > 15:
        raise ValueError(f"{what} = {v} is outside the supported range 0..15")
    return v


def _misalign(e) -> Tuple[int, Optional[List[float]]]:
    flags = 0
    if not hasattr(e, "t1"):
        return 0, None
    if np.any(e.t1 != 0.0):
        flags |= F_T1
    if np.any(e.t2 != 0.0):
        flags |= F_T2
    if np.any(e.r1 != 0.0):
        flags |= F_R1
    if np.any(e.r2 != 0.0):
        flags |= F_R2
    if not flags:
        return 0, None
    blk = list(map(float, e.t1)) + list(map(float, e.t2))
    blk += list(map(float, np.asarray(e.r1).reshape(-1, order="F")))
    blk += list(map(float, np.asarray(e.r2).reshape(-1, order="F")))
    return flags, blk


def _record(e: AbstractElement, bt: float):
    """(tag, nsteps, max_order, flags, gorder, gscale, block, misalign_block) for one element."""
    if isinstance(e, TimeVaryingElement):
        raise ValueError(
            f"element {e.name!r} is time-varying: compile_table resolves it per turn when given "
            "the turn range (turn=, nturns=, time=, dt_turn=), or call materialize_lattice first "
            "(src/time_dependence.jl:137-139)")
    if isinstance(e, (Marker, SpaceCharge)):  # SpaceCharge: identity (src/tracking.jl:1176-1180)
        return TAGS["MARKER"], 0, 0, 0, -1, 0.0, [], None
    if isinstance(e, (Drift, DriftSC)):
        mflags, mblk = _misalign(e)
        return TAGS["DRIFT"], 0, 0, mflags, -1, 0.0, [e.L, e.L * bt], mblk
    if isinstance(e, (Quadrupole, Sextupole, Octupole)):
        m = _check_order(e)
        a, b = _poly_arrays(e)
        if isinstance(e, Quadrupole):
            b[1] = e.k1                  # src/tracking.jl:395
            _kick_corrected(e, a, b)     # src/tracking.jl:398-414
            gorder, gscale = 1, 1.0
        elif isinstance(e, Sextupole):
            b[2] = e.k2 / 2              # src/tracking.jl:446
            gorder, gscale = 2, 0.5
        else:
            b[3] = e.k3 / 6              # src/tracking.jl:478
            gorder, gscale = 3, 1.0 / 6.0
        n = int(e.num_int_steps)
        if n < 0:
            raise ValueError(f"{e.name!r}: num_int_steps {n} < 0")
        L1, L2, K1, K2 = _yoshida(e.L, n)
        blk = [L1, L2, K1, K2, L1 * bt, L2 * bt] + _interleave(a, b, m)
        if gorder > m:
            gorder = -1  # the main strength is beyond max_order: the reference ignores it
        mflags, mblk = _misalign(e)
        return TAGS["MPOLE"], n, m, mflags, gorder, gscale, blk, mblk
    if isinstance(e, ThinMultipole):
        m = _check_order(e)
        a, b = _poly_arrays(e)
        mflags, mblk = _misalign(e)
        return TAGS["THINMP"], 0, m, mflags, -1, 0.0, _interleave(a, b, m), mblk
    if isinstance(e, SBend):
        m = _check_order(e)
        a, b = _poly_arrays(e)
        _kick_corrected(e, a, b)         # src/tracking.jl:670-683
        irho = e.angle / e.L             # src/tracking.jl:665
        n = int(e.num_int_steps)
        L1, L2, K1, K2 = _yoshida(e.L, n)
        fin = _fringe_method(e.fringe_bend_entrance, "fringe_bend_entrance")
        fout = _fringe_method(e.fringe_bend_exit, "fringe_bend_exit")
        blk = [L1, L2, K1, K2, L1 * bt, L2 * bt, irho, K1 * irho, K2 * irho]
        blk += _edge_block(irho, e.e1, e.fint1, e.gap, fin)
        blk += _edge_block(irho, e.e2, e.fint2, e.gap, fout)
        blk += _interleave(a, b, m)
        mflags, mblk = _misalign(e)
        flags = mflags | (fin << F_FIN_SHIFT) | (fout << F_FOUT_SHIFT)
        return TAGS["SBEND"], n, m, flags, -1, 0.0, blk, mblk
    if isinstance(e, (ExactSBend, ExactSBendLinearEdge)):
        m = _check_order(e)
        a, b = _poly_arrays(e)
        n = int(e.num_int_steps)
        if n < 0:
            raise ValueError(f"{e.name!r}: num_int_steps {n} < 0")
        irho = e.angle / e.L             # src/tracking.jl:908
        if n == 0:
            body = [irho, e.L, 0.0, 0.0, 0.0, e.L * bt, 0.0,
                    math.cos(irho * e.L), math.sin(irho * e.L), 0.0, 0.0]
        else:
            _kick_corrected(e, a, b)     # src/tracking.jl:935-942
            L1, L2, K1, K2 = _yoshida(e.L, n)
            body = [irho, L1, L2, K1, K2, L1 * bt, L2 * bt,
                    math.cos(irho * L1), math.sin(irho * L1), math.cos(irho * L2), math.sin(irho * L2)]
        fin = _fringe_method(e.fringe_bend_entrance, "fringe_bend_entrance")
        fout = _fringe_method(e.fringe_bend_exit, "fringe_bend_exit")
        flags = (fin << F_FIN_SHIFT) | (fout << F_FOUT_SHIFT)
        if abs(irho) < 1e-6:             # src/tracking.jl:823, 862
            flags |= F_DRIFTLIMIT
        if isinstance(e, ExactSBend):
            t1, t2 = -e.e1, -e.e2        # bend_edge(r, irho, -e) (src/tracking.jl:919, 956)
            edges = [math.cos(e.e1), math.sin(e.e1),
                     math.cos(t1), math.sin(t1), math.sin(2 * t1), t1,
                     math.cos(t2), math.sin(t2), math.sin(2 * t2), t2,
                     math.cos(e.e2), math.sin(e.e2), e.gk]
            tag = TAGS["ESBEND"]
        else:
            edges = _edge_block(irho, e.e1, e.fint1, e.gap, fin) + _edge_block(irho, e.e2, e.fint2, e.gap, fout)
            edges += [0.0] * 5
            tag = TAGS["ESBEND_LE"]
        blk = body + edges + _interleave(a, b, m)
        assert len(body) + len(edges) == 24
        mflags, mblk = _misalign(e)
        return tag, n, m, flags | mflags, -1, 0.0, blk, mblk
    extra = _record_extra(e, bt)
    if extra is not None:
        return extra
    if isinstance(e, RFCavity):
        beta = 1.0 / bt                  # src/tracking.jl:504
        flags = (F_RF_DRIFT if e.L > 0.0 else 0) | (F_RF_KICK if e.energy > 0.0 else 0)
        nv = e.volt / e.energy if e.energy > 0.0 else 0.0
        blk = [e.L / 2, (e.L / 2) * bt, nv, TWO_PI * e.freq, e.lag, e.philag, beta * beta]
        return TAGS["RFCAV"], 0, 0, flags, -1, 0.0, blk, None
    raise NotImplementedError(
        f"element type {type(e).__name__} ({getattr(e, 'name', '?')!r}) has no device map")


def _sc_poly(e) -> Tuple[np.ndarray, np.ndarray, int]:
    """polynom (k0, k1, k2/2, k3/6) of the SC multipoles (src/tracking.jl:993-995) and
    max_order clamped to [0, 3] (:1008)."""
    a = np.zeros(MAXORD + 1)
    b = np.zeros(MAXORD + 1)
    a[:4] = e.polynom_a
    b[:4] = [e.k0, e.k1, e.k2 / 2, e.k3 / 6]
    return a, b, min(max(int(e.max_order), 0), 3)


def _rf_k(rf) -> float:
    """src/tracking.jl:1567-1570"""
    if isinstance(rf, RFCavity):
        return TWO_PI * rf.freq / C_LIGHT
    if isinstance(rf, (AccelCavity, CrabCavity)):
        return rf.k
    return 0.0


def _rf_h(rf) -> float:
    """src/tracking.jl:1562-1565"""
    if isinstance(rf, (RFCavity, AccelCavity)):
        return rf.h
    return 1.0


def _record_extra(e, bt: float):
    """Records of the remaining element families (src/tracking.jl:529-605, 981-1218,
    1531-1663), or None."""
    if isinstance(e, (QuadrupoleSC, SextupoleSC, OctupoleSC)):
        a, b, m = _sc_poly(e)
        if isinstance(e, QuadrupoleSC):
            _kick_corrected(e, a, b)     # src/tracking.jl:1004-1007
        n = int(e.num_int_steps)
        if n < 0:
            raise ValueError(f"{e.name!r}: num_int_steps {n} < 0")
        L1, L2, K1, K2 = _yoshida(e.L, n)
        mflags, mblk = _misalign(e)
        return TAGS["MPOLE"], n, m, mflags, -1, 0.0, [L1, L2, K1, K2, L1 * bt, L2 * bt] + _interleave(a, b, m), mblk
    if isinstance(e, SBendSC):
        mflags, mblk = _misalign(e)
        if e.L == 0.0:                   # src/tracking.jl:1047-1049
            return TAGS["NAN"], 0, 0, 0, -1, 0.0, [], None
        a = np.zeros(MAXORD + 1)
        b = np.zeros(MAXORD + 1)
        a[:4], b[:4] = e.polynom_a, e.polynom_b
        m = min(max(int(e.max_order), 0), 3)
        _kick_corrected(e, a, b)
        irho = e.angle / e.L
        n = int(e.num_int_steps)
        L1, L2, K1, K2 = _yoshida(e.L, n)
        fin = _fringe_method(e.fringe_bend_entrance, "fringe_bend_entrance")
        fout = _fringe_method(e.fringe_bend_exit, "fringe_bend_exit")
        blk = [L1, L2, K1, K2, L1 * bt, L2 * bt, irho, K1 * irho, K2 * irho]
        blk += _edge_block(irho, e.e1, e.fint1, e.gap, fin) + _edge_block(irho, e.e2, e.fint2, e.gap, fout)
        blk += _interleave(a, b, m)
        return TAGS["SBEND"], n, m, mflags | (fin << F_FIN_SHIFT) | (fout << F_FOUT_SHIFT), -1, 0.0, blk, mblk
    if isinstance(e, Solenoid):
        mflags, mblk = _misalign(e)
        if e.ks == 0.0:                  # src/tracking.jl:539-540: a drift
            return TAGS["DRIFT"], 0, 0, mflags, -1, 0.0, [e.L, e.L * bt], mblk
        return TAGS["SOLENOID"], 0, 0, mflags, -1, 0.0, [e.L, e.ks], mblk
    if isinstance(e, Corrector):
        mflags, mblk = _misalign(e)
        xk, yk = e.xkick, e.ykick        # src/tracking.jl:588-596
        blk = [e.L, xk, yk, (xk * xk) / 3 + (yk * yk) / 3, xk / 2, yk / 2]
        return TAGS["CORRECTOR"], 0, 0, mflags, -1, 0.0, blk, mblk
    if isinstance(e, Translation):
        return TAGS["TRANSLATION"], 0, 0, 0, -1, 0.0, [e.dx, e.dy, e.ds], None
    if isinstance(e, YRotation):
        ang = -e.angle                   # src/tracking.jl:1195
        if ang == 0.0:
            return TAGS["MARKER"], 0, 0, 0, -1, 0.0, [], None
        return TAGS["YROTATION"], 0, 0, 0, -1, 0.0, [math.cos(ang), math.sin(ang), math.tan(ang)], None
    if isinstance(e, CrabCavity):
        volt = e.volt * (1.0 + float(e.errors[0]))   # src/tracking.jl:1532-1534
        phi = e.phi + float(e.errors[1])
        E = max(abs(e.energy), EPS)
        blk = [e.L / 2, (e.L / 2) * bt, volt / E, e.k, phi, e.k * volt / E * bt]
        return TAGS["CRAB"], 0, 0, (F_RF_DRIFT if e.L > 0.0 else 0), -1, 0.0, blk, None
    if isinstance(e, AccelCavity):
        beta = 1.0 / bt                  # src/tracking.jl:1550-1558
        beta2 = beta * beta
        if beta2 <= EPS:
            return TAGS["MARKER"], 0, 0, 0, -1, 0.0, [], None
        E = max(abs(e.energy), EPS)
        return TAGS["ACCEL"], 0, 0, 0, -1, 0.0, [e.volt / (beta2 * E), e.k, e.phis, math.sin(e.phis)], None
    if isinstance(e, LongitudinalRFMap):
        k = _rf_k(e.rf)                  # src/tracking.jl:1575-1585
        if abs(k) <= EPS:
            return TAGS["MARKER"], 0, 0, 0, -1, 0.0, [], None
        beta = 1.0 / bt
        eta = e.alphac - (1.0 - beta * beta)
        return TAGS["LONGRF"], 0, 0, 0, -1, 0.0, [TWO_PI * _rf_h(e.rf) * eta / k], None
    if isinstance(e, LorentzBoost):
        if e.mode != 0:
            return TAGS["MARKER"], 0, 0, 0, -1, 0.0, [], None
        return TAGS["BOOST"], 0, 0, 0, -1, 0.0, [e.tanang, 1.0 / e.cosang], None
    if isinstance(e, InvLorentzBoost):
        if e.mode != 0:
            return TAGS["MARKER"], 0, 0, 0, -1, 0.0, [], None
        return TAGS["INVBOOST"], 0, 0, 0, -1, 0.0, [e.sinang, e.cosang], None
    if isinstance(e, StrongThinGaussianBeam):
        sx2 = max(e.rmssizex * e.rmssizex, EPS)      # src/tracking.jl:1611-1612
        sy2 = max(e.rmssizey * e.rmssizey, EPS)
        return TAGS["GAUSSBB"], 1, 0, 0, -1, 0.0, [sx2, sy2, 1.0, e.amplitude, e.xoffset, e.yoffset], None
    if isinstance(e, StrongGaussianBeam):
        n = int(e.nzslice)
        if n <= 0:
            return TAGS["MARKER"], 0, 0, 0, -1, 0.0, [], None
        sx2 = max(float(e.beamsize[0]) ** 2, EPS)    # src/tracking.jl:1625-1627
        sy2 = max(float(e.beamsize[1]) ** 2, EPS)
        blk = [sx2, sy2, e.charge / max(abs(e.total_energy), EPS)]
        for i in range(n):
            w = float(e.zslice_npar[i]) if i < len(e.zslice_npar) else 1.0 / n
            xo = float(e.xoffsets[i]) if i < len(e.xoffsets) else 0.0
            yo = float(e.yoffsets[i]) if i < len(e.yoffsets) else 0.0
            blk += [w, xo, yo]
        return TAGS["GAUSSBB"], n, 0, 0, -1, 0.0, blk, None
    if isinstance(e, LongitudinalRLCWake):
        if e.scale == 0.0:
            return TAGS["MARKER"], 0, 0, 0, -1, 0.0, [], None
        Q0p = math.sqrt(e.Q0 * e.Q0 - 0.25) if e.Q0 * e.Q0 >= 0.25 else math.nan  # src/elements/wake.jl:25-36
        w0 = TWO_PI * e.freq
        w0p = w0 / e.Q0 * Q0p
        blk = [e.scale, e.Rshunt * w0 / e.Q0, w0p, 2 * Q0p, w0, 2 * e.Q0]
        return TAGS["RLCWAKE"], 0, 0, 0, -1, 0.0, blk, None
    if isinstance(e, LongitudinalWake):
        if e.scale == 0.0:
            return TAGS["MARKER"], 0, 0, 0, -1, 0.0, [], None
        n = len(e.times)
        blk = [e.scale, e.fliphalf, float(e.times[0]) * e.fliphalf]
        blk += list(map(float, e.times)) + list(map(float, e.wakefields))
        return TAGS["TABWAKE"], n, 0, 0, -1, 0.0, blk, None
    if isinstance(e, Wiggler):
        return _wiggler_record(e)
    if isinstance(e, LBend):
        mflags, mblk = _misalign(e)
        if e.L == 0.0:                   # src/tracking.jl:1158-1160
            return TAGS["NAN"], 0, 0, 0, -1, 0.0, [], None
        irho = e.angle / e.L
        blk = [e.L, e.K, e.angle, e.by_error, e.angle / e.L]
        blk += _edge_block(irho, e.e1, e.fint1, e.full_gap, 1) + _edge_block(irho, e.e2, e.fint2, e.full_gap, 1)
        return TAGS["LBEND"], 0, 0, mflags | (1 << F_FIN_SHIFT) | (1 << F_FOUT_SHIFT), -1, 0.0, blk, mblk
    return None


def compile_table(lat: Lattice, beam: Beam, *, aperture: Optional[float] = None,
                  exact_hamiltonian: bool = True, coord_limit: float = 1.0,
                  angle_limit: float = 1.0, info: Optional[TableInfo] = None,
                  turn: int = 0, nturns: Optional[int] = None, time: float = 0.0,
                  dt_turn: float = 0.0) -> bytes:
    """Compile a lattice into the device table blob.

    ``aperture`` is the radius R of an optional circular aperture (x^2 + y^2 > R^2 at an
    element exit marks the particle lost, in addition to check_lost).  ``exact_hamiltonian``
    mirrors the reference's global ``USE_EXACT_HAMILTONIAN`` (src/tracking.jl:38).

    A lattice with ``TimeVaryingElement``s is resolved per turn, as ``ringpass!`` resolves it
    in every ``linepass!`` call (src/lattice.jl:267, 305-317): row k of the table is the
    lattice at turn ``turn + k`` and time ``time`` advanced by ``dt_turn`` k times, for
    ``nturns`` rows (one row, at (time, turn), when ``nturns`` is None).  Static elements
    share one record across rows; identical materialised elements share records too.
    """
    bt = beti(beam)
    pool: List[float] = []
    recs: List[Tuple[int, ...]] = []
    rec_index: Dict[tuple, int] = {}
    if info is not None:
        info.seq, info.gorder, info.gscale = [], [], []

    def rec_of(e):
        tag, n, m, flags, gorder, gscale, blk, mblk = _record(e, bt)
        key = (tag, n, m, flags, gorder, tuple(struct.pack("<%dd" % len(blk), *blk)),
               None if mblk is None else tuple(struct.pack("<84d", *mblk)))
        idx = rec_index.get(key)
        if idx is None:
            off = len(pool)
            pool.extend(blk)
            moff = -1
            if mblk is not None:
                moff = len(pool)
                pool.extend(mblk)
            idx = len(recs)
            recs.append((tag, n, m, off, flags, moff, gorder, 0))
            rec_index[key] = idx
        return idx, gorder, gscale

    varying = [isinstance(e, TimeVaryingElement) for e in lat.elements]
    rows = 1 if (not any(varying) or nturns is None) else max(int(nturns), 1)
    static = [None if v else rec_of(e) for e, v in zip(lat.elements, varying)]
    seq: List[int] = []
    t = float(time)
    for k in range(rows):
        for i, (e, st) in enumerate(zip(lat.elements, static)):
            r = st if st is not None else rec_of(e.materialize(t, int(turn) + k))
            seq.append(r[0])
            if info is not None and k == 0:
                info.seq.append(r[0])
                info.gorder.append(r[1])
                info.gscale.append(r[2])
        t += float(dt_turn)  # accumulated as ringpass! does
    n_rows, row_turn0 = (rows, int(turn)) if rows > 1 else (0, 0)
    hdr = struct.pack(
        "<6IIi4d", TP_TABLE_MAGIC, TP_TABLE_VERSION, len(lat.elements), len(recs), len(pool),
        H_EXACT_HAMILTONIAN if exact_hamiltonian else 0, n_rows, row_turn0,
        bt, float(coord_limit), float(angle_limit),
        0.0 if aperture is None else float(aperture) * float(aperture))
    seq_b = struct.pack("<%di" % len(seq), *seq)
    if len(seq_b) % 8:
        seq_b += b"\0" * (8 - len(seq_b) % 8)
    rec_b = b"".join(struct.pack("<4iI3i", *r) for r in recs)
    pool_b = struct.pack("<%dd" % len(pool), *pool)
    if info is not None:
        info.n_seq, info.n_rec, info.pool_len = len(lat.elements), len(recs), len(pool)
    return hdr + seq_b + rec_b + pool_b


M_ELECTRON = 0.51099895069e6  # src/lattice.jl:42


def _julia_round(v: float) -> int:
    """round(Int, v): nearest integer, ties to even (Python's round does the same)."""
    return int(round(v))


def _wiggler_record(e):
    """TP_WIGGLER: the step plan of _wig_pass_4th / _wig_pass_4th_rad and the
    particle-independent leading factors of every harmonic term (src/tracking.jl:1224-1525)."""
    PN = int(e.Nsteps)
    Nw = _julia_round(e.L / e.lw)
    Nstep = PN * max(1, Nw)
    dl = e.lw / float(PN)
    dl1 = dl * 1.3512071919596573
    dl0 = dl * -1.7024143839193146
    gamma = e.energy / M_ELECTRON
    Po = math.sqrt(gamma * gamma - 1.0)
    Aw = 1e-9 * C_LIGHT / (M_ELECTRON * 1e-9) / TWO_PI * e.lw * e.Bmax
    srCoef = (2.0 / 3.0) * 2.8179403205e-15 * (gamma * gamma * gamma)
    H = Po * M_ELECTRON / C_LIGHT
    kw = TWO_PI / e.lw
    PB0 = e.Bmax
    nh = min(e.NHharm, len(e.By) // 6)
    nv = min(e.NVharm, len(e.Bx) // 6)
    blk = [float(Nstep), dl1, dl0, float(nh), float(nv), float(e.rad), H, srCoef, dl]
    for i in range(nh):
        b = e.By[6 * i: 6 * i + 6]
        HCw = float(b[1]) * Aw / Po
        kx, ky, kz, tz = float(b[2]) * kw, float(b[3]) * kw, float(b[4]) * kw, float(b[5])
        a = HCw * (kw / kz)
        blk += [kx, ky, kz, tz,
                a, a * ky,                                   # ax, axpy
                a * (kx / ky), a * ((kx / ky) * (kx / ky)),  # ay, aypx
                PB0 * float(b[1]) * (kx / ky), PB0 * float(b[1]),  # Bx, By of _wig_B
                1.0 if abs(kx / kw) > 1e-6 else 0.0]
    for i in range(nv):
        b = e.Bx[6 * i: 6 * i + 6]
        VCw = float(b[1]) * Aw / Po
        kx, ky, kz, tz = float(b[2]) * kw, float(b[3]) * kw, float(b[4]) * kw, float(b[5])
        a = VCw * (kw / kz)
        blk += [kx, ky, kz, tz,
                a * (ky / kx), a * ((ky / kx) * (ky / kx)),  # ax, axpy
                a, a * kx,                                   # ay, aypx
                PB0 * float(b[1]), PB0 * float(b[1]) * (ky / kx),
                1.0 if abs(ky / kw) > 1e-6 else 0.0]
    mflags, mblk = _misalign(e)
    return TAGS["WIGGLER"], Nstep, 0, mflags, -1, 0.0, blk, mblk
