"""ctypes bridge to the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the CPU-baseline
leg of bench.py, never by the product.  It converts host element objects (the field names
of the reference's structs) into the oracle's raw element records; it deliberately does not
read the device table, so the oracle restates the reference independently of the table
compiler.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# TRACKPAD_ORACLE_LIB names another build of the same source (the sanitizer build of the tests)
LIB = os.environ.get("TRACKPAD_ORACLE_LIB", os.path.join(_HERE, "liboracle.so"))
NP = 11  # MAXORD + 1

ORC_MARKER, ORC_DRIFT, ORC_QUAD, ORC_SEXT, ORC_OCT, ORC_THINMP, ORC_SBEND, ORC_EXACTSBEND, \
    ORC_RFCAVITY, ORC_EXACTSBEND_LE = range(10)
(ORC_SOLENOID, ORC_CORRECTOR, ORC_TRANSLATION, ORC_YROTATION, ORC_CRAB, ORC_ACCEL, ORC_LONGRF, ORC_BOOST,
 ORC_INVBOOST, ORC_THINBB, ORC_SGB, ORC_RLCWAKE, ORC_TABWAKE, ORC_DRIFTSC, ORC_QUADSC, ORC_SEXTSC, ORC_OCTSC,
 ORC_SBENDSC, ORC_LBEND, ORC_SPACECHARGE, ORC_WIGGLER) = range(10, 31)
NX = 128


class OrcElem(C.Structure):
    _fields_ = [
        ("kind", C.c_int32), ("max_order", C.c_int32), ("num_int_steps", C.c_int32),
        ("fringe_bend_entrance", C.c_int32), ("fringe_bend_exit", C.c_int32),
        ("has_r1", C.c_int32), ("has_r2", C.c_int32), ("pad", C.c_int32),
        ("L", C.c_double), ("k", C.c_double), ("angle", C.c_double), ("e1", C.c_double),
        ("e2", C.c_double), ("fint1", C.c_double), ("fint2", C.c_double), ("gap", C.c_double),
        ("gk", C.c_double), ("kick_angle", C.c_double * 2),
        ("polynom_a", C.c_double * NP), ("polynom_b", C.c_double * NP),
        ("t1", C.c_double * 6), ("t2", C.c_double * 6),
        ("r1", C.c_double * 36), ("r2", C.c_double * 36),
        ("volt", C.c_double), ("freq", C.c_double), ("h", C.c_double), ("lag", C.c_double),
        ("philag", C.c_double), ("energy", C.c_double), ("x", C.c_double * NX),
    ]


_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB) and "TRACKPAD_ORACLE_LIB" not in os.environ:
            build()
        L = C.CDLL(LIB)
        PD, PI64, PI32, PE = C.POINTER(C.c_double), C.POINTER(C.c_int64), C.POINTER(C.c_int32), C.POINTER(OrcElem)
        L.orc_beti.argtypes = [C.c_double, C.c_double]
        L.orc_beti.restype = C.c_double
        L.orc_pass.argtypes = [PE, PD, C.c_double, C.c_int]
        L.orc_linepass.argtypes = [PE, C.c_int, PD, C.c_int64, PI64, C.c_double, C.c_double, C.c_int, PI32, PI32, C.c_int32]
        L.orc_ringpass.argtypes = [PE, C.c_int, PD, C.c_int64, PI64, C.c_double, C.c_double, C.c_int, PI32, PI32,
                                   C.c_int32, C.c_int32, C.c_int]
        L.orc_ringpass_ld.argtypes = L.orc_ringpass.argtypes
        L.orc_ringpass_q.argtypes = L.orc_ringpass.argtypes + [PD, PD]
        L.orc_jacobian.argtypes = [PE, C.c_int, PD, C.c_int32, C.c_double, C.c_int, PD, PD]
        L.orc_loss_dir.argtypes = [PE, C.c_int, PD, C.c_int64, PD, PD, C.c_int32, C.c_double, C.c_double, C.c_int,
                                   PD, PD, PI64]
        L.orc_max_threads.restype = C.c_int
        _lib = L
    return _lib


def _kind(e) -> int:
    n = type(e).__name__
    return {"Marker": ORC_MARKER, "Drift": ORC_DRIFT, "Quadrupole": ORC_QUAD, "Sextupole": ORC_SEXT,
            "Octupole": ORC_OCT, "ThinMultipole": ORC_THINMP, "SBend": ORC_SBEND,
            "ExactSBend": ORC_EXACTSBEND, "RFCavity": ORC_RFCAVITY,
            "ExactSBendLinearEdge": ORC_EXACTSBEND_LE, "Solenoid": ORC_SOLENOID, "Corrector": ORC_CORRECTOR,
            "Translation": ORC_TRANSLATION, "YRotation": ORC_YROTATION, "CrabCavity": ORC_CRAB,
            "AccelCavity": ORC_ACCEL, "LongitudinalRFMap": ORC_LONGRF, "LorentzBoost": ORC_BOOST,
            "InvLorentzBoost": ORC_INVBOOST, "StrongThinGaussianBeam": ORC_THINBB, "StrongGaussianBeam": ORC_SGB,
            "LongitudinalRLCWake": ORC_RLCWAKE, "LongitudinalWake": ORC_TABWAKE, "DriftSC": ORC_DRIFTSC,
            "QuadrupoleSC": ORC_QUADSC, "SextupoleSC": ORC_SEXTSC, "OctupoleSC": ORC_OCTSC, "SBendSC": ORC_SBENDSC,
            "LBend": ORC_LBEND, "SpaceCharge": ORC_SPACECHARGE, "Wiggler": ORC_WIGGLER}[n]


def _extra_fields(e, o) -> None:
    """Per-kind extra fields in orc_elem.x (the raw fields of the reference structs)."""
    n = type(e).__name__
    x = []
    if n == "Solenoid":
        o.k = e.ks
    elif n == "Corrector":
        x = [e.xkick, e.ykick]
    elif n == "Translation":
        x = [e.dx, e.dy, e.ds]
    elif n == "CrabCavity":
        x = [e.k, e.phi, float(e.errors[0]), float(e.errors[1])]
    elif n == "AccelCavity":
        x = [e.k, e.phis]
    elif n == "LongitudinalRFMap":
        rf = e.rf
        kind = {"RFCavity": 1, "AccelCavity": 2, "CrabCavity": 3}.get(type(rf).__name__, 0)
        x = [e.alphac, kind, getattr(rf, "freq", 0.0), getattr(rf, "k", 0.0), getattr(rf, "h", 1.0)]
    elif n == "LorentzBoost":
        x = [e.cosang, e.tanang, e.mode]
    elif n == "InvLorentzBoost":
        x = [e.sinang, e.cosang, e.mode]
    elif n == "StrongThinGaussianBeam":
        x = [e.amplitude, e.rmssizex, e.rmssizey, e.xoffset, e.yoffset]
    elif n == "StrongGaussianBeam":
        k = int(e.nzslice)
        if 3 * k + 8 > NX:
            raise ValueError("the oracle holds at most 40 beam-beam slices")
        w = [float(e.zslice_npar[i]) if i < len(e.zslice_npar) else 1.0 / k for i in range(k)]
        xo = [float(e.xoffsets[i]) if i < len(e.xoffsets) else 0.0 for i in range(k)]
        yo = [float(e.yoffsets[i]) if i < len(e.yoffsets) else 0.0 for i in range(k)]
        x = [e.charge, e.total_energy, float(e.beamsize[0]), float(e.beamsize[1]), k, 0.0, 0.0, 0.0] + w + xo + yo
    elif n == "LongitudinalRLCWake":
        x = [e.freq, e.Rshunt, e.Q0, e.scale]
    elif n == "LongitudinalWake":
        k = len(e.times)
        if 2 * k + 3 > NX:
            raise ValueError("the oracle holds at most 62 wake samples")
        x = [e.scale, e.fliphalf, k] + list(map(float, e.times)) + list(map(float, e.wakefields))
    elif n in ("QuadrupoleSC", "SextupoleSC", "OctupoleSC"):
        x = [e.k0, e.k1, e.k2, e.k3]
    elif n == "LBend":
        x = [e.K, e.by_error]
        o.gap = e.full_gap
    elif n == "Wiggler":
        nh, nv = min(e.NHharm, len(e.By) // 6), min(e.NVharm, len(e.Bx) // 6)
        if 6 + 6 * (nh + nv) > NX:
            raise ValueError("the oracle holds at most 20 wiggler harmonics")
        x = [e.lw, e.Bmax, e.Nsteps, e.rad, nh, nv] + list(map(float, e.By[:6 * nh])) + list(map(float, e.Bx[:6 * nv]))
    for i, v in enumerate(x):
        o.x[i] = float(v)


def to_orc(e) -> OrcElem:
    o = OrcElem()
    o.kind = _kind(e)
    g = lambda k, d=0.0: float(getattr(e, k, d))
    o.max_order = int(getattr(e, "max_order", 0))
    o.num_int_steps = int(getattr(e, "num_int_steps", 0))
    o.fringe_bend_entrance = int(getattr(e, "fringe_bend_entrance", 0))
    o.fringe_bend_exit = int(getattr(e, "fringe_bend_exit", 0))
    o.L = g("L")
    o.k = g("k1") if hasattr(e, "k1") else (g("k2") if hasattr(e, "k2") else g("k3"))
    for f in ("angle", "e1", "e2", "fint1", "fint2", "gap", "gk", "volt", "freq", "h", "lag", "philag", "energy"):
        setattr(o, f, g(f))
    if hasattr(e, "kick_angle"):
        o.kick_angle[0], o.kick_angle[1] = float(e.kick_angle[0]), float(e.kick_angle[1])
    for f in ("polynom_a", "polynom_b"):
        if hasattr(e, f):
            v = np.asarray(getattr(e, f), dtype=np.float64)
            arr = getattr(o, f)
            for i, x in enumerate(v):
                arr[i] = float(x)
    for f in ("t1", "t2"):
        if hasattr(e, f):
            arr = getattr(o, f)
            for i, x in enumerate(np.asarray(getattr(e, f), dtype=np.float64)):
                arr[i] = float(x)
    for f, flag in (("r1", "has_r1"), ("r2", "has_r2")):
        if hasattr(e, f):
            M = np.asarray(getattr(e, f), dtype=np.float64)
            if np.any(M != 0.0):
                setattr(o, flag, 1)
            arr = getattr(o, f)
            for i, x in enumerate(M.reshape(-1, order="F")):
                arr[i] = float(x)
    _extra_fields(e, o)
    return o


def elems(lat) -> C.Array:
    els = list(lat.elements if hasattr(lat, "elements") else lat)
    arr = (OrcElem * max(1, len(els)))()
    for i, e in enumerate(els):
        arr[i] = to_orc(e)
    return arr, len(els)


def beti(beam) -> float:
    return float(lib().orc_beti(beam.energy, beam.mass))


def _f64(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def ringpass(coords: np.ndarray, lat, beam, lost_flags: np.ndarray, nturns: int, *, turn: int = 0,
             aperture=None, exact: bool = True, nthreads: int = 1, extended: bool = False, quad: bool = False):
    """Oracle ringpass! on host arrays; returns (coords, lost_flags, lost_turn, lost_elem) copies.
    ``extended`` runs the same algorithm in long double, ``quad`` in __float128 (results
    rounded to double)."""
    c = np.asfortranarray(np.array(coords, dtype=np.float64))
    n = c.shape[0]
    lf = np.ascontiguousarray(np.array(lost_flags, dtype=np.int64))
    lt = np.where(lf == 1, -2, -1).astype(np.int32)
    le = lt.copy()
    arr, ne = elems(lat)
    ap2 = 0.0 if aperture is None else float(aperture) ** 2
    args = (arr, ne, _f64(c), n, lf.ctypes.data_as(C.POINTER(C.c_int64)), beti(beam), ap2, int(exact),
            lt.ctypes.data_as(C.POINTER(C.c_int32)), le.ctypes.data_as(C.POINTER(C.c_int32)), int(turn), int(nturns),
            int(nthreads))
    if quad:
        lib().orc_ringpass_q(*args, None, None)
    else:
        (lib().orc_ringpass_ld if extended else lib().orc_ringpass)(*args)
    return np.ascontiguousarray(c), lf, lt, le


def loss_quad(coords, lat, beam, nturns: int, *, aperture=None, nthreads: int = 1):
    """ringpass! in quad precision: (sum over survivors of x^2 + y^2 as a double pair
    (hi, lo), number of survivors)."""
    c = np.asfortranarray(np.array(coords, dtype=np.float64))
    n = c.shape[0]
    lf = np.zeros(n, dtype=np.int64)
    lt = np.full(n, -1, dtype=np.int32)
    le = lt.copy()
    arr, ne = elems(lat)
    ap2 = 0.0 if aperture is None else float(aperture) ** 2
    hi, lo = C.c_double(), C.c_double()
    lib().orc_ringpass_q(arr, ne, _f64(c), n, lf.ctypes.data_as(C.POINTER(C.c_int64)), beti(beam), ap2, 1,
                         lt.ctypes.data_as(C.POINTER(C.c_int32)), le.ctypes.data_as(C.POINTER(C.c_int32)), 0,
                         int(nturns), int(nthreads), C.byref(hi), C.byref(lo))
    return (hi.value, lo.value), int(n - lf.sum())


def linepass(coords, lat, beam, lost_flags, *, turn: int = 0, **kw):
    return ringpass(coords, lat, beam, lost_flags, 1, turn=turn, **kw)


def pass_elem(e, r6, beam, exact: bool = True) -> np.ndarray:
    r = np.array(r6, dtype=np.float64).copy()
    o = to_orc(e)
    lib().orc_pass(C.byref(o), _f64(r), beti(beam), int(exact))
    return r


def jacobian(lat, beam, r0, nturns: int = 1, exact: bool = True):
    """(r_out, J, lost) with J[i, j] = d r_out[i] / d r0[j] by forward-mode duals."""
    arr, ne = elems(lat)
    r = np.array(r0, dtype=np.float64)
    ro = np.zeros(6)
    J = np.zeros(36)
    lost = lib().orc_jacobian(arr, ne, _f64(r), int(nturns), beti(beam), int(exact), _f64(ro), _f64(J))
    return ro, J.reshape(6, 6), bool(lost)


def loss_dir(lat, beam, coords0, nturns, dk=None, dr0=None, *, aperture=None, exact=True):
    """(loss, dloss, n_alive): loss = sum over survivors of x^2 + y^2 after nturns and its
    directional derivative along dk (per element strength k1/k2/k3) and dr0 (N x 6)."""
    arr, ne = elems(lat)
    c = np.asfortranarray(np.array(coords0, dtype=np.float64))
    n = c.shape[0]
    dkp = None
    if dk is not None:
        dk = np.ascontiguousarray(np.array(dk, dtype=np.float64))
        assert dk.shape == (ne,)
        dkp = _f64(dk)
    drp = None
    if dr0 is not None:
        dr0 = np.asfortranarray(np.array(dr0, dtype=np.float64))
        drp = _f64(dr0)
    loss, dloss, na = C.c_double(), C.c_double(), C.c_int64()
    ap2 = 0.0 if aperture is None else float(aperture) ** 2
    lib().orc_loss_dir(arr, ne, _f64(c), n, dkp, drp, int(nturns), beti(beam), ap2, int(exact), C.byref(loss),
                       C.byref(dloss), C.byref(na))
    return loss.value, dloss.value, na.value


def max_threads() -> int:
    return int(lib().orc_max_threads())
