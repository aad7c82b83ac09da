"""Device tracking: the drop-in for TrackPad's multi-particle path.

Reference entry points and their counterparts here:

    linepass!(coords, lat, beam, lost_flags; time, turn)            src/lattice.jl:260-297
        -> linepass_(coords, lat, beam, lost_flags, time=, turn=)
    ringpass!(coords, lat, beam, lost_flags, nturns; time, dt_turn, turn)
                                                                    src/lattice.jl:305-317
        -> ringpass_(coords, lat, beam, lost_flags, nturns, time=, dt_turn=, turn=)
    track!(particles, lattice)   (named in the reference's plan, MIGRATION_PLAN.md:116)
        -> track_(particles, lat, beam, nturns=1, turn=0)   on device-resident particles

The trailing underscore replaces Julia's ``!`` (in-place).  ``coords`` is an N x 6 float64
array (x, px, y, py, z, delta), ``lost_flags`` an int64 array of 0/1, both mutated in
place with the reference's semantics: a particle whose flag is 1 is skipped; a particle
lost at an element exit keeps its element-entry coordinates and gets flag 1.

Every call runs on the GPU through libtrackpad_hip; there is no CPU path.
"""

from __future__ import annotations

import atexit
import ctypes as C
import hashlib
import threading
import weakref
from typing import Dict, Optional

import numpy as np

from . import _lib as L
from .lattice import Beam, Lattice
from .table import TableInfo, compile_table

__all__ = ["HIPContext", "DeviceTable", "HIPParticles", "default_context", "linepass_",
           "ringpass_", "track_", "device_table"]

_PD = C.POINTER(C.c_double)
_PI64 = C.POINTER(C.c_int64)
_PI32 = C.POINTER(C.c_int32)


_SHUTTING_DOWN = [False]


@atexit.register
def _mark_shutdown() -> None:
    # Past this point the process is ending: device memory goes with it, and no HIP call
    # is made from objects collected during interpreter teardown.
    _SHUTTING_DOWN[0] = True


def _release(fn, handle) -> None:
    if not _SHUTTING_DOWN[0]:
        fn(handle)


def _ptr(a: Optional[np.ndarray], t):
    return None if a is None else a.ctypes.data_as(t)


class HIPContext:
    """One device and one HIP stream (tp_ctx)."""

    def __init__(self, device: int = 0):
        self._lib = L.lib()
        h = C.c_void_p()
        L.check(self._lib.tp_ctx_create(int(device), C.byref(h)))
        self.handle = h
        self.device = int(device)
        self._tables: Dict[bytes, "DeviceTable"] = {}
        self._fin = weakref.finalize(self, _release, self._lib.tp_ctx_destroy, h)
        self._fin.atexit = False

    def set_launch(self, block: int = 0, ppl: int = 0, turns_per_launch: int = 0) -> None:
        L.check(self._lib.tp_set_launch(self.handle, int(block), int(ppl), int(turns_per_launch)))

    def set_waves(self, fwd_waves: int = 0, back_waves: int = 0) -> None:
        """Register budgets (resident waves per SIMD) of the tracking and reverse kernels."""
        L.check(self._lib.tp_set_waves(self.handle, int(fwd_waves), int(back_waves)))

    def set_drift(self, form: str = "auto") -> None:
        """Drift form of the fast build for later calls: "auto" (the series for particle sets
        that fill the GPU, fixed per set), "series" or "rsqrt" (tp_set_drift)."""
        code = {"auto": L.TP_DRIFT_AUTO, "series": L.TP_DRIFT_SERIES, "rsqrt": L.TP_DRIFT_RSQRT}.get(form)
        if code is None:
            raise ValueError(f"drift form {form!r} is not 'auto', 'series' or 'rsqrt'")
        L.check(self._lib.tp_set_drift(self.handle, code))

    def autotune(self, table: "DeviceTable", particles: "HIPParticles", probe_turns: int = 2, *, fast: bool = True,
                 target_launch_ms: float = 0.0, cache: Optional[str] = None) -> dict:
        """Time the candidate launch shapes on a copy of ``particles`` and apply the fastest
        (tp_autotune); ``cache`` is a text file of earlier choices."""
        cfg = L.LaunchCfg()
        L.check(self._lib.tp_autotune(self.handle, table.handle, particles.handle, int(probe_turns),
                                      L.TP_TRACK_FAST if fast else 0, float(target_launch_ms),
                                      None if cache is None else str(cache).encode(), C.byref(cfg)))
        return _cfg_dict(cfg)

    def timing(self) -> dict:
        t = L.TimingT()
        L.check(self._lib.tp_timing(self.handle, C.byref(t)))
        return {"kernel_ms": t.kernel_ms, "launches": t.launches, "block": t.block, "ppl": t.ppl,
                "turns_per_launch": t.turns_per_launch, "forward_ms": t.forward_ms, "backward_ms": t.backward_ms}

    def sync(self) -> None:
        L.check(self._lib.tp_sync(self.handle))

    def table(self, blob: bytes) -> "DeviceTable":
        """Upload (or reuse) the device table for a compiled blob."""
        key = hashlib.sha1(blob).digest()
        t = self._tables.get(key)
        if t is None:
            t = DeviceTable(self, blob)
            if len(self._tables) > 64:
                self._tables.clear()
            self._tables[key] = t
        return t

    def close(self) -> None:
        self._tables.clear()
        self._fin()


_default: Dict[int, HIPContext] = {}
_default_lock = threading.Lock()


def default_context(device: int = 0) -> HIPContext:
    with _default_lock:
        c = _default.get(device)
        if c is None:
            c = HIPContext(device)
            _default[device] = c
        return c


class DeviceTable:
    """A validated element table resident on the device (tp_table)."""

    def __init__(self, ctx: HIPContext, blob: bytes, info: Optional[TableInfo] = None):
        self.ctx = ctx
        self.blob = bytes(blob)
        self.info = info
        h = C.c_void_p()
        buf = C.create_string_buffer(self.blob, len(self.blob))
        L.check(ctx._lib.tp_table_create(ctx.handle, buf, len(self.blob), C.byref(h)))
        self.handle = h
        ns, nr, bt = C.c_int32(), C.c_int32(), C.c_double()
        L.check(ctx._lib.tp_table_info(h, C.byref(ns), C.byref(nr), C.byref(bt)))
        self.n_seq, self.n_rec, self.beti = ns.value, nr.value, bt.value
        self._fin = weakref.finalize(self, _release, ctx._lib.tp_table_destroy, h)
        self._fin.atexit = False


def device_table(lat: Lattice, beam: Beam, ctx: Optional[HIPContext] = None, **table_kw) -> DeviceTable:
    ctx = ctx or default_context()
    return ctx.table(compile_table(lat, beam, **table_kw))


class HIPParticles:
    """Device-resident particle set: 6 planes of N doubles plus per-particle loss record."""

    def __init__(self, n: int, ctx: Optional[HIPContext] = None):
        self.ctx = ctx or default_context()
        h = C.c_void_p()
        L.check(self.ctx._lib.tp_particles_create(self.ctx.handle, int(n), C.byref(h)))
        self.handle = h
        self.n = int(n)
        self._fin = weakref.finalize(self, _release, self.ctx._lib.tp_particles_destroy, h)
        self._fin.atexit = False

    @classmethod
    def from_coords(cls, coords: np.ndarray, lost_flags: Optional[np.ndarray] = None,
                    ctx: Optional[HIPContext] = None) -> "HIPParticles":
        coords = _as_coords(coords)
        p = cls(coords.shape[0], ctx)
        p.upload(coords, lost_flags)
        return p

    def upload(self, coords: np.ndarray, lost_flags: Optional[np.ndarray] = None,
               lost_turn: Optional[np.ndarray] = None, lost_elem: Optional[np.ndarray] = None) -> None:
        coords = np.asfortranarray(_as_coords(coords))
        if coords.shape[0] != self.n:
            raise ValueError(f"coords has {coords.shape[0]} rows, particle set holds {self.n}")
        lf = None if lost_flags is None else np.ascontiguousarray(lost_flags, dtype=np.int64)
        lt = None if lost_turn is None else np.ascontiguousarray(lost_turn, dtype=np.int32)
        le = None if lost_elem is None else np.ascontiguousarray(lost_elem, dtype=np.int32)
        for a, nm in ((lf, "lost_flags"), (lt, "lost_turn"), (le, "lost_elem")):
            if a is not None and a.shape != (self.n,):
                raise ValueError(f"{nm} must have shape ({self.n},)")
        L.check(self.ctx._lib.tp_particles_upload(self.handle, _ptr(coords, _PD), _ptr(lf, _PI64),
                                                  _ptr(lt, _PI32), _ptr(le, _PI32)))

    def download(self):
        """(coords N x 6, lost_flags int64, lost_turn int32, lost_elem int32) on the host."""
        coords = np.empty((self.n, 6), dtype=np.float64, order="F")
        lf = np.empty(self.n, dtype=np.int64)
        lt = np.empty(self.n, dtype=np.int32)
        le = np.empty(self.n, dtype=np.int32)
        L.check(self.ctx._lib.tp_particles_download(self.handle, _ptr(coords, _PD), _ptr(lf, _PI64),
                                                    _ptr(lt, _PI32), _ptr(le, _PI32)))
        return coords, lf, lt, le

    @property
    def coords(self) -> np.ndarray:
        return self.download()[0]

    def copy_from(self, other: "HIPParticles") -> None:
        """Device-to-device copy of another particle set of the same size (enqueued)."""
        L.check(self.ctx._lib.tp_particles_copy(self.handle, other.handle))

    def device_ptrs(self):
        pc, pt, pe, n = _PD(), _PI32(), _PI32(), C.c_int64()
        L.check(self.ctx._lib.tp_particles_device_ptrs(self.handle, C.byref(pc), C.byref(pt), C.byref(pe), C.byref(n)))
        return C.cast(pc, C.c_void_p).value, C.cast(pt, C.c_void_p).value, C.cast(pe, C.c_void_p).value, n.value


def _cfg_dict(cfg) -> dict:
    return {"block": cfg.block, "ppl": cfg.ppl, "turns_per_launch": cfg.turns_per_launch,
            "fwd_waves": cfg.fwd_waves, "back_waves": cfg.back_waves, "candidates": cfg.candidates,
            "rate": cfg.rate}


def _as_coords(coords) -> np.ndarray:
    a = np.asarray(coords, dtype=np.float64)
    if a.ndim != 2 or a.shape[1] != 6:
        raise ValueError(f"coords must be N x 6, got shape {a.shape}")
    return a


def _flags(fast: bool, timed: bool = False) -> int:
    return (L.TP_TRACK_FAST if fast else 0) | (L.TP_TRACK_TIMED if timed else 0)


def ringpass_(coords: np.ndarray, lat: Lattice, beam: Beam, lost_flags: np.ndarray, nturns: int, *,
              time: float = 0.0, dt_turn: float = 0.0, turn: int = 0, fast: bool = False,
              aperture: Optional[float] = None, ctx: Optional[HIPContext] = None,
              lost_turn: Optional[np.ndarray] = None, lost_elem: Optional[np.ndarray] = None) -> None:
    """In-place multi-turn tracking on host arrays (ringpass!, src/lattice.jl:305-317).

    Time-varying elements are resolved at every turn k at time ``time`` advanced by
    ``dt_turn`` k times and turn ``turn + k``, as ``ringpass!`` passes them to ``linepass!``;
    on a static lattice ``time`` and ``dt_turn`` have no effect.
    """
    if not isinstance(coords, np.ndarray) or coords.dtype != np.float64 or coords.ndim != 2 or coords.shape[1] != 6:
        raise TypeError("coords must be a float64 numpy array of shape (N, 6)")
    if not isinstance(lost_flags, np.ndarray) or lost_flags.dtype != np.int64 or lost_flags.shape != (coords.shape[0],):
        raise TypeError("lost_flags must be an int64 numpy array of shape (N,)")
    nturns = int(nturns)
    if nturns < 0:
        raise ValueError(f"nturns = {nturns} < 0")
    ctx = ctx or default_context()
    tab = ctx.table(compile_table(lat, beam, aperture=aperture, turn=turn, nturns=nturns, time=time,
                                  dt_turn=dt_turn))
    n = coords.shape[0]
    work = coords if coords.flags.f_contiguous else np.asfortranarray(coords)
    lt = np.empty(n, dtype=np.int32) if lost_turn is None else lost_turn
    le = np.empty(n, dtype=np.int32) if lost_elem is None else lost_elem
    if not lost_flags.flags.c_contiguous:
        raise TypeError("lost_flags must be contiguous")
    L.check(ctx._lib.tp_track_host(ctx.handle, tab.handle, _ptr(work, _PD), n, _ptr(lost_flags, _PI64),
                                   _ptr(lt, _PI32), _ptr(le, _PI32), int(turn), nturns, _flags(fast)))
    if work is not coords:
        coords[...] = work


def linepass_(coords: np.ndarray, lat: Lattice, beam: Beam, lost_flags: np.ndarray, *,
              time: float = 0.0, turn: int = 0, **kw) -> None:
    """In-place single pass on host arrays (linepass!, src/lattice.jl:260-297)."""
    ringpass_(coords, lat, beam, lost_flags, 1, time=time, turn=turn, **kw)


def track_(particles: HIPParticles, lat, beam: Optional[Beam] = None, nturns: int = 1, *,
           turn: int = 0, fast: bool = False, timed: bool = False, aperture: Optional[float] = None,
           time: float = 0.0, dt_turn: float = 0.0, no_alloc: bool = False) -> None:
    """Track device-resident particles for ``nturns`` turns (``track!`` of the Julia package).

    ``lat`` is a :class:`Lattice` (compiled with ``beam``) or an already uploaded
    :class:`DeviceTable`.  ``no_alloc`` calls tp_track_noalloc: the same results without any
    device allocation in the call (at most TP_NOALLOC_MAX_N particles; overlapping calls on
    one device must be serialised by the caller).
    """
    ctx = particles.ctx
    if isinstance(lat, DeviceTable):
        tab = lat
    else:
        if beam is None:
            raise ValueError("track_ needs a Beam to compile a Lattice")
        tab = ctx.table(compile_table(lat, beam, aperture=aperture, turn=turn, nturns=nturns, time=time,
                                      dt_turn=dt_turn))
    fn = ctx._lib.tp_track_noalloc if no_alloc else ctx._lib.tp_track
    L.check(fn(ctx.handle, tab.handle, particles.handle, int(turn), int(nturns), _flags(fast, timed)))
