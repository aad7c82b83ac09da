"""Batched lattice variants: many (lattice, beam) pairs in one persistent launch.

No reference counterpart: a parameter sweep over lattices in the reference is a loop of
``ringpass!`` calls, one lattice each (src/lattice.jl:305-317).  :class:`HIPBatch` binds
the variants' device tables and particle sets (tp_batch_create) and tracks all of them --
or tracks them and returns each variant's adjoint gradients -- with one hipGraph launch of
the persistent kernels (tp_batch_track / tp_batch_adjoint).  The results are those of
``track_`` / ``ringpass_grad`` on each variant alone.
"""

from __future__ import annotations

import ctypes as C
import weakref
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np

from . import _lib as L
from .tracking import DeviceTable, HIPContext, HIPParticles, _cfg_dict, _flags, _release, default_context

__all__ = ["HIPBatch", "BatchGradients"]

_PD = C.POINTER(C.c_double)
_PI64 = C.POINTER(C.c_int64)
_PI32 = C.POINTER(C.c_int32)


@dataclass
class BatchGradients:
    """Per-variant results of :meth:`HIPBatch.adjoint` (lists indexed by variant).

    ``loss[v]`` is mean(x^2 + y^2) over the survivors of variant v, ``dk[v]`` has one entry
    per lattice position (dL/dk of its main strength, 0 where it has none), ``dx0[v]`` is
    the N_v x 6 gradient with respect to the initial coordinates.
    """

    loss: np.ndarray
    n_alive: np.ndarray
    dk: List[np.ndarray]
    dx0: List[np.ndarray]


class HIPBatch:
    """Variants v = 0..n-1, each a device table and a device particle set (tp_batch)."""

    def __init__(self, tables: Sequence[DeviceTable], particles: Sequence[HIPParticles],
                 ctx: Optional[HIPContext] = None):
        if len(tables) != len(particles):
            raise ValueError(f"{len(tables)} tables for {len(particles)} particle sets")
        self.ctx = ctx or (particles[0].ctx if len(particles) else default_context())
        self.tables = list(tables)
        self.particles = list(particles)
        nv = len(self.tables)
        th = (C.c_void_p * max(nv, 1))(*[t.handle.value for t in self.tables])
        ph = (C.c_void_p * max(nv, 1))(*[p.handle.value for p in self.particles])
        h = C.c_void_p()
        L.check(self.ctx._lib.tp_batch_create(self.ctx.handle, nv, C.cast(th, C.POINTER(C.c_void_p)),
                                              C.cast(ph, C.POINTER(C.c_void_p)), C.byref(h)))
        self.handle = h
        self.n_var = nv
        self._fin = weakref.finalize(self, _release, self.ctx._lib.tp_batch_destroy, h)
        self._fin.atexit = False

    def set_launch(self, block: int = 0, fwd_waves: int = 0, back_waves: int = 0) -> None:
        """Work-group size and register budgets (resident waves per SIMD) of the forward
        and reverse persistent kernels; 0 keeps the default."""
        L.check(self.ctx._lib.tp_batch_set_launch(self.handle, int(block), int(fwd_waves), int(back_waves)))

    def autotune(self, probe_turns, *, ckpt_stride: int = 10, adjoint: bool = False, fast: bool = True,
                 cache: Optional[str] = None) -> dict:
        """Time block sizes and register budgets of the persistent kernels on the bound
        beams (restored afterwards) and apply the fastest (tp_batch_autotune)."""
        nt, _ = self._turns(probe_turns, None)
        cfg = L.LaunchCfg()
        fl = (L.TP_TRACK_FAST if fast else 0) | (L.TP_TUNE_ADJOINT if adjoint else 0)
        L.check(self.ctx._lib.tp_batch_autotune(self.ctx.handle, self.handle, nt.ctypes.data_as(_PI32),
                                                int(ckpt_stride), fl, None if cache is None else str(cache).encode(),
                                                C.byref(cfg)))
        return _cfg_dict(cfg)

    def info(self) -> dict:
        ni, gf, gb, cb = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int64()
        L.check(self.ctx._lib.tp_batch_info(self.handle, C.byref(ni), C.byref(gf), C.byref(gb), C.byref(cb)))
        return {"n_items": ni.value, "grid_fwd": gf.value, "grid_back": gb.value, "ckpt_bytes": cb.value}

    def _turns(self, nturns, turn0):
        nt = np.ascontiguousarray(np.broadcast_to(np.asarray(nturns, dtype=np.int32), (self.n_var,)))
        t0 = None if turn0 is None else np.ascontiguousarray(
            np.broadcast_to(np.asarray(turn0, dtype=np.int32), (self.n_var,)))
        return nt, t0

    def track(self, nturns, turn0=None, *, fast: bool = False, timed: bool = False,
              asynchronous: bool = False) -> None:
        """Track variant v for nturns[v] turns (scalars broadcast to every variant)."""
        nt, t0 = self._turns(nturns, turn0)
        fl = _flags(fast, timed) | (L.TP_TRACK_ASYNC if asynchronous else 0)
        L.check(self.ctx._lib.tp_batch_track(self.ctx.handle, self.handle,
                                             None if t0 is None else t0.ctypes.data_as(_PI32),
                                             nt.ctypes.data_as(_PI32), fl))

    def adjoint(self, nturns, turn0=None, *, ckpt_stride: int = 10, fast: bool = False,
                timed: bool = False) -> BatchGradients:
        """Track every variant and return its loss and gradients (tp_batch_adjoint)."""
        nt, t0 = self._turns(nturns, turn0)
        nv = self.n_var
        loss = np.zeros(max(nv, 1))
        na = np.zeros(max(nv, 1), dtype=np.int64)
        nseq = [t.n_seq for t in self.tables]
        dk = np.zeros(max(sum(nseq), 1))
        sizes = [p.n for p in self.particles]
        dx0 = np.zeros(max(6 * sum(sizes), 1))
        L.check(self.ctx._lib.tp_batch_adjoint(self.ctx.handle, self.handle,
                                               None if t0 is None else t0.ctypes.data_as(_PI32),
                                               nt.ctypes.data_as(_PI32), int(ckpt_stride), _flags(fast, timed),
                                               loss.ctypes.data_as(_PD), na.ctypes.data_as(_PI64),
                                               dk.ctypes.data_as(_PD), dx0.ctypes.data_as(_PD)))
        dks, dxs = [], []
        o = 0
        for m in nseq:
            dks.append(dk[o:o + m].copy())
            o += m
        o = 0
        for n in sizes:
            dxs.append(dx0[o:o + 6 * n].reshape((n, 6), order="F").copy())
            o += 6 * n
        return BatchGradients(loss=loss[:nv].copy(), n_alive=na[:nv].copy(), dk=dks, dx0=dxs)

    def close(self) -> None:
        self._fin()
