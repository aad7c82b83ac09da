"""Reverse-mode gradients of multi-turn tracking (libtrackpad_hip's tp_adjoint).

    loss, grads = ringpass_grad(coords0, lat, beam, nturns)

returns ``L = mean over surviving particles of (x^2 + y^2)`` after ``nturns`` turns and
``grads.dk`` (dL/dk of each lattice position's main strength: k1 of a Quadrupole, k2 of a
Sextupole, k3 of an Octupole, 0 elsewhere) and ``grads.dx0`` (dL/d initial coordinates,
N x 6).  A lattice with time-varying elements is resolved per turn as in tracking (the
table holds one row per turn); the gradient of a time-varying magnet's strength is that of
an offset added to its strength at every turn.  The reference has no adjoint (its Enzyme
extension is a stub,
ext/TrackPadEnzymeExt.jl:14-17); this is the rule a Julia ChainRulesCore ``rrule`` /
Enzyme custom rule wraps, and :class:`TrackLoss` wraps it as a ``torch.autograd.Function``.
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import _lib as L
from .elements import Octupole, Quadrupole, Sextupole, TimeVaryingElement
from .lattice import Beam, Lattice
from .table import compile_table
from .tracking import HIPContext, HIPParticles, default_context

__all__ = ["Gradients", "ringpass_grad", "strength_positions", "set_strengths", "TrackLoss"]


@dataclass
class Gradients:
    loss: float
    n_alive: int
    dk: np.ndarray      # (n_positions,)
    dx0: np.ndarray     # (N, 6)
    final: np.ndarray   # tracked beam (N, 6)
    lost_flags: np.ndarray
    forward_ms: float = 0.0
    backward_ms: float = 0.0


def strength_positions(lat: Lattice) -> np.ndarray:
    """0-based positions whose element has a strength gradient (quads, sextupoles, octupoles,
    static or time-varying)."""
    def magnet(e):
        e = e.base if isinstance(e, TimeVaryingElement) else e
        return isinstance(e, (Quadrupole, Sextupole, Octupole))
    return np.array([i for i, e in enumerate(lat.elements) if magnet(e)], dtype=np.int64)


def _strength_attr(e) -> str:
    return "k1" if isinstance(e, Quadrupole) else ("k2" if isinstance(e, Sextupole) else "k3")


def set_strengths(lat: Lattice, positions, values) -> Lattice:
    """A copy of ``lat`` with the main strengths at ``positions`` replaced by ``values``."""
    import copy

    els = list(lat.elements)
    for p, v in zip(positions, values):
        e = copy.copy(els[int(p)])
        setattr(e, _strength_attr(e), float(v))
        els[int(p)] = e
    return Lattice(els, name=lat.name)


def ringpass_grad(coords0, lat: Lattice, beam: Beam, nturns: int, *, turn: int = 0, ckpt_stride: int = 1,
                  fast: bool = False, timed: bool = False, ctx: Optional[HIPContext] = None,
                  aperture: Optional[float] = None, time: float = 0.0, dt_turn: float = 0.0,
                  exact_hamiltonian: bool = True) -> Gradients:
    ctx = ctx or default_context()
    tab = ctx.table(compile_table(lat, beam, aperture=aperture, turn=turn, nturns=nturns, time=time,
                                  dt_turn=dt_turn, exact_hamiltonian=exact_hamiltonian))
    p = HIPParticles.from_coords(coords0, ctx=ctx)
    n = p.n
    loss = C.c_double()
    na = C.c_int64()
    dk = np.zeros(max(tab.n_seq, 1))
    dx0 = np.zeros((n, 6), order="F")
    flags = (L.TP_TRACK_FAST if fast else 0) | (L.TP_TRACK_TIMED if timed else 0)
    L.check(ctx._lib.tp_adjoint(ctx.handle, tab.handle, p.handle, int(turn), int(nturns), int(ckpt_stride), flags,
                                C.byref(loss), C.byref(na), dk.ctypes.data_as(C.POINTER(C.c_double)),
                                dx0.ctypes.data_as(C.POINTER(C.c_double))))
    fin, lf, _, _ = p.download()
    g = Gradients(loss.value, na.value, dk[: tab.n_seq], np.ascontiguousarray(dx0), np.ascontiguousarray(fin), lf)
    if timed:
        t = ctx.timing()
        g.forward_ms, g.backward_ms = t["forward_ms"], t["backward_ms"]
    return g


_TRACKLOSS = None


def _make_trackloss():
    import torch

    class TrackLoss(torch.autograd.Function):
        """loss = TrackLoss.apply(k, coords0, lat, beam, nturns, positions)

        The autograd analogue of the Julia rrule: ``k`` (float64 tensor) holds the
        strengths at ``positions``; gradients flow to ``k`` and ``coords0`` through the
        hand-written adjoint kernel."""

        @staticmethod
        def forward(ctx_, k, coords0, lat, beam, nturns, positions):
            lat_k = set_strengths(lat, positions, k.detach().cpu().numpy())
            g = ringpass_grad(coords0.detach().cpu().numpy(), lat_k, beam, int(nturns))
            ctx_.save_for_backward(torch.from_numpy(g.dk[np.asarray(positions)].copy()), torch.from_numpy(g.dx0))
            return torch.tensor(g.loss, dtype=torch.float64)

        @staticmethod
        def backward(ctx_, gout):
            dk, dx0 = ctx_.saved_tensors
            return gout * dk, gout * dx0, None, None, None, None

    return TrackLoss


def __getattr__(name):  # TrackLoss is built on first use, so the package does not need torch
    global _TRACKLOSS
    if name == "TrackLoss":
        if _TRACKLOSS is None:
            _TRACKLOSS = _make_trackloss()
        return _TRACKLOSS
    raise AttributeError(name)
