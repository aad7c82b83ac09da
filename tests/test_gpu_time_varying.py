"""Time-varying lattices: per-turn sequence rows in the device table.

The reference resolves every TimeVaryingElement at the start of each linepass! call, and
ringpass! advances time by dt_turn and the turn by 1 per call (src/lattice.jl:260-317).
The device must give what that loop gives: checked against the oracle run one turn at a
time on the lattice materialised for that turn.
"""

import math

import numpy as np
import pytest

from oracle import pyoracle as O
from trackpad_hip import (Beam, Drift, HIPBatch, HIPParticles, Lattice, Quadrupole, Sextupole,
                          TrackPadHIPError, compile_table, materialize_lattice, ringpass_, timed,
                          track_)
from trackpad_hip import workloads as W

pytestmark = pytest.mark.gpu


def _oracle_turns(c0, lat, beam, nturns, *, time=0.0, dt_turn=0.0, turn=0):
    """ringpass! as the reference runs it: one linepass! per turn on the lattice resolved at
    that turn's (time, turn)."""
    c = np.array(c0, dtype=np.float64)
    lf = np.zeros(len(c), np.int64)
    lt = np.full(len(c), -1, np.int32)
    le = np.full(len(c), -1, np.int32)
    t = time
    for k in range(nturns):
        c, lf, ltk, lek = O.ringpass(c, materialize_lattice(lat, time=t, turn=turn + k), beam, lf, 1, turn=turn + k)
        new = (ltk >= 0) & (lt == -1)
        lt[new], le[new] = ltk[new], lek[new]
        t += dt_turn
    return c, lf, lt, le


def _ref_quad_line():
    # the reference's own time-dependence test (test/verify_time_dependence.jl:26-70)
    q0 = Quadrupole(0.4, 0.6, name="QTD", num_int_steps=10)
    qtd = timed(q0, k1=lambda t, n: 0.6 + 0.2 * math.sin(t) + 0.01 * n,
                num_int_steps=lambda t, n: 8 if t < 0.5 else 12)
    d = Drift(1.0, name="D")
    return Lattice([d, qtd, d])


def test_reference_time_dependence_case(ctx):
    beam = Beam(3.0e9)
    lat = _ref_quad_line()
    r0 = np.array([[1e-3, 2e-4, 5e-4, -1e-4, 0.0, 0.0]])
    # ringpass over 3 turns from time 0, dt 0.1, turn 4 == manual per-turn linepass
    c = r0.copy()
    lf = np.zeros(1, np.int64)
    ringpass_(c, lat, beam, lf, 3, time=0.0, dt_turn=0.1, turn=4, ctx=ctx)
    m = r0.copy()
    for k in range(3):
        mk = materialize_lattice(lat, time=0.1 * k, turn=4 + k)
        assert mk[1].k1 == 0.6 + 0.2 * math.sin(0.1 * k) + 0.01 * (4 + k) and mk[1].num_int_steps == 8
        ringpass_(m, mk, beam, np.zeros(1, np.int64), 1, turn=4 + k, ctx=ctx)
    assert np.array_equal(c, m)
    oc, olf, _, _ = _oracle_turns(r0, lat, beam, 3, time=0.0, dt_turn=0.1, turn=4)
    assert np.array_equal(c, oc) and lf[0] == olf[0] == 0
    # resolved at (1.0, 5): 12 steps and the turn-5 strength; differs from (0, 0)
    a, b = r0.copy(), r0.copy()
    ringpass_(a, lat, beam, np.zeros(1, np.int64), 1, time=1.0, turn=5, ctx=ctx)
    ringpass_(b, lat, beam, np.zeros(1, np.int64), 1, time=0.0, turn=0, ctx=ctx)
    assert np.linalg.norm(a - b) > 1e-12


def test_dba_ring_with_ramped_sextupoles(ctx):
    """DBA ring whose SF sextupoles ramp with the turn and whose first quadrupole wobbles in
    time, 40 turns from turn 10: device vs the oracle's turn-by-turn loop."""
    beam = Beam(W.ENERGY)
    base = W.dba_ring(beam=beam)
    els = list(base.elements)
    for i, e in enumerate(els):
        if isinstance(e, Sextupole) and e.name == "SF" and i < 50:
            els[i] = timed(e, k2=lambda t, n, k0=e.k2: k0 * (1.0 + 0.002 * n))
    q = els[1]
    els[1] = timed(q, k1=lambda t, n, k0=q.k1: k0 * (1.0 + 1e-3 * math.sin(2 * math.pi * 50.0 * t)))
    lat = Lattice(els)
    c0 = W.dba_particles(512, seed=51)
    c0[:3, 0] = 0.03  # beyond the dynamic aperture: lost on the way
    T, turn, time, dt = 40, 10, 0.0, 1.0e-4
    gc = c0.copy()
    glf = np.zeros(512, np.int64)
    glt = np.empty(512, np.int32)
    gle = np.empty(512, np.int32)
    ringpass_(gc, lat, beam, glf, T, time=time, dt_turn=dt, turn=turn, ctx=ctx, lost_turn=glt, lost_elem=gle)
    oc, olf, olt, ole = _oracle_turns(c0, lat, beam, T, time=time, dt_turn=dt, turn=turn)
    assert np.array_equal(glf, olf) and np.array_equal(glt, olt) and np.array_equal(gle, ole)
    assert olf.sum() >= 1
    alive = olf == 0
    scale = np.maximum(np.abs(oc), np.sqrt(np.mean(oc[alive] ** 2, axis=0)))
    assert np.max((np.abs(gc - oc) / scale)[alive]) <= 1e-12
    # the static ring at the same turns differs
    sc = c0.copy()
    ringpass_(sc, base, beam, np.zeros(512, np.int64), T, ctx=ctx)
    assert np.max(np.abs(sc[alive] - gc[alive])) > 1e-9


def test_rows_bound_the_turn_range(ctx):
    """Tracking, the adjoint and Jacobians refuse turns outside the table's rows."""
    beam = Beam(3.0e9)
    lat = _ref_quad_line()
    tab = ctx.table(compile_table(lat, beam, turn=4, nturns=3, time=0.0, dt_turn=0.1))
    p = HIPParticles.from_coords(np.zeros((8, 6)), ctx=ctx)
    track_(p, tab, nturns=3, turn=4)  # inside the rows
    with pytest.raises(TrackPadHIPError, match="TP_EINVAL"):
        track_(p, tab, nturns=3, turn=5)
    import ctypes as C

    q = HIPParticles.from_coords(np.zeros((8, 6)), ctx=ctx)
    loss, na = C.c_double(), C.c_int64()
    rc = ctx._lib.tp_adjoint(ctx.handle, tab.handle, q.handle, 5, 3, 1, 0, C.byref(loss), C.byref(na), None, None)
    assert rc == -1 and b"time-varying rows" in ctx._lib.tp_last_error()
    rc = ctx._lib.tp_adjoint(ctx.handle, tab.handle, q.handle, 4, 3, 1, 0, C.byref(loss), C.byref(na), None, None)
    assert rc == 0 and na.value == 8


def _ramped_ring():
    """FODO ring (8 cells, 4 Yoshida steps) whose focusing quads ramp with the turn, whose
    first defocusing quad wobbles in time (with a step count that changes at t = 0.3), with
    a sextupole and an octupole ramped in time."""
    from trackpad_hip import Octupole
    base = W.fodo_lattice()
    els = list(base.elements)
    for i, e in enumerate(els):
        if e.name == "QF":
            els[i] = timed(e, k1=lambda t, n, k0=e.k1: k0 * (1.0 + 0.004 * n))
    q = els[2]
    els[2] = timed(q, k1=lambda t, n, k0=q.k1: k0 * (1.0 + 0.01 * math.sin(7.0 * t)),
                   num_int_steps=lambda t, n: 4 if t < 0.3 else 6)
    els[5] = timed(Sextupole(0.2, 5.0, name="S"), k2=lambda t, n: 5.0 + 20.0 * t)
    els[9] = timed(Octupole(0.2, 50.0, name="O"), k3=lambda t, n: 50.0 * math.cos(t))
    return Lattice(els)


@pytest.mark.parametrize("fast", [False, True])
def test_adjoint_time_varying(ctx, fast):
    """Reverse mode over a time-varying table (one row per turn): the gradient equals that of
    the static lattice made of the T materialised turns in a row, tracked for one turn, with
    each position's strength gradient summed over its T copies -- the derivative with
    respect to an offset added to the element's strength at every turn -- and the oracle's
    dual-number derivatives of that unrolled lattice along random directions (relative
    1e-9).  With one checkpoint segment both runs undo the same maps from the same end state,
    so dL/dx0 agrees bit for bit in the parity build; with segments of 4 turns the reverse
    pass restarts from stored states (reconstruction differs in the last bits: 1e-12)."""
    from trackpad_hip import materialize_lattice, ringpass_grad, strength_positions
    beam = Beam(W.ENERGY)
    lat = _ramped_ring()
    n, T, turn, time, dt = 48, 6, 3, 0.2, 0.05
    c0 = W.gaussian_particles(n, 53, 1e-3, 1e-4, 1e-3, 1e-4, 1e-3)
    g = ringpass_grad(c0, lat, beam, T, turn=turn, time=time, dt_turn=dt, ckpt_stride=4, fast=fast, ctx=ctx)
    assert g.n_alive == n
    ns = len(lat)
    els, t = [], time
    for k in range(T):  # time accumulated per turn, as ringpass! and the table writer do
        els += materialize_lattice(lat, time=t, turn=turn + k).elements
        t += dt
    unrolled = Lattice(els)
    u = ringpass_grad(c0, unrolled, beam, 1, fast=fast, ctx=ctx)
    assert u.n_alive == n and abs(g.loss - u.loss) <= 1e-14 * abs(u.loss)
    assert np.max(np.abs(g.dx0 - u.dx0)) <= 1e-12 * np.max(np.abs(u.dx0))
    g1 = ringpass_grad(c0, lat, beam, T, turn=turn, time=time, dt_turn=dt, ckpt_stride=T, fast=fast, ctx=ctx)
    if fast:
        assert np.max(np.abs(g1.dx0 - u.dx0)) <= 1e-12 * np.max(np.abs(u.dx0))
    else:
        assert np.array_equal(g1.dx0, u.dx0)
    summed = u.dk.reshape(T, ns).sum(axis=0)
    pos = strength_positions(lat)
    assert len(pos) == 18 and np.all(g.dk[np.setdiff1d(np.arange(ns), pos)] == 0.0)
    assert np.max(np.abs(g.dk - summed)) <= 1e-12 * np.max(np.abs(summed))
    rng = np.random.default_rng(9)
    for trial in range(2):
        dk = np.zeros(ns)
        dk[pos] = rng.normal(0, 1, len(pos))
        dr0 = rng.normal(0, 1, (n, 6)) * np.array([5e-5, 2e-5, 1e-5, 2e-6, 1e-4, 1e-3])
        _, dl, na = O.loss_dir(unrolled, beam, c0, 1, dk=np.tile(dk, T), dr0=dr0)
        ref = dl / na
        got = float(np.dot(g.dk, dk) + np.sum(g.dx0 * dr0))
        assert abs(got - ref) <= 1e-9 * abs(ref), (trial, got, ref)


def test_batch_with_time_varying_variant(ctx):
    beam = Beam(3.0e9)
    lat = _ref_quad_line()
    static = Lattice([Drift(1.0), Quadrupole(0.4, 0.6), Drift(1.0)])
    c0 = np.random.default_rng(52).normal(0, 1e-4, (100, 6))
    tabs = [ctx.table(compile_table(lat, beam, turn=4, nturns=5, time=0.0, dt_turn=0.1)),
            ctx.table(compile_table(static, beam))]
    parts = [HIPParticles.from_coords(c0, ctx=ctx) for _ in range(2)]
    b = HIPBatch(tabs, parts)
    b.track([5, 5], [4, 0])
    for tab, p, t0 in zip(tabs, parts, (4, 0)):
        r = HIPParticles.from_coords(c0, ctx=ctx)
        track_(r, tab, nturns=5, turn=t0)
        assert np.array_equal(r.download()[0], p.download()[0])
    with pytest.raises(TrackPadHIPError, match="TP_EINVAL"):
        b.track([5, 5], [0, 0])  # variant 0 has rows for turns 4..8 only
    # the batched adjoint against the per-variant calls (parity build: dL/dx0 bit for bit)
    from trackpad_hip import ringpass_grad
    parts = [HIPParticles.from_coords(c0, ctx=ctx) for _ in range(2)]
    b = HIPBatch(tabs, parts)
    gb = b.adjoint([5, 5], [4, 0], ckpt_stride=2)
    for v, (l, t0) in enumerate(((lat, 4), (static, 0))):
        r = ringpass_grad(c0, l, beam, 5, turn=t0, time=0.0, dt_turn=0.1, ckpt_stride=2, ctx=ctx)
        assert np.array_equal(gb.dx0[v], r.dx0)
        assert np.max(np.abs(gb.dk[v] - r.dk)) <= 1e-12 * np.max(np.abs(r.dk))
