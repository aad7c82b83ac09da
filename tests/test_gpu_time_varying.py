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


def test_rows_bound_the_turn_range_and_adjoint_refuses(ctx):
    beam = Beam(3.0e9)
    lat = _ref_quad_line()
    tab = ctx.table(compile_table(lat, beam, turn=4, nturns=3, time=0.0, dt_turn=0.1))
    p = HIPParticles.from_coords(np.zeros((8, 6)), ctx=ctx)
    track_(p, tab, nturns=3, turn=4)  # inside the rows
    with pytest.raises(TrackPadHIPError, match="TP_EINVAL"):
        track_(p, tab, nturns=3, turn=5)
    from trackpad_hip import _lib as L
    import ctypes as C

    q = HIPParticles.from_coords(np.zeros((8, 6)), ctx=ctx)
    loss, na = C.c_double(), C.c_int64()
    rc = ctx._lib.tp_adjoint(ctx.handle, tab.handle, q.handle, 4, 3, 1, 0, C.byref(loss), C.byref(na), None, None)
    assert rc == -2 and b"time-varying" in ctx._lib.tp_last_error()


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
    with pytest.raises(TrackPadHIPError, match="TP_EUNSUPPORTED"):
        b.adjoint([5, 5], [4, 0])
