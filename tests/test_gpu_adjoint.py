"""Adjoint kernel (tp_adjoint) against the oracle.

Checks, with the tolerance written at each:
  * exact forward-mode derivatives of the oracle (dual numbers through the same reference
    algorithm -- the CPU counterpart of an AD tool) along random directions in
    (strengths, initial coordinates): relative 1e-9;
  * central finite differences with h = 1e-7 evaluated in quad precision (the oracle run
    with __float128): relative 1e-7.  In double the reference's own rounding noise
    divided by h limits such differences to ~1e-5..1e-4 (measured; long double to
    ~1e-5 on small gradients), so the differences are taken in quad precision;
  * bit-reproducibility, checkpoint-stride independence, survivors-only loss, the torch
    autograd wrapper and the refusal of unsupported elements.
"""

import copy

import numpy as np
import pytest

from oracle import pyoracle as O
from trackpad_hip import (Beam, Drift, ExactSBend, Lattice, Octupole, Quadrupole, Sextupole, TrackPadHIPError, ringpass_grad,
                          set_strengths, strength_positions)
from trackpad_hip import workloads as W

import lattices as LT

pytestmark = pytest.mark.gpu


def _dba(beam):
    return W.dba_ring(beam=beam)


def _directional(g, dk, dr0):
    return float(np.dot(g.dk, dk) + np.sum(g.dx0 * dr0))


@pytest.mark.parametrize("fast", [False, True])
def test_adjoint_vs_forward_mode(ctx, fast):
    beam = Beam(W.ENERGY)
    lat = _dba(beam)
    n, T = 64, 5
    c0 = W.dba_particles(n, seed=21)
    g = ringpass_grad(c0, lat, beam, T, fast=fast, ctx=ctx)
    assert g.n_alive == n
    rng = np.random.default_rng(5)
    pos = strength_positions(lat)
    for trial in range(3):
        dk = np.zeros(len(lat))
        dk[pos] = rng.normal(0, 1, len(pos))
        dr0 = rng.normal(0, 1, (n, 6)) * np.array([5e-5, 2e-5, 1e-5, 2e-6, 1e-4, 1e-3])
        loss, dl, na = O.loss_dir(lat, beam, c0, T, dk=dk, dr0=dr0)
        assert na == n and abs(loss / na - g.loss) <= 1e-12 * abs(g.loss) * (1e3 if fast else 1)
        ref = dl / na
        got = _directional(g, dk, dr0)
        assert abs(got - ref) <= 1e-9 * abs(ref), (trial, got, ref)
    # per-component check of the strength gradient (the unit directions of 5 positions)
    for p in pos[:: max(1, len(pos) // 5)]:
        dk = np.zeros(len(lat))
        dk[p] = 1.0
        _, dl, na = O.loss_dir(lat, beam, c0, T, dk=dk)
        assert abs(g.dk[p] - dl / na) <= 1e-9 * abs(dl / na) + 1e-22, (p, g.dk[p], dl / na)
    # positions without a strength carry no gradient
    mask = np.ones(len(lat), bool)
    mask[pos] = False
    assert np.all(g.dk[mask] == 0.0)


def test_adjoint_vs_quad_precision_fd(ctx):
    """100 turns of the DBA ring (the differentiation workload's turn count), 8 particles:
    central differences with h = 1e-7, evaluated in quad precision, agree to 1e-7 with
    the adjoint for the largest quad and sextupole gradients and for initial coordinates."""
    import mpmath as mp

    beam = Beam(W.ENERGY)
    lat = _dba(beam)
    n, T, h = 8, 100, 1e-7
    c0 = W.dba_particles(n, seed=22)
    g = ringpass_grad(c0, lat, beam, T, ctx=ctx)

    def loss_q(lat_, c_):
        (hi, lo), na = O.loss_quad(c_, lat_, beam, T, nthreads=8)
        assert na == n
        return (mp.mpf(hi) + mp.mpf(lo)) / na

    pos = strength_positions(lat)
    quads = sorted((p for p in pos if isinstance(lat[p], Quadrupole)), key=lambda p: -abs(g.dk[p]))
    sexts = sorted((p for p in pos if isinstance(lat[p], Sextupole)), key=lambda p: -abs(g.dk[p]))
    for p in (quads[0], sexts[0]):
        k0 = lat[p].k1 if isinstance(lat[p], Quadrupole) else lat[p].k2
        fd = float((loss_q(set_strengths(lat, [p], [k0 + h]), c0) - loss_q(set_strengths(lat, [p], [k0 - h]), c0))
                   / (2 * h))
        assert abs(g.dk[p] - fd) <= 1e-7 * abs(fd), (p, g.dk[p], fd)
    for (i, j, scale) in ((0, 0, 1e-3), (5, 3, 1e-3)):
        hh = h * scale  # the same relative step for a coordinate of size ~1e-4
        cp, cm = c0.copy(), c0.copy()
        cp[i, j] += hh
        cm[i, j] -= hh
        fd = float((loss_q(lat, cp) - loss_q(lat, cm)) / (2 * hh))
        assert abs(g.dx0[i, j] - fd) <= 1e-7 * abs(fd), (i, j, g.dx0[i, j], fd)


def test_adjoint_deterministic_and_stride_independent(ctx):
    beam = Beam(W.ENERGY)
    lat = _dba(beam)
    c0 = W.dba_particles(600, seed=23)
    a = ringpass_grad(c0, lat, beam, 12, ctx=ctx)
    b = ringpass_grad(c0, lat, beam, 12, ctx=ctx)
    assert np.array_equal(a.dk, b.dk) and np.array_equal(a.dx0, b.dx0) and a.loss == b.loss
    for stride in (5, 12):
        c = ringpass_grad(c0, lat, beam, 12, ckpt_stride=stride, ctx=ctx)
        assert c.loss == a.loss
        np.testing.assert_allclose(c.dk, a.dk, rtol=1e-9, atol=1e-9 * np.max(np.abs(a.dk)))
        np.testing.assert_allclose(c.dx0, a.dx0, rtol=1e-8, atol=1e-9 * np.max(np.abs(a.dx0)))


@pytest.mark.parametrize("fast", [False, True])
def test_adjoint_zero_strength_magnets(ctx, fast):
    """A magnet at zero strength is the identity map, but its strength gradient is not zero
    (dpx'/dk1 = -K x): quadrupoles, a sextupole and an octupole switched off, beside a
    sextupole with only a quadrupole error (its b2 gradient through the general path), each
    against the oracle's forward-mode derivative, 1e-9 relative."""
    beam = Beam(W.ENERGY)
    sx_err = Sextupole(0.2, 0.0, num_int_steps=6, polynom_b=[0.0, 0.3, 0.0, 0.0])
    lat = Lattice([Quadrupole(0.4, 1.2, num_int_steps=6), Drift(1.0), Sextupole(0.2, 0.0, num_int_steps=6),
                   Drift(1.0), Quadrupole(0.4, -1.2, num_int_steps=6), Drift(1.0),
                   Quadrupole(0.3, 0.0, num_int_steps=6), Octupole(0.2, 0.0, num_int_steps=4), sx_err,
                   Drift(1.0), Sextupole(0.2, 5.0, num_int_steps=6), Drift(0.5)])
    n, T = 64, 10
    c0 = W.dba_particles(n, seed=27)
    g = ringpass_grad(c0, lat, beam, T, fast=fast, ctx=ctx)
    assert g.n_alive == n
    for p in strength_positions(lat):
        dk = np.zeros(len(lat))
        dk[p] = 1.0
        _, dl, na = O.loss_dir(lat, beam, c0, T, dk=dk)
        ref = dl / na
        assert ref != 0.0
        assert abs(g.dk[p] - ref) <= 1e-9 * abs(ref), (p, g.dk[p], ref)


def test_adjoint_survivors_only(ctx):
    """Lost particles leave the loss and get zero gradient; the loss is the survivors' mean."""
    beam = Beam(W.ENERGY)
    lat = _dba(beam)
    c0 = W.dba_particles(100, seed=24)
    c0[:5, 1] = 0.9  # lost in the first drift (|px| large)
    g = ringpass_grad(c0, lat, beam, 4, ctx=ctx)
    assert g.n_alive == 95 and np.all(g.lost_flags[:5] == 1)
    alive = g.lost_flags == 0
    assert g.loss == pytest.approx(np.mean(g.final[alive, 0] ** 2 + g.final[alive, 2] ** 2), rel=1e-14)
    assert np.all(g.dx0[:5] == 0.0)
    oc, olf, _, _ = O.ringpass(c0, lat, beam, np.zeros(100, np.int64), 4)
    assert np.array_equal(olf, g.lost_flags)


def test_trackloss_autograd(ctx):
    torch = pytest.importorskip("torch")
    from trackpad_hip import TrackLoss

    beam = Beam(W.ENERGY)
    lat = _dba(beam)
    pos = strength_positions(lat)[:10]
    k = torch.tensor([lat[int(p)].k1 if isinstance(lat[int(p)], Quadrupole) else lat[int(p)].k2 for p in pos],
                     dtype=torch.float64, requires_grad=True)
    c0 = torch.tensor(W.dba_particles(50, seed=25), requires_grad=True)
    loss = TrackLoss.apply(k, c0, lat, beam, 3, pos)
    loss.backward()
    g = ringpass_grad(c0.detach().numpy(), lat, beam, 3, ctx=ctx)
    assert float(loss.detach()) == g.loss
    np.testing.assert_array_equal(k.grad.numpy(), g.dk[pos])
    np.testing.assert_array_equal(c0.grad.numpy(), g.dx0)


def test_adjoint_refuses_unsupported(ctx):
    beam = Beam(W.ENERGY)
    # a singular r1 matrix cannot be undone
    lat = Lattice([Drift(1.0), Quadrupole(0.4, 1.0, r1=np.diag([1.0, 1.0, 1.0, 1.0, 1.0, 0.0]))])
    with pytest.raises(TrackPadHIPError, match="TP_EUNSUPPORTED"):
        ringpass_grad(W.dba_particles(8), lat, beam, 2, ctx=ctx)
    # an exact bend in its drift limit (|angle / L| < 1e-6)
    lat = Lattice([Drift(1.0), ExactSBend(1.0, 1e-8), Quadrupole(0.4, 1.0)])
    with pytest.raises(TrackPadHIPError, match="TP_EUNSUPPORTED"):
        ringpass_grad(W.dba_particles(8), lat, beam, 2, ctx=ctx)


@pytest.mark.parametrize("fast", [False, True])
def test_adjoint_forest_exact_bends(ctx, fast):
    """Reverse mode through the reference's ExactSBend (yrot, bend_fringe, bend_edge undone by
    their inverses, VJPs by dual numbers at the reconstructed inputs): gradients against the
    oracle's forward-mode derivatives along random directions and per strength, relative
    1e-9, both builds, 5 turns of 64 particles."""
    beam = Beam(W.ENERGY)
    lat = LT.forest_ring(W.ENERGY)
    n, T = 64, 5
    c0 = W.gaussian_particles(n, 31, 2e-4, 2e-5, 2e-4, 2e-5, 1e-3)
    g = ringpass_grad(c0, lat, beam, T, fast=fast, ctx=ctx)
    assert g.n_alive == n
    rng = np.random.default_rng(8)
    pos = strength_positions(lat)
    assert len(pos) == 12
    for trial in range(3):
        dk = np.zeros(len(lat))
        dk[pos] = rng.normal(0, 1, len(pos))
        dr0 = rng.normal(0, 1, (n, 6)) * np.array([5e-5, 2e-5, 1e-5, 2e-6, 1e-4, 1e-3])
        loss, dl, na = O.loss_dir(lat, beam, c0, T, dk=dk, dr0=dr0)
        assert na == n
        ref = dl / na
        got = _directional(g, dk, dr0)
        assert abs(got - ref) <= 1e-9 * abs(ref), (trial, got, ref)
    for p in pos[::3]:
        dk = np.zeros(len(lat))
        dk[p] = 1.0
        _, dl, na = O.loss_dir(lat, beam, c0, T, dk=dk)
        assert abs(g.dk[p] - dl / na) <= 1e-9 * abs(dl / na) + 1e-22, (p, g.dk[p], dl / na)
    # every initial coordinate of two particles, against unit directions
    for q in (0, 37):
        for j in range(6):
            dr0 = np.zeros((n, 6))
            dr0[q, j] = 1.0
            _, dl, na = O.loss_dir(lat, beam, c0, T, dr0=dr0)
            assert abs(g.dx0[q, j] - dl / na) <= 1e-9 * abs(dl / na) + 1e-20, (q, j, g.dx0[q, j], dl / na)


def test_adjoint_diff_workload(ctx):
    """The differentiation workload at full size: DBA ring, 4,096 particles, 100 turns,
    200 strengths + 4,096 x 6 coordinates.  Against double-precision central differences
    of the device forward with h = 1e-5 (noise-limited to ~1e-5, measured on the oracle):
    relative 2e-4.  Reports the reverse/forward time ratio."""
    beam = Beam(W.ENERGY)
    lat = _dba(beam)
    n, T = 4096, 100
    c0 = W.dba_particles(n, seed=26)
    g = ringpass_grad(c0, lat, beam, T, timed=True, ctx=ctx)
    pos = strength_positions(lat)
    assert len(pos) == 200 and g.n_alive == n
    from trackpad_hip import HIPParticles, compile_table, track_

    def loss_gpu(lat_):
        tab = ctx.table(compile_table(lat_, beam))
        p = HIPParticles.from_coords(c0, ctx=ctx)
        track_(p, tab, nturns=T)
        c, lf, _, _ = p.download()
        a = lf == 0
        return float(np.sum(c[a, 0] ** 2 + c[a, 2] ** 2) / a.sum())

    h = 1e-5
    for p in (pos[0], pos[50], pos[120], pos[199]):
        k0 = lat[p].k1 if isinstance(lat[p], Quadrupole) else lat[p].k2
        fd = (loss_gpu(set_strengths(lat, [p], [k0 + h])) - loss_gpu(set_strengths(lat, [p], [k0 - h]))) / (2 * h)
        assert abs(g.dk[p] - fd) <= 2e-4 * abs(fd), (p, g.dk[p], fd)
    # the whole gradient at full size along one random direction of (strengths, initial
    # coordinates), against the oracle's exact forward-mode derivative: relative 1e-9
    rng = np.random.default_rng(8)
    dk = np.zeros(len(lat))
    dk[pos] = rng.normal(0, 1, len(pos))
    dr0 = rng.normal(0, 1, (n, 6)) * np.array([5e-5, 2e-5, 1e-5, 2e-6, 1e-4, 1e-3])
    _, dl, na = O.loss_dir(lat, beam, c0, T, dk=dk, dr0=dr0)
    assert na == n
    ref = dl / na
    got = _directional(g, dk, dr0)
    assert abs(got - ref) <= 1e-9 * abs(ref), (got, ref)
    print(f"adjoint diff workload: forward {g.forward_ms:.1f} ms, reverse {g.backward_ms:.1f} ms, "
          f"ratio {g.backward_ms / max(g.forward_ms, 1e-9):.2f}")
