"""Jacobians from reverse passes (tp_jacobian) and the optics functions built on them.

The oracle's forward-mode dual numbers give the exact derivative of the same algorithm
on the CPU; the GPU Jacobian must agree with it, have unit determinant, and reproduce the
reference's tune / chromaticity / closed-orbit formulas (src/optics.jl) evaluated on the
oracle's maps.
"""

import numpy as np
import pytest

from oracle import pyoracle as O
from trackpad_hip import (Beam, Lattice, Quadrupole, fastfindm66_refpts, find_closed_orbit_4d,
                          find_closed_orbit_6d, getchrom, gettune, jacobians, one_turn_map, tune_from_map)
from trackpad_hip import workloads as W

pytestmark = pytest.mark.gpu


def _dba():
    beam = Beam(W.ENERGY)
    return beam, W.dba_ring(beam=beam)


@pytest.mark.parametrize("nturns,fast,tol", [(1, False, 1e-11), (5, False, 1e-10), (5, True, 1e-9)])
def test_jacobian_matches_forward_mode_oracle(ctx, nturns, fast, tol):
    beam, lat = _dba()
    c0 = W.dba_particles(6, seed=41)
    c0[:, :4] *= 20  # transverse amplitudes of ~1 mm: nonlinear maps
    J, fin, lf = jacobians(c0, lat, beam, nturns, ckpt_stride=2, fast=fast, ctx=ctx)
    assert not lf.any()
    for q in range(6):
        ro, Jo, lost = O.jacobian(lat, beam, c0[q], nturns)
        assert not lost
        scale = np.max(np.abs(Jo))
        assert np.max(np.abs(J[q] - Jo)) <= tol * scale, (q, np.max(np.abs(J[q] - Jo)) / scale)
        np.testing.assert_allclose(fin[q], ro, rtol=1e-9, atol=1e-15)


def test_jacobian_determinant_is_one(ctx):
    beam, lat = _dba()
    c0 = W.dba_particles(64, seed=42)
    c0[:, :4] *= 10
    for T, tol in ((1, 1e-12), (50, 1e-10)):
        J, _, lf = jacobians(c0, lat, beam, T, ckpt_stride=5, ctx=ctx)
        det = np.linalg.det(J[lf == 0])
        assert np.max(np.abs(det - 1.0)) <= tol, (T, np.max(np.abs(det - 1.0)))


@pytest.mark.parametrize("fast", [False, True])
def test_jacobian_forest_exact_bends(ctx, fast):
    """Jacobians through the reference's ExactSBend (yrot, Forest fringe with gK != 0, wedges):
    against the oracle's dual numbers after 3 turns (1e-10 of the largest entry; 1e-9 fast),
    and unit determinant to 1e-11 for 32 particles."""
    import lattices as LT
    beam = Beam(W.ENERGY)
    lat = LT.forest_ring(W.ENERGY)
    c0 = W.gaussian_particles(32, 44, 1e-3, 1e-4, 1e-3, 1e-4, 1e-3)
    J, fin, lf = jacobians(c0, lat, beam, 3, fast=fast, ctx=ctx)
    assert not lf.any()
    tol = 1e-9 if fast else 1e-10
    for q in (0, 13, 31):
        ro, Jo, lost = O.jacobian(lat, beam, c0[q], 3)
        assert not lost
        scale = np.max(np.abs(Jo))
        assert np.max(np.abs(J[q] - Jo)) <= tol * scale, (q, np.max(np.abs(J[q] - Jo)) / scale)
    det = np.linalg.det(J)
    assert np.max(np.abs(det - 1.0)) <= 1e-11, np.max(np.abs(det - 1.0))


def test_jacobian_of_lost_particle_is_zero(ctx):
    beam, lat = _dba()
    c0 = W.dba_particles(4, seed=43)
    c0[1, 1] = 0.9
    J, _, lf = jacobians(c0, lat, beam, 2, ctx=ctx)
    assert lf.tolist() == [0, 1, 0, 0]
    assert np.all(J[1] == 0.0) and np.any(J[0] != 0.0)


def test_one_turn_map_and_tunes(ctx):
    beam, lat = _dba()
    M = one_turn_map(lat, beam, ctx=ctx)
    _, Mo, _ = O.jacobian(lat, beam, np.zeros(6), 1)
    assert np.max(np.abs(M - Mo)) <= 1e-12 * np.max(np.abs(Mo))
    qx, qy = gettune(lat, beam, ctx=ctx)
    qxo, qyo = tune_from_map(Mo)
    assert abs(qx - qxo) <= 1e-12 and abs(qy - qyo) <= 1e-12
    assert abs(qx - 0.335) < 0.01 and abs(qy - 0.639) < 0.01  # the ring's design tunes 9.335 / 6.639
    # segment maps, each about the same reference (as the reference does), against the
    # oracle's; maps about the orbit propagated segment to segment compose to the one-turn map
    refpts = [100, 250, len(lat)]
    maps = fastfindm66_refpts(lat, 0.0, refpts, E0=beam.energy, ctx=ctx)
    prev, r, P = 0, np.zeros(6), np.eye(6)
    for i, rp in enumerate(refpts):
        seg = Lattice(list(lat.elements[prev:rp]))
        _, Ms, _ = O.jacobian(seg, beam, np.zeros(6), 1)
        assert np.max(np.abs(maps[:, :, i] - Ms)) <= 1e-12 * np.max(np.abs(Ms))
        J, fin, _ = jacobians(r, seg, beam, 1, ctx=ctx)
        P = J[0] @ P
        r, prev = fin[0], rp
    assert np.max(np.abs(P - M)) <= 1e-12 * np.max(np.abs(M))


def test_chromaticity_against_oracle_maps(ctx):
    beam, lat = _dba()
    xi = getchrom(lat, beam, ctx=ctx)
    _, M0, _ = O.jacobian(lat, beam, np.zeros(6), 1)
    r1 = np.zeros(6)
    r1[5] = 1e-8
    _, M1, _ = O.jacobian(lat, beam, r1, 1)
    q0, q1 = tune_from_map(M0), tune_from_map(M1)
    ref = ((q1[0] - q0[0]) / 1e-8, (q1[1] - q0[1]) / 1e-8)
    assert abs(xi[0] - ref[0]) <= 1e-5 and abs(xi[1] - ref[1]) <= 1e-5, (xi, ref)


def test_closed_orbits(ctx):
    beam, lat = _dba()
    # 4-D off-momentum orbit of the DBA ring: a fixed point of the oracle's map
    dp = 1e-3
    x4 = find_closed_orbit_4d(lat, beam, dp=dp, tol=1e-14, maxiter=30, ctx=ctx)
    ro, _, _ = O.jacobian(lat, beam, np.array([x4[0], x4[1], x4[2], x4[3], 0.0, dp]), 1)
    assert np.max(np.abs(ro[:4] - x4)) <= 1e-12 and abs(x4[0]) > 1e-5
    # 6-D orbit of the DBA ring (with its RF cavity) with one displaced quadrupole
    els = list(lat.elements)
    q0 = els[1]
    els[1] = Quadrupole(q0.L, q0.k1, num_int_steps=10, t1=[1e-4, 0, -5e-5, 0, 0, 0], t2=[-1e-4, 0, 5e-5, 0, 0, 0])
    ring = Lattice(els)
    x6 = find_closed_orbit_6d(ring, beam, x0=np.zeros(6), tol=1e-14, maxiter=30, ctx=ctx)
    ro, _, _ = O.jacobian(ring, beam, x6, 1)
    assert np.max(np.abs(ro - x6)) <= 1e-11 and np.max(np.abs(x6[:4])) > 1e-6, (ro - x6, x6)
