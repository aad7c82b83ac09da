"""Closed forms and independent constructions that pin oracle maps whose other checks only
restate the reference's formula (CPU, no GPU).

* RF cavity (src/tracking.jl:502-518): the synchrotron tune of a ring, taken from the
  eigenvalues of the oracle's one-turn Jacobian (forward-mode dual numbers through the whole
  ring), against the small-oscillation closed form built from the ring's longitudinal
  dispersion R56 and the cavity's voltage slope (V/E)(2 pi f / c) cos(phi_s) / beta^2.  At
  1 MeV (beta = 0.94) a 1/beta factor in place of 1/beta^2 moves the tune by 3 %, a missing
  2 pi by a factor 2.5; the sign of the kick decides stability (trace of the 2 x 2 block
  below or above 2), and R56 itself is pinned by relativistic kinematics, -C/(beta gamma)^2
  for a ring without bends (z = c * delay).
* Linear edges (src/tracking.jl:160-221): the zero-amplitude Jacobian of a rectangular bend
  (e1 = e2 = angle/2) against the textbook hard-edge matrices (TRANSPORT: horizontal edge
  focusing tan(e)/rho cancels the sector's, leaving [[1, rho sin(angle)], [0, 1]]; vertical
  lens -tan(e - psi)/rho with psi = gap fint (1 + sin^2 e) / (rho cos e)), for the three
  edge methods (SOLEIL and THOMX reduce to Brown at delta = 0 and px = 0).
* Forest's rotated faces (yrot, bend_edge, src/tracking.jl:727-745, 822-851): see
  ``wedge_geometry`` in tests/golden/make_golden_mpmath.py and the ``ExactSBendWedge*``
  cases of tests/test_oracle_pinning.py::test_oracle_vs_mpmath.
* Gaussian strong-beam kicks (src/tracking.jl:1610-1641): the kick is the gradient of the
  potential U = A exp(-(dx^2/sx^2 + dy^2/sy^2)/2), taken at 50 digits by mpmath's numerical
  differentiation of U (not of the kick formula).
* Lorentz boost (src/tracking.jl:1587-1604): its map equals the Jacobian at zero amplitude
  of Hirata's boost for a crossing half-angle phi (K. Hirata, H. Moshammer, F. Ruggiero,
  Part. Accel. 40 (1993) 205; the full nonlinear map, written here in mpmath and
  differentiated numerically), which fixes the sign of every coupling term; it is
  conformally symplectic (J^T S J = S / cos(phi)), which ties the x-z and delta-px terms'
  signs together.
"""

import math

import mpmath as mp
import numpy as np
import pytest

from oracle import pyoracle as O
from trackpad_hip import (Beam, Drift, Lattice, LorentzBoost, InvLorentzBoost, Quadrupole, RFCavity, SBend,
                          StrongGaussianBeam, StrongThinGaussianBeam, total_length)
from trackpad_hip import workloads as W

C_LIGHT = 2.99792458e8


def _jac(elems, beam, r0=None):
    _, J, lost = O.jacobian(Lattice(list(elems)), beam, np.zeros(6) if r0 is None else np.asarray(r0, float), 1)
    assert not lost
    return J


def _tunes(J):
    return np.sort(np.abs(np.angle(np.linalg.eigvals(J))) / (2 * math.pi))


# --------------------------------------------------------------------------------------
# RF cavity: synchrotron tune
# --------------------------------------------------------------------------------------
def test_rf_synchrotron_tune_low_energy_ring():
    beam = Beam(1.0e6)  # 1 MeV kinetic electrons: beta = 0.941, gamma = 2.96
    b, g = beam.beta, beam.gamma
    cell = [Drift(2.0), Quadrupole(0.3, 1.0), Drift(2.0), Quadrupole(0.3, -1.0)]
    ring = cell * 4
    C = total_length(Lattice(ring))
    h, V, E = 4.0, 2.0e3, 1.0e6
    f = h * b * C_LIGHT / C
    J0 = _jac(ring, beam)
    # longitudinal dispersion of a ring without bends: a particle of energy deviation delta
    # = dE/(p0 c) is faster by d(beta)/beta = delta / (beta gamma^2), so it arrives earlier by
    # C delta / (beta^2 gamma^2) in units of c * time
    R56 = -C / (b * b * g * g)
    assert J0[4, 5] == pytest.approx(R56, rel=1e-13)
    for philag, stable in ((math.pi, True), (0.0, False)):
        J = _jac(ring + [RFCavity(0.0, V, f, 0.0, h=h, philag=philag, energy=E)], beam)
        # no transverse-longitudinal coupling without bends
        assert np.abs(J[:4, 4:]).max() == 0.0 and np.abs(J[4:, :4]).max() == 0.0
        R65 = -(V / E) * (2 * math.pi * f / C_LIGHT) * math.cos(philag) / (b * b)
        # thin cavity after the ring: trace of [[1, R56], [R65, 1 + R56 R65]]
        tr = J[4, 4] + J[5, 5]
        assert tr == pytest.approx(2.0 + R56 * R65, rel=1e-13, abs=1e-13)
        assert (abs(tr) < 2.0) == stable
        if stable:
            qs = math.acos(tr / 2) / (2 * math.pi)
            # small-oscillation form with the slip factor eta_t = 1/(beta gamma)^2 of the ring:
            # Qs^2 = h eta_t V / (2 pi beta E)  (the cavity's 1/beta^2 times f = h beta c / C)
            eta_t = 1.0 / (b * b * g * g)
            closed = math.sqrt(h * eta_t * V / (2 * math.pi * b * E))
            assert qs == pytest.approx(closed, rel=1e-3)  # (2 pi Qs)^2 / 24 = 3e-4 apart
            # a kick with 1/beta in place of 1/beta^2 would give sqrt(beta) times the tune
            assert abs(closed * math.sqrt(b) - qs) > 1e-2 * qs


def test_rf_synchrotron_tune_dba_ring():
    """The DBA workload ring at 3 GeV: the one-turn Jacobian with the cavity equals the ring's
    (without it) times the thin-kick matrix of the cavity's closed-form slope, so their
    eigen-tunes agree; the synchrotron tune is near the 2 x 2 small-oscillation value
    sqrt(-R56 R65) / (2 pi) (the dispersion at the cavity couples it to x at the 2 % level)."""
    beam = Beam(W.ENERGY)
    lat = W.dba_ring(beam=beam)
    ring, rf = lat.elements[:-1], lat.elements[-1]
    J0 = _jac(ring, beam)
    J = _jac(lat.elements, beam)
    R56 = J0[4, 5]
    R65 = -(rf.volt / rf.energy) * (2 * math.pi * rf.freq / C_LIGHT) * math.cos(rf.philag) / beam.beta ** 2
    K = np.eye(6)
    K[5, 4] = R65
    np.testing.assert_allclose(J, K @ J0, rtol=0, atol=1e-12 * np.abs(J).max())
    t, tc = _tunes(J), _tunes(K @ J0)
    np.testing.assert_allclose(t, tc, rtol=1e-10)
    qs = t[0]
    assert R56 * R65 < 0 and 0.005 < qs < 0.05  # above transition: stable at philag = 0
    assert qs == pytest.approx(math.sqrt(-R56 * R65) / (2 * math.pi), rel=0.05)


# --------------------------------------------------------------------------------------
# Linear edges: rectangular bend against the textbook matrices
# --------------------------------------------------------------------------------------
@pytest.mark.parametrize("method", [1, 2, 3])
@pytest.mark.parametrize("fint,gap", [(0.0, 0.0), (0.5, 0.04)])
def test_rectangular_bend_textbook(method, fint, gap):
    beam = Beam(W.ENERGY)
    L, th = 1.2, 0.2
    rho = L / th
    e = th / 2
    bend = SBend(L, th, e, e, num_int_steps=60, fint1=fint, fint2=fint, gap=gap, fringe_bend_entrance=method,
                 fringe_bend_exit=method)
    J = _jac([bend], beam)
    # body: Yoshida-4 of x'' = -x/rho^2, error O((th/n)^4); the vertical body is a drift
    np.testing.assert_allclose(J[:2, :2], [[1.0, rho * math.sin(th)], [0.0, 1.0]], rtol=0, atol=1e-10)
    psi = gap * fint * (1 + math.sin(e) ** 2) / (rho * math.cos(e))
    Ev = np.array([[1.0, 0.0], [-math.tan(e - psi) / rho, 1.0]])
    Dv = np.array([[1.0, L], [0.0, 1.0]])
    np.testing.assert_allclose(J[2:4, 2:4], Ev @ Dv @ Ev, rtol=0, atol=1e-13)
    # without the edges the horizontal block is the sector bend's own focusing
    sector = SBend(L, th, e, e, num_int_steps=60, fringe_bend_entrance=0, fringe_bend_exit=0)
    Js = _jac([sector], beam)
    np.testing.assert_allclose(Js[:2, :2], [[math.cos(th), rho * math.sin(th)], [-math.sin(th) / rho, math.cos(th)]],
                               rtol=0, atol=1e-10)
    np.testing.assert_allclose(Js[2:4, 2:4], Dv, rtol=0, atol=1e-14)  # 240 drift pieces


# --------------------------------------------------------------------------------------
# Gaussian strong beam: kick = gradient of the potential
# --------------------------------------------------------------------------------------
def _grad_potential(A, sx, sy, x0, y0, x, y):
    mp.mp.dps = 50
    U = lambda X, Y: A * mp.exp(-((X - x0) ** 2 / mp.mpf(sx) ** 2 + (Y - y0) ** 2 / mp.mpf(sy) ** 2) / 2)
    return (float(mp.diff(lambda X: U(X, mp.mpf(y)), mp.mpf(x))),
            float(mp.diff(lambda Y: U(mp.mpf(x), Y), mp.mpf(y))))


def test_thin_gaussian_beam_kick_is_gradient():
    beam = Beam(W.ENERGY)
    A, sx, sy, x0, y0 = 3.0e-7, 2.0e-4, 5.0e-5, 1.0e-5, -2.0e-5
    e = StrongThinGaussianBeam(A, sx, sy, xoffset=x0, yoffset=y0)
    rng = np.random.default_rng(4)
    for _ in range(12):
        r = np.array([rng.normal(0, 3e-4), rng.normal(0, 1e-5), rng.normal(0, 1e-4), rng.normal(0, 1e-5), 1e-4, 1e-3])
        out = O.pass_elem(e, r, beam)
        gx, gy = _grad_potential(A, sx, sy, x0, y0, r[0], r[2])
        assert out[1] - r[1] == pytest.approx(gx, rel=1e-9, abs=1e-22)
        assert out[3] - r[3] == pytest.approx(gy, rel=1e-9, abs=1e-22)
        assert out[0] == r[0] and out[2] == r[2] and out[4] == r[4] and out[5] == r[5]


def test_sliced_gaussian_beam_kick_is_gradient():
    beam = Beam(W.ENERGY)
    q, E0, sx, sy = 2.0e3, 3.0e9, 1.5e-4, 4.0e-5
    w, xo, yo = [0.2, 0.5, 0.3], [1e-5, 0.0, -2e-5], [0.0, 3e-6, 1e-5]
    e = StrongGaussianBeam(q, 0.51099895069e6, 1.0, 10 ** 10, E0, [sx, sy], nzslice=3, zslice_npar=w,
                           xoffsets=xo, yoffsets=yo)
    r = np.array([2e-4, 1e-5, -5e-5, 2e-6, 0.0, 0.0])
    out = O.pass_elem(e, r, beam)
    gx = gy = 0.0
    for wi, x0, y0 in zip(w, xo, yo):
        a, b = _grad_potential(q / E0 * wi, sx, sy, x0, y0, r[0], r[2])
        gx, gy = gx + a, gy + b
    assert out[1] - r[1] == pytest.approx(gx, rel=1e-9)
    assert out[3] - r[3] == pytest.approx(gy, rel=1e-9)


# --------------------------------------------------------------------------------------
# Lorentz boost: Hirata's boost at zero amplitude
# --------------------------------------------------------------------------------------
def _hirata_boost(phi, r):
    """Hirata, Moshammer, Ruggiero (1993), eq. of the boost to the head-on frame: h =
    1 + delta - sqrt((1 + delta)^2 - px^2 - py^2), then px* = (px - h tan phi)/cos phi,
    py* = py/cos phi, delta* = delta - px tan phi + h tan^2 phi, and with
    p_s* = sqrt((1 + delta*)^2 - px*^2 - py*^2), hx* = px*/p_s*, hy* = py*/p_s*,
    h_sigma* = 1 - (1 + delta*)/p_s*: x* = tan phi z + (1 + hx* sin phi) x,
    y* = y + hy* sin phi x, z* = z / cos phi + h_sigma* sin phi x."""
    x, px, y, py, z, d = r
    t, c, s = mp.tan(phi), mp.cos(phi), mp.sin(phi)
    h = 1 + d - mp.sqrt((1 + d) ** 2 - px ** 2 - py ** 2)
    pxs = (px - h * t) / c
    pys = py / c
    ds = d - px * t + h * t ** 2
    ps = mp.sqrt((1 + ds) ** 2 - pxs ** 2 - pys ** 2)
    hx, hy, hs = pxs / ps, pys / ps, 1 - (1 + ds) / ps
    return [t * z + (1 + hx * s) * x, pxs, y + hy * s * x, pys, z / c + hs * s * x, ds]


def test_lorentz_boost_is_hirata_linearised():
    mp.mp.dps = 50
    beam = Beam(W.ENERGY)
    phi = 0.0125
    e = LorentzBoost(phi)
    Jh = np.zeros((6, 6))
    for j in range(6):
        def f(eps, j=j):
            r = [mp.mpf(0)] * 6
            r[j] = eps
            return _hirata_boost(mp.mpf(phi), r)
        col = [mp.diff(lambda s, i=i: f(s)[i], mp.mpf(0)) for i in range(6)]
        Jh[:, j] = [float(v) for v in col]
    J = _jac([e], beam, np.zeros(6))
    np.testing.assert_allclose(J, Jh, rtol=0, atol=1e-15)
    assert J[0, 4] > 0 and J[5, 1] < 0  # x gains +tan(phi) z, delta loses tan(phi) px
    S = np.zeros((6, 6))
    for k in range(3):
        S[2 * k, 2 * k + 1], S[2 * k + 1, 2 * k] = 1.0, -1.0
    np.testing.assert_allclose(J.T @ S @ J, S / math.cos(phi), rtol=0, atol=1e-15)
    # the inverse boost undoes it
    Ji = _jac([InvLorentzBoost(phi)], beam, np.zeros(6))
    np.testing.assert_allclose(Ji @ J, np.eye(6), rtol=0, atol=1e-15)
