"""Pins the oracle's restatements of the element families beyond the hot lattices
(src/tracking.jl:529-605, 981-1218, 1531-1663) without retyping their formulas:

* independent formulations in 50-digit mpmath: the solenoid as the integrated helix of a
  particle in a uniform longitudinal field with hard-edge fringe kicks, the corrector as
  the integrated trajectory under a uniform kick per unit length, the linear bend's
  transverse part as the solution of its linear ODE, the RLC wake as a solution of the
  damped-oscillator equation;
* group identities: YRotation(a) then YRotation(-a), LorentzBoost then InvLorentzBoost;
* special cases that reduce to maps pinned elsewhere: the SC wrappers against Drift /
  Quadrupole / Sextupole / Octupole / SBend, LBend of zero angle and gradient against the
  linearised drift, Translation along s against a drift of negative length, a multi-slice
  strong beam of zero offsets against the thin beam, the tabulated wake at its samples.
"""

import math

import numpy as np
import pytest

from oracle import pyoracle as O
from trackpad_hip import (AccelCavity, Beam, CrabCavity, Drift, DriftSC, InvLorentzBoost, Lattice, LBend,
                          LongitudinalRFMap, LongitudinalRLCWake, LongitudinalWake, LorentzBoost, Octupole, OctupoleSC,
                          Quadrupole, QuadrupoleSC, RFCavity, SBend, SBendSC, Sextupole, SextupoleSC, Solenoid,
                          SpaceCharge, StrongGaussianBeam, StrongThinGaussianBeam, Translation, YRotation, Corrector)

from conftest import ENERGY_VAL, PARTICLES_INITIAL

mp = pytest.importorskip("mpmath")
P = np.array(PARTICLES_INITIAL, dtype=np.float64)
BEAM = Beam(ENERGY_VAL)


def line(els, c=P, beam=BEAM, exact=True):
    oc, lf, _, _ = O.linepass(c, Lattice(list(els)), beam, np.zeros(len(c), np.int64), exact=exact)
    assert not lf.any()
    return oc


def test_solenoid_against_integrated_helix():
    L, ks = 0.5, 0.8
    out = line([Solenoid(L, ks)])
    mp.mp.dps = 50
    for r0, r1 in zip(P, out):
        x0, px0, y0, py0, z0, d = (mp.mpf(float(v)) for v in r0)
        pn = 1 / (1 + d)
        H = mp.mpf(ks) * pn / 2
        # mechanical slopes inside after the entrance fringe kick; Larmor motion x'' = 2H y',
        # y'' = -2H x' (uniform field, slopes of constant length)
        u0, v0 = px0 * pn + H * y0, py0 * pn - H * x0
        w = 2 * H

        def traj(s):
            c, sn = mp.cos(w * s), mp.sin(w * s)
            u = u0 * c + v0 * sn
            v = -u0 * sn + v0 * c
            x = x0 + (u0 * sn - v0 * (c - 1)) / w
            y = y0 + (v0 * sn + u0 * (c - 1)) / w
            return x, y, u, v

        x, y, u, v = traj(mp.mpf(L))
        # exit fringe kick back to the scaled canonical slopes, then to momenta
        xpr, ypr = u - H * y, v + H * x
        dz = mp.quad(lambda s: (traj(s)[2] ** 2 + traj(s)[3] ** 2) / 2, [0, L])
        ref = [x, xpr / pn, y, ypr / pn, z0 + dz, d]
        for k in range(6):
            assert abs(float(r1[k]) - float(ref[k])) <= 4e-18 + 1e-15 * abs(float(ref[k])), (k, r1[k], ref[k])


def test_corrector_against_integrated_trajectory():
    L, hk, vk = 0.4, 1.5e-4, -2.2e-4
    out = line([Corrector(L, hk, vk)])
    mp.mp.dps = 50
    for r0, r1 in zip(P, out):
        x0, px0, y0, py0, z0, d = (mp.mpf(float(v)) for v in r0)
        pn = 1 / (1 + d)
        # uniform kick per unit length: slope (p + k s / L) pn
        sx = lambda s: (px0 + mp.mpf(hk) * s / L) * pn
        sy = lambda s: (py0 + mp.mpf(vk) * s / L) * pn
        x = x0 + mp.quad(sx, [0, L])
        y = y0 + mp.quad(sy, [0, L])
        dz = mp.quad(lambda s: (sx(s) ** 2 + sy(s) ** 2) / 2, [0, L])
        ref = [x, px0 + hk, y, py0 + vk, z0 + dz, d]
        for k in range(6):
            assert abs(float(r1[k]) - float(ref[k])) <= 4e-18 + 1e-15 * abs(float(ref[k])), (k, r1[k], ref[k])


def test_lbend_transverse_against_linear_ode():
    L, ang, K, bye = 0.7, 0.05, 0.8, 1e-4
    out = line([LBend(L, ang, K=K, by_error=bye)])  # no edges (e1 = e2 = 0, fint = 0)
    mp.mp.dps = 40
    Kx = mp.mpf(ang) / L
    for r0, r1 in zip(P, out):
        x0, px0, y0, py0, z0, d = (mp.mpf(float(v)) for v in r0)
        pn = 1 / (1 + d)
        G1, G2 = (Kx * Kx + K) * pn, -mp.mpf(K) * pn
        dterm = d * pn - bye
        # x'' = -G1 x + Kx dterm, y'' = -G2 y (slopes xpr = px pn)
        fx = lambda s, X: [X[1], -G1 * X[0] + Kx * dterm]
        fy = lambda s, Y: [Y[1], -G2 * Y[0]]
        X = mp.odefun(fx, 0, [x0, px0 * pn])(L)
        Y = mp.odefun(fy, 0, [y0, py0 * pn])(L)
        ref = [X[0], X[1] / pn, Y[0], Y[1] / pn]
        for k in range(4):
            assert abs(float(r1[k]) - float(ref[k])) <= 4e-18 + 1e-14 * abs(float(ref[k])), (k, r1[k], ref[k])


def test_rlc_wake_solves_damped_oscillator():
    f, R, Q = 1.0e9, 1.0e6, 1.2
    w0 = 2 * math.pi * f
    scale = 1e-12
    c = 2.99792458e8
    mp.mp.dps = 30
    # the reference's wake, read back from the oracle: delta change / -scale at z = c t
    def W(t):
        r = np.zeros((1, 6))
        r[0, 4] = t * c
        return -line([LongitudinalRLCWake(freq=f, Rshunt=R, Q0=Q, scale=scale)], r)[0, 5] / scale

    assert W(0.0) == pytest.approx(R * w0 / Q, rel=1e-15)
    assert W(1e-12) == W(0.0)  # min(z / c, 0): no wake ahead of the source
    # W'' - (w0/Q) W' + w0^2 W = 0 for t < 0 (roots w0/(2Q) +- i w0'), by central differences
    for t in (-0.3e-9, -1.1e-9, -2.6e-9):
        h = 2e-13
        w, wp, wm = W(t), W(t + h), W(t - h)
        d1 = (wp - wm) / (2 * h)
        d2 = (wp - 2 * w + wm) / (h * h)
        res = d2 - (w0 / Q) * d1 + w0 * w0 * w
        assert abs(res) <= 1e-5 * w0 * w0 * max(abs(w), 1e-3 * R * w0 / Q), (t, res)


def test_group_identities():
    a = 0.02
    back = line([YRotation(0.0, angle=a), YRotation(0.0, angle=-a)])
    assert np.max(np.abs(back - P)) <= 2e-17 and not np.array_equal(line([YRotation(0.0, angle=a)]), P)
    back = line([LorentzBoost(0.03), InvLorentzBoost(0.03)])
    assert np.max(np.abs(back - P)) <= 2e-19
    assert np.array_equal(line([LorentzBoost(0.03, mode=1)]), P) and np.array_equal(line([SpaceCharge(0.9)]), P)


def test_special_cases_reduce_to_pinned_maps():
    # SC wrappers: the same maps as the plain magnets (bit for bit)
    assert np.array_equal(line([DriftSC(0.5, a=0.01, b=0.02)]), line([Drift(0.5)]))
    assert np.array_equal(line([QuadrupoleSC(0.4, k1=1.3)]), line([Quadrupole(0.4, 1.3)]))
    assert np.array_equal(line([SextupoleSC(0.3, k2=2.4)]), line([Sextupole(0.3, 2.4)]))
    assert np.array_equal(line([OctupoleSC(0.2, k3=3.6)]), line([Octupole(0.2, 3.6)]))
    assert np.array_equal(line([SBendSC(0.9, 0.15, 0.03, 0.02)]), line([SBend(0.9, 0.15, 0.03, 0.02)]))
    # LBend without angle and gradient: the linearised drift6 (USE_EXACT_HAMILTONIAN = false)
    lb = line([LBend(0.7, 0.0, K=0.0)])
    dr = line([Drift(0.7)], exact=False)
    assert np.max(np.abs(lb - dr)) <= 4e-19  # one rounding of z (different operation order)
    # Translation along s moves x, y as a drift of negative length does
    tr = line([Translation(0.0, ds=3e-3)])
    dn = line([Drift(-3e-3)])
    assert np.max(np.abs(tr[:, [0, 2]] - dn[:, [0, 2]])) <= 1e-19
    tr = line([Translation(0.0, dx=1e-3, dy=-2e-3)])
    assert np.array_equal(tr[:, 0], P[:, 0] - 1e-3) and np.array_equal(tr[:, 2], P[:, 2] + 2e-3)
    # a strong beam of equal-weight centred slices kicks as the thin beam of the same charge/E
    E = ENERGY_VAL
    sgb = line([StrongGaussianBeam(3e-3, 0.511e6, 1.0, 10 ** 6, E, [1e-3, 1.2e-3], nzslice=4)])
    thin = line([StrongThinGaussianBeam(3e-3 / E, 1e-3, 1.2e-3)])
    assert np.max(np.abs(sgb - thin)) <= 1e-18 and np.max(np.abs(sgb - P)) > 1e-14
    # accelerating cavity: no change on the synchronous phase (z = 0)
    c = P.copy()
    c[:, 4] = 0.0
    acc = line([AccelCavity(0.0, volt=2.5e6, freq=5e8, h=1200.0, phis=0.2, energy=E)], c)
    assert np.array_equal(acc, c)
    # crab cavity (L = 0): no energy change on axis, px kick volt/E sin(phi beti) at z = 0
    c[:, 0] = 0.0
    cr = line([CrabCavity(0.0, volt=2.0e6, freq=5e8, phi=0.1, energy=E)], c)
    assert np.array_equal(cr[:, 5], c[:, 5])
    bt = O.beti(BEAM)
    assert np.allclose(cr[:, 1] - c[:, 1], 2.0e6 / E * math.sin(0.1 * bt), rtol=1e-12)
    # longitudinal RF map: slip -2 pi h eta / k per unit delta, eta = alphac - 1/gamma^2
    rf = RFCavity(0.0, 1e6, 5e8, h=1200.0)
    lr = line([LongitudinalRFMap(2e-4, rf)])
    eta = 2e-4 - (1 - (1 / bt) ** 2)
    k = 2 * math.pi * 5e8 / 2.99792458e8
    assert np.allclose((lr[:, 4] - P[:, 4]) / P[:, 5], -2 * math.pi * 1200.0 * eta / k, rtol=1e-9)


def test_tabulated_wake_interpolates():
    times = np.array([0.0, 1e-10, 2e-10, 4e-10])
    wakes = np.array([5.0, 3.0, -1.0, 0.5])
    scale = 1e-9
    c = 2.99792458e8
    for tf, want in ((1e-10, 3.0), (2e-10, -1.0), (1.5e-10, 1.0), (3e-10, -0.25), (5e-10, 0.5 + 1.5 * 0.5)):
        r = np.zeros((1, 6))
        r[0, 4] = -tf * c  # fliphalf -1: t * fliphalf = tf
        got = -line([LongitudinalWake(times, wakes, scale=scale)], r)[0, 5] / scale
        assert got == pytest.approx(want, rel=1e-12, abs=1e-12), (tf, got, want)
    r = np.zeros((1, 6))
    r[0, 4] = 1e-3  # ahead of the source: t > times[0] * fliphalf
    assert line([LongitudinalWake(times, wakes, scale=scale)], r)[0, 5] == 0.0


def test_crab_cavity_kicks_are_one_potential():
    """A thin crab cavity is a kick from one potential, V = -(volt/E) x sin((k z + phi) beti) /
    beti-scaled, so its px kick and its delta kick must be each other's derivative
    (Panofsky-Wenzel): with the reference's canonical pairs (x, px), (y, py), (delta, z), the
    map is symplectic, J^T S J = S for the oracle's dual-number Jacobian -- which a wrong sign
    or a dropped k beti factor in the delta kick breaks, while the thin kick's det J = 1 holds
    either way.  For L > 0 the reference evaluates the kick's phase from the entry z, before
    the first half-drift (src/tracking.jl:1535-1542); that map does not even keep volume
    (det J - 1 ~ 2e-7 here): the oracle keeps the reference's order, and the test records the
    defect."""
    S = np.zeros((6, 6))
    for a, b in ((0, 1), (2, 3), (5, 4)):
        S[a, b], S[b, a] = 1.0, -1.0
    beam = Beam(3e9)
    r0 = np.array([1e-3, 2e-4, -5e-4, 1e-4, 2e-3, 1e-3])
    for volt, freq, phi in ((2.0e6, 5e8, 0.3), (-1.5e6, 1.3e9, -1.1)):
        lat = Lattice([Drift(0.5), CrabCavity(0.0, volt=volt, freq=freq, phi=phi, energy=3e9), Quadrupole(0.4, 1.0)])
        _, J, lost = O.jacobian(lat, beam, r0, 1)
        assert not lost
        d = J.T @ S @ J - S
        assert np.max(np.abs(d)) <= 1e-13 * np.max(np.abs(J)) ** 2, np.max(np.abs(d))
        # the delta kick is there: J[delta, x] = -k volt/E beti cos(...) is far from zero
        assert abs(J[5, 0]) > 1e-3
    lat = Lattice([CrabCavity(0.3, volt=2.0e6, freq=5e8, phi=0.3, energy=3e9)])
    _, J, lost = O.jacobian(lat, beam, r0, 1)
    assert not lost
    assert abs(np.linalg.det(J) - 1.0) > 1e-9 and np.max(np.abs(J.T @ S @ J - S)) > 1e-9
