"""Pin the CPU oracle before trusting it (CPU only).

1. 50-digit golden vectors (tests/golden/mpmath_golden.json, written by
   tests/golden/make_golden_mpmath.py from independent formulations: complex power sums,
   circle-arc geometry, straight lines) on the reference tests' 10 x 6 input.  Tolerances
   are absolute, per coordinate, and follow the conditioning of the reference formulas
   themselves: z += NormL (beti + delta) - L beti cancels O(L) terms (ulp(L) ~ 1e-16 per
   drift), and Forest's exact-body x = (pzn - pzmx cs + px sn - 1)/irho cancels O(1) terms
   (ulp(1)/irho ~ 1e-15 per body).  A dropped term or a wrong sign moves results by orders
   of magnitude more than these bounds.
2. Closed forms: Yoshida coefficients, Beam, 4th-order convergence of the Yoshida thick
   quad, Horner kicks vs complex evaluation up to order 10, exact dipole body n=0 vs n=10,
   drift composition.
3. Invariants: det of the one-turn Jacobian = 1, J^T S J = S for the reference's exact bend
   with Forest's yrot / fringe / wedge maps and for a ring of them, the fringe's updates as
   the derivatives of its kick, reversibility of the symmetric Yoshida step, on-axis
   particles stay on axis, k = 0 quad == drift.
4. Loss-boundary semantics of check_lost (src/tracking.jl:49-52) and the aperture.
"""

import json
import math
import os

import mpmath as mp
import numpy as np
import pytest

from conftest import PARTICLES_INITIAL
from oracle import pyoracle as O
from trackpad_hip import (Beam, Drift, ExactSBend, ExactSBendLinearEdge, Lattice, Octupole, Quadrupole,
                          RFCavity, SBend, Sextupole, ThinMultipole)
from trackpad_hip.table import DRIFT1, DRIFT2, KICK1, KICK2
from trackpad_hip import workloads as W

HERE = os.path.dirname(os.path.abspath(__file__))
mp.mp.dps = 50
GOLD = json.load(open(os.path.join(HERE, "golden", "mpmath_golden.json")))

# absolute tolerances per coordinate (x, px, y, py, z, delta)
TOL_PLAIN = np.array([4e-18, 4e-18, 4e-18, 4e-18, 4e-16, 4e-19])
TOL_EXACT_BEND = np.array([4e-14, 4e-16, 4e-18, 4e-18, 4e-15, 4e-19])


def _elem(c):
    p, k = c["params"], c["kind"]
    if k == "drift":
        return Drift(p["L"])
    if k == "quad":
        return Quadrupole(p["L"], p["k1"], num_int_steps=p["n"])
    if k == "sext":
        return Sextupole(p["L"], p["k2"], num_int_steps=p["n"])
    if k == "oct":
        return Octupole(p["L"], p["k3"], num_int_steps=p["n"])
    if k == "thinmp":
        return ThinMultipole(0.0, p["a"], p["b"], max_order=p["m"])
    if k == "sbend":
        return SBend(p["L"], p["angle"], p["e1"], p["e2"], num_int_steps=p["n"], fint1=p.get("fint1", 0.0),
                     fint2=p.get("fint2", 0.0), gap=p.get("gap", 0.0), fringe_bend_entrance=p.get("fbe", 1),
                     fringe_bend_exit=p.get("fbx", 1))
    if k == "esbend_body":
        return ExactSBendLinearEdge(p["L"], p["angle"], 0.0, 0.0, num_int_steps=p["n"], fringe_bend_entrance=0,
                                    fringe_bend_exit=0)
    if k == "esbend_le":
        return ExactSBendLinearEdge(p["L"], p["angle"], p["e1"], p["e2"], num_int_steps=p["n"], fint1=p["fint1"],
                                    fint2=p["fint2"], gap=p["gap"])
    if k == "forest_esbend":
        return ExactSBend(p["L"], p["angle"], p["e1"], p["e2"], num_int_steps=p["n"], gk=p["gk"])
    if k == "rf":
        return RFCavity(p["L"], p["volt"], p["freq"], p["lag"], philag=p["philag"], energy=p["energy"])
    if k == "wedge":
        return ExactSBend(p["L"], p["angle"], p["e1"], p["e2"], num_int_steps=p["n"], fringe_bend_entrance=0,
                          fringe_bend_exit=0)
    raise ValueError(k)


@pytest.mark.parametrize("case", [c["name"] for c in GOLD["cases"]])
def test_oracle_vs_mpmath(case):
    c = next(c for c in GOLD["cases"] if c["name"] == case)
    beam = Beam(GOLD["energy"])
    e = _elem(c)
    ref = np.array([[float(v) for v in row] for row in c["out"]])
    out = np.array([O.pass_elem(e, r, beam) for r in GOLD["particles"]])
    tol = TOL_EXACT_BEND if c["kind"] in ("esbend_body", "esbend_le", "forest_esbend", "wedge") else TOL_PLAIN
    err = np.max(np.abs(out - ref), axis=0)
    assert np.all(err <= tol), (err, tol)
    # and the map is not the identity (the fixture exercises it)
    assert np.max(np.abs(ref - np.array(GOLD["particles"]))) > 1e-12


def test_yoshida_literals_vs_closed_form():
    """src/tracking.jl:29-32 against Yoshida's closed form: within 8 ulp, and the
    consistency conditions 2(d1 + d2) = 1, 2 k1 + k2 = 1 hold exactly in double."""
    c = mp.cbrt(2)
    closed = [1 / (2 * (2 - c)), (1 - c) / (2 * (2 - c)), 1 / (2 - c), -c / (2 - c)]
    for lit, v in zip((DRIFT1, DRIFT2, KICK1, KICK2), closed):
        assert abs(mp.mpf(lit) - v) <= 8 * np.spacing(abs(lit))
    assert 2 * (DRIFT1 + DRIFT2) == 1.0 and 2 * KICK1 + KICK2 == 1.0
    assert GOLD["yoshida_closed_form"] == [float(v) for v in closed]


def test_beam_beti():
    """Beam(E): gamma = (E + m)/m with E kinetic (src/lattice.jl:50-54), beti = 1/beta."""
    beam = Beam(GOLD["energy"])
    assert O.beti(beam) == GOLD["beti"] == 1.0 / beam.beta
    assert abs(mp.mpf(GOLD["beti"]) - mp.mpf(GOLD["beti_mp"])) < 4e-16


def test_yoshida_fourth_order():
    """A thick quad converges as h^4: halving the step divides the error by ~16."""
    beam = Beam(3e9)
    r0 = np.array([1e-3, -2e-4, 5e-4, 1e-4, 0.0, 1e-3])
    ref = O.pass_elem(Quadrupole(1.0, 2.0, num_int_steps=512), r0, beam)
    e1 = np.max(np.abs(O.pass_elem(Quadrupole(1.0, 2.0, num_int_steps=2), r0, beam)[:4] - ref[:4]))
    e2 = np.max(np.abs(O.pass_elem(Quadrupole(1.0, 2.0, num_int_steps=4), r0, beam)[:4] - ref[:4]))
    assert 12.0 < e1 / e2 < 20.0, e1 / e2


def test_thick_quad_linear_limit():
    """Small amplitude, delta = 0: the quad converges to the cos/cosh matrix of sqrt(k)L."""
    beam = Beam(1e12)  # beti -> 1 (ultra-relativistic)
    k, L = 1.5, 0.7
    r0 = np.array([1e-7, 2e-8, -1e-7, 3e-8, 0.0, 0.0])
    out = O.pass_elem(Quadrupole(L, k, num_int_steps=200), r0, beam)
    s = math.sqrt(k)
    Mx = np.array([[math.cos(s * L), math.sin(s * L) / s], [-s * math.sin(s * L), math.cos(s * L)]])
    My = np.array([[math.cosh(s * L), math.sinh(s * L) / s], [s * math.sinh(s * L), math.cosh(s * L)]])
    np.testing.assert_allclose(out[:2], Mx @ r0[:2], rtol=1e-9, atol=1e-20)
    np.testing.assert_allclose(out[2:4], My @ r0[2:4], rtol=1e-9, atol=1e-20)


def test_horner_vs_complex_order10():
    """strthinkick's Horner recurrence extended to order 10 equals
    Re/Im sum (b_n + i a_n)(x + i y)^n (numpy complex128)."""
    rng = np.random.default_rng(3)
    beam = Beam(3e9)
    for m in range(0, 11):
        a = rng.normal(0, 1, 11) * 10.0 ** rng.uniform(-2, 4, 11)
        b = rng.normal(0, 1, 11) * 10.0 ** rng.uniform(-2, 4, 11)
        e = ThinMultipole(0.0, a, b, max_order=m)
        for _ in range(5):
            r = np.array([rng.normal(0, 0.01), rng.normal(0, 1e-3), rng.normal(0, 0.01), rng.normal(0, 1e-3), 0.0, 0.0])
            w = complex(r[0], r[2])
            f = sum(complex(b[n], a[n]) * w ** n for n in range(m + 1))
            out = O.pass_elem(e, r, beam)
            scale = sum(abs(complex(b[n], a[n])) * abs(w) ** n for n in range(m + 1))
            assert abs(out[1] - (r[1] - f.real)) <= 1e-14 * scale + 1e-300
            assert abs(out[3] - (r[3] + f.imag)) <= 1e-14 * scale + 1e-300


def test_exact_dipole_body_split_invariance():
    """A pure dipole exact body is exact for any sub-step split: n = 0 and n = 10 agree."""
    beam = Beam(3e9)
    for p in PARTICLES_INITIAL:
        a = O.pass_elem(ExactSBendLinearEdge(1.5, 0.157, num_int_steps=0, fringe_bend_entrance=0, fringe_bend_exit=0), p, beam)
        b = O.pass_elem(ExactSBendLinearEdge(1.5, 0.157, num_int_steps=10, fringe_bend_entrance=0, fringe_bend_exit=0), p, beam)
        assert np.all(np.abs(a - b) <= np.array([1e-13, 1e-15, 1e-17, 1e-17, 1e-14, 0]))


def test_drift_composition_and_zero_quad():
    """Drifts compose (L = L1 + L2); a k = 0 quad (kicks are zero) is a drift."""
    beam = Beam(3e9)
    for p in PARTICLES_INITIAL:
        d = O.pass_elem(Drift(1.3), p, beam)
        dd = O.pass_elem(Drift(0.6), O.pass_elem(Drift(0.7), p, beam), beam)
        assert np.all(np.abs(d - dd) <= np.array([1e-18, 0, 1e-18, 0, 4e-16, 0]))
        q = O.pass_elem(Quadrupole(1.3, 0.0, num_int_steps=7), p, beam)
        assert np.all(np.abs(d - q) <= np.array([1e-18, 0, 1e-18, 0, 4e-15, 0]))


def test_reversibility():
    """The symmetric Yoshida step is time-reversible: flip px, py, track, flip back."""
    beam = Beam(1e12)  # beti = 1 to roundoff, so the z update is symmetric too
    lat = Lattice([Quadrupole(0.4, 1.7, num_int_steps=6), Drift(0.3), Sextupole(0.3, 40.0, num_int_steps=5),
                   Drift(0.2)])
    rev = Lattice(list(reversed(lat.elements)))
    for p in PARTICLES_INITIAL:
        c, lf, _, _ = O.ringpass(np.array([p]), lat, beam, np.zeros(1, np.int64), 1)
        q = c[0].copy()
        q[1], q[3] = -q[1], -q[3]
        back, _, _, _ = O.ringpass(np.array([q]), rev, beam, np.zeros(1, np.int64), 1)
        b = back[0]
        b[1], b[3] = -b[1], -b[3]
        assert np.all(np.abs(b[:4] - np.array(p[:4])) <= 1e-17)


@pytest.mark.parametrize("name", ["fodo", "dba", "forest_bend", "sbend"])
def test_symplectic_det(name):
    """det of the one-turn Jacobian (forward-mode duals, not finite differences) is 1."""
    beam = Beam(3e9)
    if name == "fodo":
        lat = W.fodo_lattice()
    elif name == "dba":
        lat = W.dba_ring(beam=beam)
    elif name == "forest_bend":
        lat = Lattice([Drift(0.5), ExactSBend(1.0, 0.2, 0.05, 0.03, gk=0.02), Quadrupole(0.4, 1.0)])
    else:
        lat = Lattice([SBend(0.9, 0.15, 0.03, 0.02, fint1=0.5, fint2=0.4, gap=0.05), Quadrupole(0.4, -1.0)])
    r0 = np.array([1e-4, -2e-5, 5e-5, 1e-5, 0.0, 1e-4])
    _, J, lost = O.jacobian(lat, beam, r0, 1)
    assert not lost
    assert abs(np.linalg.det(J) - 1.0) < 1e-12, np.linalg.det(J) - 1.0


# (x, px), (y, py) and (delta, z) are the canonical pairs of the reference's coordinates:
# its drift advances z by NormL (beti + delta) - L beti, i.e. z is conjugate to -delta.
_S6 = np.zeros((6, 6))
for _i, (_a, _b) in enumerate(((0, 1), (2, 3), (5, 4))):
    _S6[_a, _b], _S6[_b, _a] = 1.0, -1.0


@pytest.mark.parametrize("gk,e1,e2,r0", [
    (0.0, 0.05, 0.03, [1e-3, 2e-3, -5e-4, 1e-3, 0.0, 1e-3]),
    (0.02, 0.05, 0.03, [1e-3, 2e-3, -5e-4, 1e-3, 0.0, 1e-3]),
    (0.1, -0.04, 0.07, [-2e-3, -4e-3, 3e-3, -2e-3, 1e-4, -5e-3]),
    (0.05, 0.1, 0.0, [5e-3, 1e-2, -4e-3, 6e-3, 0.0, 2e-2]),
])
def test_forest_exact_bend_symplectic(gk, e1, e2, r0):
    """The oracle's ExactSBend -- yrot, Forest fringe, wedge, exact body with Yoshida kicks,
    mirrored exit -- is a symplectic map: J^T S J = S for its dual-number Jacobian, to 1e-13.
    This pins yrot, bend_fringe and bend_edge beyond det J = 1 (15 independent conditions
    instead of one): e.g. the fringe's x and z updates must carry exactly the derivatives of
    its py kick with respect to px and delta (checked at 50 digits below)."""
    beam = Beam(3e9)
    lat = Lattice([ExactSBend(1.0, 0.2, e1, e2, gk=gk, polynom_b=[0.0, 0.1, 2.0, 0.0], max_order=2)])
    _, J, lost = O.jacobian(lat, beam, np.array(r0), 1)
    assert not lost
    d = J.T @ _S6 @ J - _S6
    assert np.max(np.abs(d)) <= 1e-13 * np.max(np.abs(J)) ** 2, np.max(np.abs(d))
    # the map is not the identity's neighbour trivially: the fringe and wedges act
    assert np.max(np.abs(J - np.eye(6))) > 0.1


def test_forest_ring_one_turn_map_symplectic():
    """The one-turn map of a ring of Forest bends, quads, a sextupole and RF: J^T S J = S."""
    import lattices as LT
    beam = Beam(3e9)
    lat = LT.forest_ring(3e9)
    for r0 in ([1e-3, 1e-4, -1e-3, 2e-4, 0.0, 1e-3], [0.0] * 6):
        _, J, lost = O.jacobian(lat, beam, np.array(r0), 3)
        assert not lost
        d = J.T @ _S6 @ J - _S6
        assert np.max(np.abs(d)) <= 1e-12 * np.max(np.abs(J)) ** 2, np.max(np.abs(d))


def test_fringe_updates_are_derivatives_of_its_kick():
    """Forest's fringe (src/tracking.jl:752-815) moves py by -phi yf and x, z by dpx yf^2/2 and
    -dd yf^2/2 with y -> yf = 2y / (1 + sqrt(1 - 2 dpy y)).  The map is symplectic exactly when
    dpx, dpy, dd are the partial derivatives of phi(px, py, dp1) with respect to px, py and
    dp1 = beti + delta; the reference's long closed forms for them are checked against
    50-digit derivatives of its phi, as restated in the golden-vector writer (phi, dpx, dpy,
    dd are read back from the map's outputs).  A single wrong coefficient in any of the four
    closed forms breaks the identity by many orders of magnitude."""
    import importlib.util
    spec = importlib.util.spec_from_file_location(
        "make_golden_mpmath", os.path.join(os.path.dirname(__file__), "golden", "make_golden_mpmath.py"))
    G = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(G)
    mp.mp.dps = 50
    bt = mp.mpf(1) + mp.mpf("1.4e-8")
    y0 = mp.mpf("2e-3")

    def coeffs(px, py, d, b0, gk):
        x1, _, yf, py1, z1, _ = G.bend_fringe([mp.mpf(0), px, y0, py, mp.mpf(0), d], b0, gk, bt)
        phi = (py - py1) / yf
        return phi, 2 * x1 / yf ** 2, 2 * (yf - y0) / yf ** 2, -2 * z1 / yf ** 2

    for px, py, d, b0, gk in ((mp.mpf("3e-3"), mp.mpf("-2e-3"), mp.mpf("1e-3"), mp.mpf("0.1"), mp.mpf("0.02")),
                              (mp.mpf("-1e-2"), mp.mpf("5e-3"), mp.mpf("-2e-2"), mp.mpf("-0.3"), mp.mpf("0.1")),
                              (mp.mpf("2e-3"), mp.mpf("1e-3"), mp.mpf("0"), mp.mpf("0.05"), mp.mpf("0"))):
        phi, dpx, dpy, dd = coeffs(px, py, d, b0, gk)
        f = lambda a, b, c: coeffs(a, b, c, b0, gk)[0]
        gx = mp.diff(lambda a: f(a, py, d), px)
        gy = mp.diff(lambda b: f(px, b, d), py)
        gd = mp.diff(lambda c: f(px, py, c), d)
        for got, ref in ((dpx, gx), (dpy, gy), (dd, gd)):
            assert abs(got - ref) <= mp.mpf("1e-30") * (abs(ref) + mp.mpf("1e-30")), (got, ref)


def test_on_axis_stays_on_axis():
    beam = Beam(3e9)
    lat = Lattice([Quadrupole(0.5, 1.2), Drift(1.0), Sextupole(0.2, 30.0), Octupole(0.2, 100.0), Drift(0.5)])
    c, lf, _, _ = O.ringpass(np.zeros((1, 6)), lat, beam, np.zeros(1, np.int64), 50)
    assert np.all(c[0, :4] == 0.0) and lf[0] == 0


def test_loss_boundary_exact():
    """check_lost uses strict '>' against 1.0 (src/tracking.jl:49-52); the aperture uses
    x^2 + y^2 > R^2 at element exit.  Marker keeps coordinates unchanged."""
    beam = Beam(3e9)
    from trackpad_hip import Marker
    lat = Lattice([Marker()])
    one_up = np.nextafter(1.0, 2.0)
    c = np.zeros((6, 6))
    c[0, 0] = 1.0
    c[1, 0] = one_up
    c[2, 2] = -one_up
    c[3, 1] = 1.0
    c[4, 3] = -one_up
    c[5, 4] = 5.0  # z is never checked
    _, lf, lt, le = O.ringpass(c, lat, beam, np.zeros(6, np.int64), 1)
    assert list(lf) == [0, 1, 1, 0, 1, 0]
    R = 0.02
    c = np.zeros((3, 6))
    c[0, 0] = R
    c[1, 0] = np.nextafter(R, 1.0)
    c[2, 0], c[2, 2] = 0.015, 0.014
    _, lf, _, _ = O.ringpass(c, lat, beam, np.zeros(3, np.int64), 1, aperture=R)
    assert list(lf) == [0, 1, 1]


def test_oracle_threads_identical():
    """The multi-threaded oracle (CPU baseline) gives the same bits as one thread."""
    beam = Beam(3e9)
    lat = W.dba_ring(beam=beam)
    c0 = W.dba_particles(64, seed=12)
    a = O.ringpass(c0, lat, beam, np.zeros(64, np.int64), 3, nthreads=1)
    b = O.ringpass(c0, lat, beam, np.zeros(64, np.int64), 3, nthreads=4)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def _plane_err(a, ref):
    rms = np.sqrt(np.mean(ref * ref, axis=0, keepdims=True))
    scale = np.maximum(np.abs(ref), rms)
    return np.max(np.abs(a - ref) / np.where(scale > 0, scale, 1.0), axis=0)


@pytest.mark.parametrize("nturns", [1, 20])
def test_long_double_yardstick(nturns):
    """Implementations that round differently from the reference are measured against the
    same algorithm run in long double (``extended=True``).  That yardstick's own per-plane
    error against the __float128 run is below 1 % of the double run's on the DBA ring
    (measured: 3e-14 against 1e-10 in x after one turn), so it resolves differences of the
    double run's size."""
    beam = Beam(W.ENERGY)
    lat = W.dba_ring(beam=beam)
    n = 32
    c0 = W.dba_particles(n, seed=3)
    lf = np.zeros(n, np.int64)
    oc = O.ringpass(c0, lat, beam, lf, nturns, nthreads=8)[0]
    ld = O.ringpass(c0, lat, beam, lf, nturns, nthreads=8, extended=True)[0]
    q = O.ringpass(c0, lat, beam, lf, nturns, nthreads=8, quad=True)[0]
    eo, el = _plane_err(oc, q), _plane_err(ld, q)
    assert np.all(el <= 0.01 * eo), (el, eo)
    # the double run's error is what the tolerance tests of the fast build say it is
    assert eo[0] > 1e-12 and eo[3] < 1e-11, eo


def test_adjoint_fixture_T2000_vs_oracle_dual():
    """The T = 2,000 derivative fixture (quad-precision central differences at h and h/2,
    Richardson-extrapolated; tests/golden/make_adjoint_golden.py) against the oracle's
    forward-mode derivative in double along the same directions: two independent
    computations of the same number agree to 1e-8 relative (measured 7.5e-10 and 3.3e-9).
    The unextrapolated differences are ~1e-5..1e-6 away: at this turn count the h^2 term of a
    central difference is not small."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "adjoint_dba_T2000.json")
    with open(path) as f:
        fx = json.load(f)
    beam = Beam(W.ENERGY)
    rings, _, _ = W.variant_rings(nvar=2, seed=fx["variant_seed"])
    for lat, rec in zip(rings, fx["variants"]):
        c = W.dba_particles(fx["particles"], seed=rec["particle_seed"])
        dk = np.zeros(len(lat))
        dk[rec["positions"]] = rec["dk"]
        _, orc, na = O.loss_dir(lat, beam, c, fx["turns"], dk=dk, dr0=np.array(rec["dx0"]))
        exact = float(rec["dS"])
        assert na == fx["particles"]
        assert abs(orc - exact) <= 1e-8 * abs(exact), (orc, exact)
        assert abs(float(rec["dS_fd_h"]) - exact) > 1e-7 * abs(exact)
