"""Pins the oracle's Wiggler (src/tracking.jl:1224-1525) without retyping its formulas:

* no field (Bmax = 0): the step plan (Nsteps x round(L / lw) Yoshida steps of lw / Nsteps)
  is the linearised drift of the element's length;
* symplecticity: each 2nd-order map shifts the momenta by the gradient of a potential
  (px, py) -> (px - d/dx F, py - a_y) with a_y = d/dy F, so the ratios of the ax / axpy and
  ay / aypx terms are fixed by calculus; a wrong ratio, sign or index breaks J^T S J = S;
* fourth-order convergence in the step (the Yoshida coefficients);
* the textbook planar wiggler: the path lengthening L K^2 / (4 (beta gamma)^2) of an
  on-axis electron, K = e B0 lw / (2 pi m c) (the K^2 / 2 term of the undulator resonance
  condition), and the radiated energy (2/3) r_e gamma^3 times the sampled integral of
  B^2 / (B rho)^2 with B rho = p / e.
"""

import math

import numpy as np
import pytest

from oracle import pyoracle as O
from trackpad_hip import Beam, Drift, Lattice, Wiggler

from conftest import PARTICLES_INITIAL

P = np.array(PARTICLES_INITIAL, dtype=np.float64)
BEAM = Beam(3.5e9)
M_E = 0.51099895069e6  # eV
C = 2.99792458e8
R_E = 2.8179403205e-15  # classical electron radius, m


def line(els, c=P, extended=False, exact=True):
    oc, lf, _, _ = O.ringpass(c, Lattice(list(els)), BEAM, np.zeros(len(c), np.int64), 1, exact=exact,
                              extended=extended)
    assert not lf.any()
    return oc


def test_no_field_is_the_linearised_drift():
    w = line([Wiggler(1.2, lw=0.2, Bmax=0.0, Nsteps=8)])
    d = line([Drift(1.2)], exact=False)
    assert np.max(np.abs(w - d)) <= 2e-17
    # round(L / lw) periods of at least one: L = 0.05 still tracks one period of 0.2
    w = line([Wiggler(0.05, lw=0.2, Bmax=0.0, Nsteps=4)])
    assert np.max(np.abs(w - line([Drift(0.2)], exact=False))) <= 2e-17


HELICAL = dict(By=[1, 1, 0, 1, 1, 0, 2, 1, 3, 5, 4, 1], Bx=[1, 1, 5, 3, 4, 0])


@pytest.mark.parametrize("w", [Wiggler(1.2, lw=0.2, Bmax=0.8, Nsteps=8),
                               Wiggler(0.6, lw=0.2, Bmax=1.5, Nsteps=8, energy=3e9, **HELICAL)],
                         ids=["planar", "three harmonics"])
def test_symplectic(w):
    # canonical pairs (x, px), (y, py), (delta, z) in the reference's sign of z
    S = np.zeros((6, 6))
    for i, (a, b) in enumerate(((0, 1), (2, 3), (5, 4))):
        S[a, b], S[b, a] = 1.0, -1.0
    for r0 in P[:4]:
        ro, J, lost = O.jacobian(Lattice([w]), BEAM, r0)
        assert not lost
        assert np.max(np.abs(J.T @ S @ J - S)) <= 4e-15
        assert np.max(np.abs(ro - r0)) > 1e-7  # the field acts


def test_fourth_order_in_the_step():
    r = np.array([[3e-4, 1e-5, 2e-4, -1e-5, 0.0, 1e-3]])
    mk = lambda n: [Wiggler(1.2, lw=0.2, Bmax=1.5, Nsteps=n, energy=1e9)]
    ref = line(mk(256), r, extended=True)
    err = [np.max(np.abs(line(mk(n), r) - ref)) for n in (4, 8, 16)]
    for a, b in zip(err, err[1:]):
        assert 13.0 <= a / b <= 19.0, err


@pytest.mark.parametrize("n", [8, 64])
def test_planar_path_lengthening(n):
    E, lw, B0, L = 1e9, 0.1, 1.0, 1.0
    out = line([Wiggler(L, lw=lw, Bmax=B0, Nsteps=n, energy=E)], np.zeros((1, 6)))[0]
    K = C * B0 * lw / (2 * math.pi * M_E)
    bg = math.sqrt((E / M_E) ** 2 - 1.0)
    assert out[4] == pytest.approx(L * K * K / (4 * bg * bg), rel=1e-9)
    assert abs(out[0]) <= 1e-17 and out[1] == 0.0 and out[2] == 0.0 and out[3] == 0.0


def test_planar_radiation_loss():
    E, lw, B0, L, n = 1e9, 0.1, 1.0, 1.0, 8
    out = line([Wiggler(L, lw=lw, Bmax=B0, Nsteps=n, energy=E, rad=1)], np.zeros((1, 6)))[0]
    gamma = E / M_E
    brho = math.sqrt(E * E - M_E * M_E) / C
    # the kick samples B^2 = B0^2 cos^2(kw s) at s = 0 and after each of the L / lw * n steps
    # of lw / n, each weighted by lw / n
    nstep = n * round(L / lw)
    sampled = sum(math.cos(2 * math.pi * k / n) ** 2 for k in range(nstep + 1)) * lw / n
    assert sampled == pytest.approx(L / 2 + lw / n, rel=1e-12)
    want = -(2.0 / 3.0) * R_E * gamma ** 3 * sampled * B0 * B0 / (brho * brho)
    assert out[5] == pytest.approx(want, rel=1e-5)
    # without radiation the energy is untouched
    assert line([Wiggler(L, lw=lw, Bmax=B0, Nsteps=n, energy=E)], np.zeros((1, 6)))[0][5] == 0.0
