"""The oracle's SOLEIL (method 2) and THOMX (method 3) linear edges (src/tracking.jl:160-221)
pinned to Brown's (method 1) through the relations the reference's formulas state, measured
on the oracle's own outputs rather than retyped:

* SOLEIL is Brown's vertical focusing divided by (1 + delta): the py kicks of the two
  methods on one particle are in that ratio;
* THOMX is Brown's with the edge angle shifted by the particle's slope px / (1 + delta): plus
  at the entrance, minus at the exit -- the same py as Brown's at the shifted angle.

Brown's edge itself is pinned by the textbook edge matrices (test_oracle_pinning.py).  An
SBend without multipoles leaves py unchanged in its body, so py after the element minus py
before is exactly the entrance edge's kick (exit method 0) or the exit edge's (entrance
method 0), with y at that edge.  No GPU."""

import numpy as np

from oracle import pyoracle as O
from trackpad_hip import Beam, Lattice, SBend

BEAM = Beam(3e9)
L, ANGLE, E1, E2, FINT, GAP = 0.6, 0.12, 0.05, -0.03, 0.5, 0.04

# particles with y, py, px and delta all nonzero
R0 = np.array([[1e-3, 2e-3, 4e-3, -1e-3, 0.0, 1e-2],
               [-2e-3, -5e-3, -3e-3, 2e-3, 1e-4, -2e-2],
               [5e-4, 1e-2, 2e-3, 5e-4, 0.0, 5e-3],
               [-1e-3, -8e-3, -6e-3, -2e-3, -1e-4, 3e-2]])


def run(r0, **kw):
    lat = Lattice([SBend(L, ANGLE, **kw)])
    out, lf, _, _ = O.linepass(np.ascontiguousarray(r0, dtype=np.float64), lat, BEAM,
                               np.zeros(len(r0), np.int64))
    assert not lf.any()
    return out


def test_soleil_is_brown_over_one_plus_delta():
    for fi in (FINT, 0.0):
        kw = dict(e1=E1, e2=E2, fint1=fi, gap=GAP, fringe_bend_exit=0)
        brown = run(R0, fringe_bend_entrance=1, **kw)
        soleil = run(R0, fringe_bend_entrance=2, **kw)
        kb = brown[:, 3] - R0[:, 3]
        ks = soleil[:, 3] - R0[:, 3]
        assert np.all(np.abs(kb) > 1e-7)
        np.testing.assert_allclose(ks, kb / (1.0 + R0[:, 5]), rtol=1e-14, atol=0)
        # and the horizontal edge is Brown's: px after the element differs only through the
        # body's dependence on py (second order), not by the edge's first-order x fx
        assert np.max(np.abs(soleil[:, 1] - brown[:, 1])) < 1e-8


def test_thomx_is_brown_at_the_slope_shifted_angle():
    # entrance: THOMX at e1 equals Brown at e1 + px / (1 + delta) (no fringe integral, so the
    # fringe correction is zero in both)
    thomx = run(R0, e1=E1, e2=E2, fringe_bend_entrance=3, fringe_bend_exit=0)
    for i, r in enumerate(R0):
        e1s = E1 + r[1] / (1.0 + r[5])
        brown = run(r[None, :], e1=e1s, e2=E2, fringe_bend_entrance=1, fringe_bend_exit=0)
        assert thomx[i, 3] == brown[0, 3], i
        assert abs(thomx[i, 3] - r[3]) > 1e-7
    # exit: the slope enters with the other sign; px and y at the exit edge from the same
    # element without its exit edge
    before = run(R0, e1=E1, e2=E2, fringe_bend_entrance=0, fringe_bend_exit=0)
    thomx = run(R0, e1=E1, e2=E2, fringe_bend_entrance=0, fringe_bend_exit=3)
    for i, r in enumerate(R0):
        e2s = E2 - before[i, 1] / (1.0 + r[5])
        brown = run(r[None, :], e1=E1, e2=e2s, fringe_bend_entrance=0, fringe_bend_exit=1)
        assert thomx[i, 3] == brown[0, 3], i
        # the opposite sign (the entrance's) gives another kick
        e2w = E2 + before[i, 1] / (1.0 + r[5])
        wrong = run(r[None, :], e1=E1, e2=e2w, fringe_bend_entrance=0, fringe_bend_exit=1)
        assert abs(wrong[0, 3] - thomx[i, 3]) > 1e-9
