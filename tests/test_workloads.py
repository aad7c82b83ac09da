"""The benchmark workloads are what trackpad_hip.workloads documents (CPU only): element counts
and strengths of the DBA and DA rings, and the periodic beta functions DBA_BETA that
dba_particles matches the beam to, recomputed from the oracle's one-turn Jacobian."""

import math

import numpy as np

from oracle import pyoracle as O
from trackpad_hip import Beam, Drift, ExactSBendLinearEdge, Quadrupole, RFCavity, Sextupole, total_length
from trackpad_hip import workloads as W


def test_dba_ring_layout():
    beam = Beam(W.ENERGY)
    lat = W.dba_ring(beam=beam)
    els = lat.elements
    assert len(els) == 501
    cell = els[:25]
    kinds = [type(e) for e in cell]
    assert kinds.count(ExactSBendLinearEdge) == 2 and kinds.count(Quadrupole) == 6
    assert kinds.count(Sextupole) == 4 and kinds.count(Drift) == 13
    k1 = sorted(abs(e.k1) for e in cell if isinstance(e, Quadrupole))
    assert k1[0] == 0.5 and k1[-1] == 2.4
    assert sorted({e.k2 for e in cell if isinstance(e, Sextupole)}) == [-32.0, 27.0]
    b = [e for e in cell if isinstance(e, ExactSBendLinearEdge)]
    assert all(e.L == 1.5 and e.angle == 2 * math.pi / 40 and e.fint1 == 0.5 and e.gap == 0.03 for e in b)
    rf = els[-1]
    assert isinstance(rf, RFCavity) and rf.volt == 1e6 and rf.h == 400
    assert rf.freq == 400 * beam.beta * 2.99792458e8 / total_length(lat)
    da = W.da_ring(beam=beam)
    assert len(da.elements) == 80 * 25 + 1


def test_dba_beta_from_the_one_turn_jacobian():
    """beta = J12 / sin(mu), cos(mu) = (J11 + J22) / 2 of each transverse 2 x 2 block of the
    oracle's one-turn map at the origin, as dba_particles uses it; alpha = (J11 - J22) / (2 sin(mu))
    vanishes at the symmetry point the ring starts at."""
    beam = Beam(W.ENERGY)
    ro, J, lost = O.jacobian(W.dba_ring(beam=beam), beam, np.zeros(6), 1)
    assert not lost
    for plane, beta in ((0, W.DBA_BETA[0]), (2, W.DBA_BETA[1])):
        m = J[plane:plane + 2, plane:plane + 2]
        assert abs(np.linalg.det(m) - 1.0) < 1e-12
        c = 0.5 * (m[0, 0] + m[1, 1])
        assert abs(c) < 1.0  # stable
        s = math.copysign(math.sqrt(1.0 - c * c), m[0, 1])
        assert m[0, 1] / s == np.float64(beta) or abs(m[0, 1] / s - beta) <= 4e-16 * beta
        assert abs((m[0, 0] - m[1, 1]) / (2 * s)) < 1e-6
