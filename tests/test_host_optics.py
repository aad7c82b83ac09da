"""Host-side logic of the optics functions (no GPU): the reference's tune formula on maps
with known phase advances, the dp / orbit convention and argument validation."""

import math

import numpy as np
import pytest

from trackpad_hip import Drift, Lattice, fastfindm66_refpts, tune_from_map
from trackpad_hip.optics import _reference6


def _rot(mu, beta=3.0, alpha=0.5):
    """Twiss-parametrised 2 x 2 one-turn block with phase advance mu."""
    g = (1 + alpha * alpha) / beta
    c, s = math.cos(mu), math.sin(mu)
    return np.array([[c + alpha * s, beta * s], [-g * s, c - alpha * s]])


@pytest.mark.parametrize("qx,qy", [(0.335, 0.639), (0.1, 0.9), (0.5 - 1e-6, 0.25)])
def test_tune_from_map_recovers_phase_advance(qx, qy):
    M = np.eye(6)
    M[:2, :2] = _rot(2 * math.pi * qx)
    M[2:4, 2:4] = _rot(2 * math.pi * qy, beta=7.0, alpha=-1.2)
    gx, gy = tune_from_map(M)
    assert abs(gx - qx) < 1e-12 and abs(gy - qy) < 1e-12


def test_tune_from_map_zero_offdiagonal():
    # sign(0) = 0 in the reference: a block with M12 = 0 gives sin(mu) = 0
    M = np.eye(6)
    M[0, 0] = M[1, 1] = -1.0
    assert tune_from_map(M)[0] == pytest.approx(0.5)


def test_reference6_dp_convention():
    # src/optics.jl:74-84: dp = 0 with a nonzero orbit delta takes the orbit's delta
    assert _reference6(0.0, [1, 2, 3, 4, 5, 6e-3])[5] == 6e-3
    assert _reference6(1e-3, [1, 2, 3, 4, 5, 6e-3])[5] == 1e-3
    with pytest.raises(ValueError):
        _reference6(0.0, [0, 0, 0])


def test_refpts_validation():
    lat = Lattice([Drift(1.0), Drift(2.0)])
    with pytest.raises(ValueError, match="refpts entries"):
        fastfindm66_refpts(lat, 0.0, [0, 2])
    with pytest.raises(ValueError, match="non-decreasing"):
        fastfindm66_refpts(lat, 0.0, [2, 1])
