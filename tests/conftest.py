"""Test configuration: import paths and the ``gpu`` marker.

``-m "not gpu"`` runs the oracle, host-logic and library-load tests on any machine;
``-m gpu`` runs the device parity tests (they call libtrackpad_hip through its C ABI).
"""

import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "trackpad.jl_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs libtrackpad_hip kernels)")


@pytest.fixture(scope="session")
def ctx():
    from trackpad_hip import HIPContext

    return HIPContext(0)


# The reference tests' hand-written 10 x 6 input (test/verify_fodo.jl:7-18, repeated in
# verify_dba.jl, verify_bend.jl, verify_exact_bend.jl, verify_elements.jl) and energy.
PARTICLES_INITIAL = [
    [1.0e-4, 2.0e-4, 3.0e-4, -1.0e-4, 5.0e-5, 2.0e-4],
    [-2.2e-4, 1.7e-4, -1.1e-4, 2.3e-4, -7.0e-5, 1.0e-4],
    [3.5e-4, -2.1e-4, 8.0e-5, -1.8e-4, 1.4e-4, -2.6e-4],
    [-4.0e-4, 2.9e-4, 1.6e-4, 9.0e-5, -1.2e-4, 3.1e-4],
    [5.2e-4, -3.3e-4, -2.4e-4, 1.1e-4, 2.6e-4, -3.7e-4],
    [-6.1e-4, 4.4e-4, 2.7e-4, -2.5e-4, -3.0e-4, 4.5e-4],
    [7.0e-4, -5.2e-4, -3.1e-4, 3.4e-4, 3.3e-4, -5.0e-4],
    [-8.0e-4, 6.0e-4, 3.9e-4, -4.2e-4, -3.8e-4, 5.8e-4],
    [9.1e-4, -6.7e-4, -4.6e-4, 5.1e-4, 4.4e-4, -6.3e-4],
    [-9.8e-4, 7.3e-4, 5.3e-4, -5.9e-4, -4.9e-4, 7.1e-4],
]
ENERGY_VAL = 3.5e9
PARITY_ATOL = 1e-15  # test/verify_elements.jl:6
