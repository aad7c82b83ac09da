"""Write tests/golden/tables/*.bin: table blobs of the Python writer for a few lattices.

Run from the repository root:  python tests/golden/make_tables.py

The blobs pin the table format: tests/test_table_writer.py checks that the Python writer
still produces them byte for byte, and the Julia writer's test
(trackpad.jl_amd/julia/TrackPadHIP/test/runtests.jl) compares its output with them.
"""

import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "tests"), os.path.join(ROOT, "trackpad.jl_amd")]

import math  # noqa: E402

import lattices as LT  # noqa: E402
from trackpad_hip import Beam, Lattice, Quadrupole, Sextupole, compile_table, read_madx, timed, write_madx  # noqa: E402
from trackpad_hip import workloads as W  # noqa: E402

# The turn range of the time-varying golden table: rows for turns 5..8, the first at
# time 2e-4 s, each later one 1e-6 s after the previous (as ringpass! advances time).
TV_TURN, TV_NTURNS, TV_TIME, TV_DT = 5, 4, 2e-4, 1e-6


def dba_ring_time_varying():
    """The DBA workload ring with every sextupole ramped by 0.2 % per turn and the first
    QF1 modulated at 50 Hz in time (test/runtests.jl builds the same lattice in Julia)."""
    els = list(W.dba_ring().elements)
    for i, e in enumerate(els):
        if isinstance(e, Sextupole):
            els[i] = timed(e, k2=lambda t, n, k0=e.k2: k0 * (1.0 + 0.002 * n))
    q = els[1]
    assert isinstance(q, Quadrupole)
    els[1] = timed(q, k1=lambda t, n, k0=q.k1: k0 * (1.0 + 1e-3 * math.sin(2 * math.pi * 50.0 * t)))
    return Lattice(els, name="DBA_TV")


MADX = os.path.join(HERE, "madx")


def madx_files():
    """MAD-X text the Julia reader is checked on: the DBA workload ring as the Python writer
    prints it.  (madx/hand_cell.madx, a hand-written cell with variables, deferred expressions,
    inheritance, a sequence with refer = centre and from =, an RBEND given by its chord, a
    multipole, a monitor, a kicker and a cavity given by its harmonic number, is written by
    hand; its table is generated below.)"""
    yield "dba_workload.madx", write_madx(W.dba_ring(), Beam(W.ENERGY))


def cases():
    b35 = Beam(3.5e9)
    yield "verify_fodo_line", compile_table(LT.verify_fodo_line(), b35)
    yield "verify_dba_line", compile_table(LT.verify_dba_line(), b35)
    yield "verify_exact_bend_line", compile_table(LT.verify_exact_bend_line(), b35)
    yield "verify_elements", compile_table(Lattice([e for _, e in LT.element_cases()]), b35)
    yield "fodo_workload", compile_table(W.fodo_lattice(), Beam(W.ENERGY), aperture=0.02)
    yield "dba_workload", compile_table(W.dba_ring(), Beam(W.ENERGY))
    yield "families_ring", compile_table(LT.families_ring(), Beam(3e9))
    yield "dba_time_varying", compile_table(dba_ring_time_varying(), Beam(W.ENERGY), turn=TV_TURN,
                                            nturns=TV_NTURNS, time=TV_TIME, dt_turn=TV_DT)
    yield "madx_hand_cell", compile_table(*read_madx(os.path.join(MADX, "hand_cell.madx"), bend="ExactSBendLinearEdge"))


def main():
    os.makedirs(os.path.join(HERE, "tables"), exist_ok=True)
    os.makedirs(MADX, exist_ok=True)
    for name, text in madx_files():
        with open(os.path.join(MADX, name), "w") as f:
            f.write(text)
        print(name, len(text))
    for name, blob in cases():
        with open(os.path.join(HERE, "tables", name + ".bin"), "wb") as f:
            f.write(blob)
        print(name, len(blob))


if __name__ == "__main__":
    main()
