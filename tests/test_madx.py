"""MAD-X lattice I/O (trackpad_hip.madx): round trips through the table compiler, a hand-written
sequence file against the lattice built from constructors, and the refusals (CPU only)."""

import math
import os

import numpy as np
import pytest

from oracle import pyoracle as O
from trackpad_hip import (Beam, Corrector, Drift, ExactSBend, ExactSBendLinearEdge, Lattice, M_PROTON, MadxError,
                          Marker, Octupole, Quadrupole, RFCavity, SBend, Sextupole, Solenoid, ThinMultipole,
                          compile_table, read_madx, total_length, write_madx)
from trackpad_hip import workloads as W

import lattices as LT


@pytest.mark.parametrize("which", ["dba", "fodo", "verify_dba", "exact_bend", "mixed"])
def test_round_trip_same_table(which, tmp_path):
    """write_madx then read_madx gives the same elements: the compiled tables are the same
    bytes (every value prints to a string that reads back to the same double, unit changes
    included)."""
    beam = Beam(W.ENERGY)
    if which == "dba":
        lat = W.dba_ring(beam=beam)
    elif which == "fodo":
        lat = W.fodo_lattice()
    elif which == "verify_dba":
        lat = LT.verify_dba_line()
    elif which == "exact_bend":
        lat = Lattice([Drift(1.0), ExactSBend(1.0, 0.1, 0.03, 0.02, gk=0.01), Quadrupole(0.4, -1.2, num_int_steps=7)])
    else:
        lat = Lattice([Drift(0.5, name="D0"), ThinMultipole(0.0, [0.0, 1e-4, 0.0, 3e-3], [0.0, 0.3, -2.0, 17.0, 41.0],
                                                           max_order=4, name="MP"),
                       Octupole(0.2, -120.0, name="O1"), Marker(name="BPM"), Corrector(0.1, 1e-4, -2e-4),
                       Solenoid(0.5, 0.2), SBend(1.0, 0.05, 0.01, 0.0, polynom_b=[0, 0.4, 0, 0], max_order=1),
                       RFCavity(0.2, 2.5e5, 499.654e6, 0.0, h=328, philag=0.3, energy=W.ENERGY)])
    path = tmp_path / "lat.madx"
    text = write_madx(lat, beam, str(path))
    assert path.read_text() == text
    lat2, beam2 = read_madx(str(path), energy=beam.energy)
    assert len(lat2) == len(lat)
    if which != "mixed":
        assert compile_table(lat2, beam) == compile_table(lat, beam)
    else:
        # a thin-multipole coefficient that no knl reproduces after the division by n! reads
        # back one ulp away (3e-3 / 3! here); the tracked coordinates agree to rounding
        c0 = W.dba_particles(32, seed=5)
        a = O.ringpass(c0, lat, beam, np.zeros(32, np.int64), 2)[0]
        b = O.ringpass(c0, lat2, beam, np.zeros(32, np.int64), 2)[0]
        np.testing.assert_allclose(a, b, rtol=1e-14, atol=1e-18)
    # BEAM's total energy in GeV reads back to the kinetic energy within rounding
    _, beam3 = read_madx(text)
    assert beam3.energy == pytest.approx(beam.energy, rel=1e-15)


# a hand-written cell (tests/golden/madx/hand_cell.madx): variables, deferred expressions,
# inheritance, element references, a sequence with refer = centre and from =, and a line
with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "madx", "hand_cell.madx")) as _f:
    SEQ = _f.read()


def test_hand_written_sequence():
    lat, beam = read_madx(SEQ, bend="ExactSBendLinearEdge")
    m = 0.51099895069e6
    assert beam.energy == pytest.approx(math.sqrt(3e9 ** 2 + m * m) - m, rel=1e-15)
    ang = 2 * math.pi / 40
    names = [e.name for e in lat.elements]
    assert names == ["QF", "DRIFT_1", "QD", "DRIFT_2", "B", "SD", "MP", "DRIFT_3", "QF2", "SF", "BPM", "DRIFT_4",
                     "SF2", "DRIFT_5", "RB", "DRIFT_6", "KICK", "DRIFT_7", "RF"]
    e = {x.name: x for x in lat.elements}
    assert isinstance(e["QF"], Quadrupole) and e["QF"].k1 == 2.1 and e["QF"].L == 0.3 and e["QF"].num_int_steps == 10
    assert e["QD"].k1 == -2.4 and e["QD"].L == 0.3
    assert e["QF2"].num_int_steps == 6
    assert isinstance(e["B"], ExactSBendLinearEdge) and e["B"].angle == ang and e["B"].gap == 0.03
    assert e["B"].fint1 == 0.5 and e["B"].fint2 == 0.5
    assert e["SF"].k2 == 27.0 and e["SD"].k2 == -32.0 and e["SF2"].k2 == 20.0
    assert e["RB"].L == pytest.approx(1.5, rel=1e-15) and e["RB"].e1 == pytest.approx(ang / 2, rel=1e-15)
    assert isinstance(e["MP"], ThinMultipole) and e["MP"].polynom_b[2] == 0.1 and e["MP"].polynom_a[1] == 1e-3
    assert isinstance(e["BPM"], Marker) and isinstance(e["KICK"], Corrector) and e["KICK"].xkick == 2e-5
    C = total_length(lat)
    assert C == pytest.approx(10.45, abs=1e-12)
    assert isinstance(e["RF"], RFCavity) and e["RF"].freq == pytest.approx(4 * beam.beta * 2.99792458e8 / C, rel=1e-15)
    assert e["RF"].energy == beam.energy and e["RF"].volt == 1e6
    drifts = [x.L for x in lat.elements if isinstance(x, Drift)]
    np.testing.assert_allclose(drifts, [0.25, 0.4, 0.1, 0.1, 0.2, 0.3, 4.4], atol=1e-12)
    # tracked like the same cell built from constructors (the drift lengths differ by rounding)
    ref = Lattice([Quadrupole(0.3, 2.1), Drift(0.25), Quadrupole(0.3, -2.4), Drift(0.4),
                   ExactSBendLinearEdge(1.5, ang, fint1=0.5, fint2=0.5, gap=0.03),
                   Sextupole(0.2, -32.0), ThinMultipole(0.0, [0, 1e-3, 0, 0], [0, 0, 0.1, 0], max_order=2),
                   Drift(0.1), Quadrupole(0.4, 0.5, num_int_steps=6), Sextupole(0.2, 27.0), Marker(), Drift(0.1),
                   Sextupole(0.2, 20.0), Drift(0.2),
                   ExactSBendLinearEdge(1.5, ang, ang / 2, ang / 2), Drift(0.3), Corrector(0.1, 2e-5, 0.0),
                   Drift(4.4), RFCavity(0.0, 1e6, 4 * beam.beta * 2.99792458e8 / 10.45, 0.0, h=4, energy=beam.energy)])
    c0 = W.dba_particles(64, seed=3)
    a = O.ringpass(c0, lat, beam, np.zeros(64, np.int64), 3)[0]
    b = O.ringpass(c0, ref, beam, np.zeros(64, np.int64), 3)[0]
    np.testing.assert_allclose(a, b, rtol=0, atol=1e-13)


def test_lines_repeat_invert_and_nest():
    src = """
    d: drift, l = 1; q: quadrupole, l = 0.5, k1 = 1; s: sextupole, l = 0.2, k2 = 3;
    half: line = (d, q, s);
    ring: line = (2*half, -half, (d, 2*(q, s)));
    beam, particle = proton, energy = 7000;
    use, period = ring;
    """
    lat, beam = read_madx(src)
    kinds = "".join({"Drift": "d", "Quadrupole": "q", "Sextupole": "s"}[type(e).__name__] for e in lat.elements)
    assert kinds == "dqs" * 2 + "sqd" + "d" + "qs" * 2
    assert beam.mass == M_PROTON and beam.energy == pytest.approx(7000e9 - M_PROTON, rel=1e-15)


@pytest.mark.parametrize("src,msg", [
    ("x: foo, l = 1; r: line = (x); beam, energy = 3; use, period = r;", "undefined element or class"),
    ("q: quadrupole, l = 1, k1 = a; r: line = (q); beam, energy = 3; use, period = r;", "undefined name"),
    ("q: quadrupole, l = 1, k1 = 1, k1s = 0.1; r: line = (q); beam, energy = 3; use, period = r;", "k1s"),
    ("a := b; b := a; q: quadrupole, l = 1, k1 = a; r: line = (q); beam, energy = 3; use, period = r;", "circular"),
    ("q: quadrupole, l = 1; s: sequence, l = 3; q, at = 1; q, at = 1.5; endsequence; beam, energy = 3; use, sequence = s;",
     "overlaps"),
    ("q: quadrupole, l = 1; s: sequence, l = 3; q, at = 1;", "ENDSEQUENCE"),
    ("w: wiggler, l = 1; r: line = (w); beam, energy = 3; use, period = r;", "undefined element or class"),
    ("q: quadrupole, l = 1, k1 = 2 +* 3; r: line = (q); beam, energy = 3; use, period = r;", "expression"),
])
def test_refusals(src, msg):
    with pytest.raises(MadxError, match=msg):
        read_madx(src)


def test_writer_refuses_what_madx_cannot_hold():
    beam = Beam(W.ENERGY)
    with pytest.raises(MadxError, match="multipole errors"):
        write_madx(W.da_ring(ncell=2, beam=beam), beam)
    with pytest.raises(MadxError, match="edge method"):
        write_madx(Lattice([SBend(1.0, 0.1, fringe_bend_entrance=3)]), beam)
    with pytest.raises(MadxError, match="one bend model"):
        write_madx(Lattice([SBend(1.0, 0.1), ExactSBend(1.0, 0.1)]), beam)


def test_golden_madx_files():
    """The MAD-X fixtures the Julia reader is checked on (tests/golden/madx): the writer still
    prints the DBA ring as committed, and both files read back to the committed tables."""
    import sys

    here = os.path.dirname(os.path.abspath(__file__))
    sys.path.insert(0, os.path.join(here, "golden"))
    import make_tables

    g = os.path.join(here, "golden")
    with open(os.path.join(g, "madx", "dba_workload.madx")) as f:
        assert f.read() == dict(make_tables.madx_files())["dba_workload.madx"]
    lat, beam = read_madx(os.path.join(g, "madx", "dba_workload.madx"), energy=W.ENERGY)
    with open(os.path.join(g, "tables", "dba_workload.bin"), "rb") as f:
        assert compile_table(lat, beam) == f.read()
    lat, beam = read_madx(os.path.join(g, "madx", "hand_cell.madx"), bend="ExactSBendLinearEdge")
    with open(os.path.join(g, "tables", "madx_hand_cell.bin"), "rb") as f:
        assert compile_table(lat, beam) == f.read()
