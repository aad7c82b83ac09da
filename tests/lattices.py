"""Lattices of the reference's own tests (test/verify_*.jl) and element cases of
test/verify_elements.jl:127-147, built with the host mirror."""

import math

from trackpad_hip import (Drift, ERBend, ExactSBend, Lattice, Marker, Octupole, Quadrupole, RBend,
                          RFCavity, SBend, Sextupole, ThinMultipole)

ENERGY_VAL = 3.5e9


def verify_fodo_line():
    """test/verify_fodo.jl:30-39"""
    return Lattice([Drift(1.0, name="D1"), Quadrupole(1.0, -0.9, name="Q1", num_int_steps=10),
                    Drift(1.0, name="D2"), SBend(0.6, math.pi / 15.0, name="B1", num_int_steps=10),
                    Drift(1.0, name="D3"), Quadrupole(1.0, 0.3, name="Q2", num_int_steps=10),
                    Drift(1.0, name="D4")])


def verify_dba_line():
    """test/verify_dba.jl:31-44"""
    D_L = Drift(2.0, name="D_L")
    D_S = Drift(0.5, name="D_S")
    a = math.pi / 9
    B = SBend(1.0, a, a / 2, a / 2, name="B_DBA")
    QF = Quadrupole(0.4, 1.5, name="Q_F", num_int_steps=10)
    QD = Quadrupole(0.4, -1.2, name="Q_D", num_int_steps=10)
    return Lattice([D_L, B, D_S, QF, D_S, QD, D_S, QF, D_S, B, D_L])


def verify_bend_line():
    """test/verify_bend.jl:109-112"""
    a = math.pi / 4
    return Lattice([SBend(1.0, a, a / 2, a / 2, name="B_DBA")])


def verify_exact_bend_line():
    """test/verify_exact_bend.jl:162-166"""
    a = math.pi / 9
    return Lattice([ExactSBend(1.0, a, a / 2, a / 2, name="B_EXACT", num_int_steps=10)])


def element_cases():
    """Hot-path element cases of test/verify_elements.jl:127-147 (name, element)."""
    return [
        ("Marker", Marker()),
        ("Drift", Drift(0.7)),
        ("Quadrupole", Quadrupole(0.4, 1.3, num_int_steps=10)),
        ("Sextupole", Sextupole(0.4, 2.0, num_int_steps=10)),
        ("Octupole", Octupole(0.4, -3.0, num_int_steps=10)),
        ("ThinMultipole", ThinMultipole(0.0, [0.0, 0.0, 0.0, 0.0], [0.0, 0.2, -0.1, 0.05], max_order=3)),
        ("SBend", SBend(0.9, 0.15, 0.03, 0.02, num_int_steps=10)),
        ("RBend", RBend(0.9, 0.15, num_int_steps=10)),
        ("ExactSBend", ExactSBend(0.9, 0.15, 0.03, 0.02, num_int_steps=10)),
        ("ERBend", ERBend(0.9, 0.15, num_int_steps=10)),
        ("RFCavity", RFCavity(0.3, 2.0e6, 500.0e6, 0.002, h=1200.0, philag=0.3, energy=ENERGY_VAL)),
    ]


# Elements whose map uses only + - * / sqrt on particle coordinates: the parity build
# must match the oracle bit for bit.  The rest call asin/sin/tan/atan/cos per particle.
BITEXACT = {"Marker", "Drift", "Quadrupole", "Sextupole", "Octupole", "ThinMultipole", "SBend", "RBend"}


def forest_ring(energy):
    """Four cells of the reference's ExactSBend with yrot (e1, e2 != 0), Forest fringe (gK != 0)
    and wedges, one bend with a sextupole component (non-identity kicks between the body's
    sub-steps), quads, a sextupole, and an RF cavity (h = 400)."""
    cell = [Drift(1.0), Quadrupole(0.4, 1.2), Drift(0.5), ExactSBend(1.0, 0.1, 0.03, 0.02, gk=0.01), Drift(0.3),
            Sextupole(0.2, 10.0), Drift(0.2), Quadrupole(0.4, -1.2), Drift(0.5),
            ExactSBend(1.0, 0.1, 0.0, 0.05, gk=0.02, polynom_b=[0.0, 0.0, 3.0, 0.0], max_order=2, num_int_steps=4),
            Drift(1.0)]
    els = cell * 4
    circ = sum(e.L for e in els)
    els.append(RFCavity(0.0, 1e6, 400 * 2.99792458e8 / circ, energy=energy))
    return Lattice(els)
