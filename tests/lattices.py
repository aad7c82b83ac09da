"""Lattices of the reference's own tests (test/verify_*.jl) and element cases of
test/verify_elements.jl:127-147, built with the host mirror."""

import math

import numpy as np

from trackpad_hip import (M_ELECTRON, AccelCavity, Corrector, CrabCavity, Drift, DriftSC, ERBend, ExactSBend,
                          InvLorentzBoost, Lattice, LBend, LongitudinalRFMap, LongitudinalRLCWake, LongitudinalWake,
                          LorentzBoost, Marker, Octupole, Quadrupole, QuadrupoleSC, RBend, RFCavity, SBend, SBendSC,
                          Sextupole, SextupoleSC, Solenoid, SpaceCharge, StrongGaussianBeam, StrongThinGaussianBeam,
                          ThinMultipole, Translation, YRotation)

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


def tilt(theta):
    """6 x 6 roll about the s axis by theta (x, y and px, py rotated together)."""
    c, s_ = np.cos(theta), np.sin(theta)
    R = np.eye(6)
    for i, j in ((0, 2), (1, 3)):
        R[i, i], R[i, j], R[j, i], R[j, j] = c, s_, -s_, c
    return R


def families_ring():
    """Every family the reverse pass covers beyond the hot lattices: solenoid (misaligned),
    corrector, linear bend with edges, the SC wrappers, Lorentz boosts around a thin beam-beam
    lens, a sliced strong beam, translations, y-rotations, a crab cavity with length (its
    phase taken before the first half drift), an accelerating cavity, the longitudinal RF map,
    both wakes, space charge (a marker), and quadrupoles rolled by r1 / r2 with t1 / t2
    offsets, in FODO cells."""
    q = lambda k: Quadrupole(0.4, k, num_int_steps=8)
    rolled = Quadrupole(0.4, 1.2, num_int_steps=8, r1=tilt(2e-3), r2=tilt(-2e-3) * 1.0001,
                        t1=[1e-4, 0, -5e-5, 0, 0, 0], t2=[-1e-4, 0, 5e-5, 0, 0, 0])
    els = [q(1.2), Drift(1.0), Solenoid(0.4, 0.3, t1=[1e-4, 0, -2e-4, 0, 0, 0], t2=[-1e-4, 0, 2e-4, 0, 0, 0]),
           Drift(0.5), q(-1.2), Drift(1.0), Corrector(0.2, 1e-5, -2e-5),
           LBend(0.8, 0.05, 0.01, 0.02, K=0.1, fint1=0.3, fint2=0.3, full_gap=0.02), Drift(0.5), rolled, Drift(1.0),
           QuadrupoleSC(0.4, k1=-1.2), DriftSC(1.0), SextupoleSC(0.2, k2=3.0), SBendSC(0.8, 0.05, 0.025, 0.025),
           Drift(0.5), LorentzBoost(0.01), StrongThinGaussianBeam(1e-7, 1e-3, 1e-3, xoffset=1e-4),
           InvLorentzBoost(0.01),
           StrongGaussianBeam(3e-3, M_ELECTRON, 1.0, 10 ** 6, 3e9, [1e-3, 1.2e-3], nzslice=3,
                              xoffsets=[0, 1e-4, -1e-4], zslice_npar=[0.2, 0.5, 0.3]),
           Translation(0.0, dx=1e-4, dy=-2e-4, ds=1e-3), YRotation(0.0, angle=1e-3), Drift(0.3),
           YRotation(0.0, angle=-1.5e-3), CrabCavity(0.3, volt=1e5, freq=5e8, phi=0.1, energy=3e9),
           AccelCavity(0.0, volt=1e6, freq=5e8, h=400.0, phis=0.1, energy=3e9),
           LongitudinalRFMap(1e-3, AccelCavity(0.0, volt=1e6, freq=5e8, h=400.0)),
           LongitudinalRLCWake(freq=1e9, Rshunt=1e4, Q0=2.0, scale=1e-20),
           LongitudinalWake([-4e-12, -1e-12, 0.0, 2e-12, 5e-12], [0.5, -1.0, 3.0, 5.0, 1.0], scale=1e-7),
           SpaceCharge(0.0)]
    return Lattice(els)
