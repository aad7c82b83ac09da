"""Algorithmic FP64 operation counts of the reference maps (the roofline's numerator).

Counting rule: every IEEE + - * / sqrt applied to a particle-dependent value as written in
the reference is one flop; subexpressions of element or beam constants are folded (the
reference recomputes them per call, the device table holds them); transcendental calls
(sin, tan, asin, atan, cos) are not flops.  Per-primitive counts, with the reference lines
they follow:

    drift6                      19            src/tracking.jl:87-99
    strthinkick (order m)       8m + 4, m = 0: 2     :284-305
    bndthinkick (order m)       8m + 11, m = 0: 10   :243-270
    exact_bend_body             39            :858-894
    linear edge (one side)      8             :160-221
    RF kick                     7             :508-513
    yrot                        23            :727-745
    bend_edge                   40            :822-851
    bend_fringe                 128           :752-815
    Solenoid                    83            :542-560
    Corrector                   25            :586-596
    Translation / YRotation     17 / 26       :1182-1218
    Crab / Accel cavity         6 / 4         :1531-1559
    LongitudinalRFMap, boosts   2, 7          :1575-1604
    Gaussian beam-beam          15 per slice  :1610-1641
    RLC / tabulated wake        12 / 6        :1647-1663
    LBend                       81 + edges    :1076-1170
    Wiggler                     per step 3 x (90 per harmonic + 35), transcendentals not counted
                                              :1224-1525

so one Yoshida-4 step of a thick multipole is 4 x 19 + 3 x (8m + 4).  With these counts
the 20-cell DBA ring of the benchmark is 313,587 flops per particle-turn and the FODO
workload 7,472.
"""

from __future__ import annotations

from .elements import (AccelCavity, Corrector, CrabCavity, Drift, DriftSC, ExactSBend, ExactSBendLinearEdge,
                       InvLorentzBoost, LBend, LongitudinalRFMap, LongitudinalRLCWake, LongitudinalWake, LorentzBoost,
                       Marker, Octupole, OctupoleSC, Quadrupole, QuadrupoleSC, RFCavity, SBend, SBendSC, Sextupole,
                       SextupoleSC, Solenoid, SpaceCharge, StrongGaussianBeam, StrongThinGaussianBeam, ThinMultipole,
                       Translation, Wiggler, YRotation)
from .lattice import Lattice

__all__ = ["elem_flops", "flops_per_turn"]

DRIFT, EBODY, EDGE, RF, YROT, BEDGE, BFRINGE = 19, 39, 8, 7, 23, 40, 128


def _kick(m: int) -> int:
    return 8 * m + 4 if m >= 1 else 2


def _bkick(m: int) -> int:
    return 8 * m + 11 if m >= 1 else 10


def elem_flops(e) -> int:
    if isinstance(e, (Marker, SpaceCharge)):
        return 0
    if isinstance(e, (Drift, DriftSC)):
        return DRIFT
    if isinstance(e, (Quadrupole, Sextupole, Octupole)):
        return e.num_int_steps * (4 * DRIFT + 3 * _kick(e.max_order))
    if isinstance(e, (QuadrupoleSC, SextupoleSC, OctupoleSC)):
        return e.num_int_steps * (4 * DRIFT + 3 * _kick(min(max(e.max_order, 0), 3)))
    if isinstance(e, SBendSC):
        edges = EDGE * ((e.fringe_bend_entrance != 0) + (e.fringe_bend_exit != 0))
        return edges + e.num_int_steps * (4 * DRIFT + 3 * _bkick(min(max(e.max_order, 0), 3)))
    if isinstance(e, Solenoid):
        return DRIFT if e.ks == 0 else 83
    if isinstance(e, Corrector):
        return 25
    if isinstance(e, Translation):
        return 17
    if isinstance(e, YRotation):
        return 26 if e.angle != 0 else 0
    if isinstance(e, CrabCavity):
        return 6 + (2 * DRIFT if e.L > 0 else 0)
    if isinstance(e, AccelCavity):
        return 4
    if isinstance(e, LongitudinalRFMap):
        return 2
    if isinstance(e, (LorentzBoost, InvLorentzBoost)):
        return 7 if e.mode == 0 else 0
    if isinstance(e, StrongThinGaussianBeam):
        return 15
    if isinstance(e, StrongGaussianBeam):
        return 15 * max(e.nzslice, 0)
    if isinstance(e, LongitudinalRLCWake):
        return 12 if e.scale != 0 else 0
    if isinstance(e, LongitudinalWake):
        return 6 if e.scale != 0 else 0
    if isinstance(e, LBend):
        return 81 + 2 * EDGE
    if isinstance(e, Wiggler):
        # per step three 2nd-order maps, each six potential evaluations of ~15 operations per
        # harmonic and ~35 for the drifts; the radiation kick adds two potentials, the field
        # (~12 per harmonic) and ~15
        nhv = min(e.NHharm, len(e.By) // 6) + min(e.NVharm, len(e.Bx) // 6)
        nstep = e.Nsteps * max(1, round(e.L / e.lw)) if e.lw else 0
        rad = (42 * nhv + 15) if e.rad == 1 else 0
        return nstep * (3 * (90 * nhv + 35) + rad) + rad
    if isinstance(e, ThinMultipole):
        return _kick(e.max_order)
    if isinstance(e, SBend):
        edges = EDGE * ((e.fringe_bend_entrance != 0) + (e.fringe_bend_exit != 0))
        return edges + e.num_int_steps * (4 * DRIFT + 3 * _bkick(e.max_order))
    if isinstance(e, (ExactSBend, ExactSBendLinearEdge)):
        body = EBODY if e.num_int_steps == 0 else e.num_int_steps * (4 * EBODY + 3 * _kick(e.max_order))
        if isinstance(e, ExactSBendLinearEdge):
            return body + EDGE * ((e.fringe_bend_entrance != 0) + (e.fringe_bend_exit != 0))
        extra = 2 * BEDGE + BFRINGE * ((e.fringe_bend_entrance != 0) + (e.fringe_bend_exit != 0))
        extra += YROT * ((e.e1 != 0.0) + (e.e2 != 0.0))
        return body + extra
    if isinstance(e, RFCavity):
        return (2 * DRIFT if e.L > 0 else 0) + (RF if e.energy > 0 else 0)
    raise NotImplementedError(type(e).__name__)


def flops_per_turn(lat: Lattice) -> int:
    return sum(elem_flops(e) for e in lat.elements)
