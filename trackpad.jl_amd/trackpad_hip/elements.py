"""Element types of the tracking path, mirroring TrackPad.jl's element structs.

Field names, positional arguments, keyword arguments and defaults follow the reference
constructors so that a lattice written for TrackPad reads the same here:

    Drift          src/elements/drift.jl:6-50
    Marker         src/elements/marker.jl
    Quadrupole     src/elements/quad.jl:6-111
    Sextupole      src/elements/multipole.jl:6-47
    Octupole       src/elements/multipole.jl:61-102
    ThinMultipole  src/elements/multipole.jl:116-155
    SBend, RBend   src/elements/bend.jl:6-74, 243
    ExactSBend     src/elements/bend.jl:88-149
    ERBend         src/elements/SCelements.jl:339
    RFCavity       src/elements/rf.jl:6-21
    CrabCavity, AccelCavity, LongitudinalRFMap      src/elements/rf.jl:35-109
    Solenoid       src/elements/solenoid.jl:6-30
    Corrector, HKicker, VKicker                     src/elements/corrector.jl:6-51
    DriftSC, QuadrupoleSC, SextupoleSC, OctupoleSC, SBendSC, RBendSC, LBend,
    SpaceCharge, Translation, YRotation             src/elements/SCelements.jl:6-460
    LorentzBoost, InvLorentzBoost, StrongThinGaussianBeam, StrongGaussianBeam
                                                    src/elements/beambeam.jl:6-125
    LongitudinalRLCWake, LongitudinalWake           src/elements/wake.jl:6-90
    Wiggler        src/elements/wiggler.jl:6-44

Two deliberate extensions, both absent from the reference:

* polynomials may hold up to ``MAXORD + 1 = 11`` coefficients (the reference stores
  ``SVector{4}``, i.e. ``max_order <= 3``, src/elements/quad.jl:10-11); the kick is the
  same Horner recurrence (src/tracking.jl:290-298) run to ``max_order``;
* :class:`ExactSBendLinearEdge`, the exact sector-bend body of ``ExactSBend``
  (src/tracking.jl:858-894, 923-953) entered and left through the linear edge maps of
  ``SBend`` (src/tracking.jl:160-221) instead of Forest's fringe/wedge maps.

Elements are plain mutable Python objects; nothing here touches a device.
"""

from __future__ import annotations

import numpy as np

MAXORD = 10

__all__ = [
    "MAXORD", "AbstractElement", "AbstractMagnet", "AbstractDrift", "AbstractCavity",
    "Drift", "Marker", "Quadrupole", "Sextupole", "Octupole", "ThinMultipole",
    "SBend", "RBend", "ExactSBend", "ERBend", "ExactSBendLinearEdge", "RFCavity",
    "TimeVaryingElement", "timed", "C_LIGHT", "get_length",
    "CrabCavity", "AccelCavity", "LongitudinalRFMap", "Solenoid", "Corrector", "HKicker", "VKicker",
    "DriftSC", "QuadrupoleSC", "SextupoleSC", "OctupoleSC", "SBendSC", "RBendSC", "LBend", "SpaceCharge",
    "Translation", "YRotation", "LorentzBoost", "InvLorentzBoost", "StrongThinGaussianBeam",
    "StrongGaussianBeam", "LongitudinalRLCWake", "LongitudinalWake", "Wiggler",
]

# src/elements/rf.jl:349
C_LIGHT = 2.99792458e8


class AbstractElement:
    """Root of the element hierarchy (src/elements/core.jl:1)."""

    name: str = "ELEMENT"

    def __repr__(self) -> str:
        fields = ", ".join(f"{k}={v!r}" for k, v in vars(self).items()
                           if not isinstance(v, np.ndarray))
        return f"{type(self).__name__}({fields})"


class AbstractMagnet(AbstractElement):
    """src/elements/core.jl:2"""


class AbstractDrift(AbstractElement):
    """src/elements/core.jl:3"""


class AbstractCavity(AbstractElement):
    """src/elements/core.jl:4"""


def _vec(v, n: int, what: str) -> np.ndarray:
    """_default_vec (src/elements/core.jl:19): zeros when not given."""
    if v is None:
        return np.zeros(n, dtype=np.float64)
    a = np.array(v, dtype=np.float64).reshape(-1)
    if a.shape[0] != n:
        raise ValueError(f"{what} must have {n} entries, got {a.shape[0]}")
    return a


def _poly(v, what: str) -> np.ndarray:
    """Polynomial coefficients: the reference's SVector{4} default, extendable to 11."""
    if v is None:
        return np.zeros(4, dtype=np.float64)
    a = np.array(v, dtype=np.float64).reshape(-1)
    if a.shape[0] < 1 or a.shape[0] > MAXORD + 1:
        raise ValueError(f"{what} must have 1..{MAXORD + 1} coefficients, got {a.shape[0]}")
    return a


def _mat(v, what: str) -> np.ndarray:
    """_default_mat (src/elements/core.jl:20): 6x6 zeros when not given."""
    if v is None:
        return np.zeros((6, 6), dtype=np.float64)
    a = np.array(v, dtype=np.float64)
    if a.shape != (6, 6):
        raise ValueError(f"{what} must be 6x6, got {a.shape}")
    return a


class _Misaligned:
    def _init_misalign(self, t1, t2, r1, r2, r_apertures, e_apertures):
        self.t1 = _vec(t1, 6, "t1")
        self.t2 = _vec(t2, 6, "t2")
        self.r1 = _mat(r1, "r1")
        self.r2 = _mat(r2, "r2")
        # stored but never read by the tracker (src/elements/drift.jl:13-14; grep)
        self.r_apertures = _vec(r_apertures, 6, "r_apertures")
        self.e_apertures = _vec(e_apertures, 6, "e_apertures")


class Drift(AbstractDrift, _Misaligned):
    """Drift(L; name=:DRIFT, t1, t2, r1, r2, r_apertures, e_apertures) -- src/elements/drift.jl:37-50"""

    def __init__(self, L, *, name="DRIFT", t1=None, t2=None, r1=None, r2=None,
                 r_apertures=None, e_apertures=None):
        self.name = str(name)
        self.L = float(L)
        self._init_misalign(t1, t2, r1, r2, r_apertures, e_apertures)


class Marker(AbstractElement):
    """Marker(; name=:MARKER) -- src/elements/marker.jl"""

    def __init__(self, *, name="MARKER"):
        self.name = str(name)


class _ThickMultipole(AbstractMagnet, _Misaligned):
    _default_max_order = 1

    def _init_mp(self, L, polynom_a, polynom_b, max_order, num_int_steps, rad,
                 fringe_entrance, fringe_exit, t1, t2, r1, r2, r_apertures, e_apertures,
                 kick_angle):
        self.L = float(L)
        self.polynom_a = _poly(polynom_a, "polynom_a")
        self.polynom_b = _poly(polynom_b, "polynom_b")
        self.max_order = int(max_order)
        self.num_int_steps = int(num_int_steps)
        self.rad = int(rad)
        self.fringe_entrance = int(fringe_entrance)
        self.fringe_exit = int(fringe_exit)
        self._init_misalign(t1, t2, r1, r2, r_apertures, e_apertures)
        self.kick_angle = _vec(kick_angle, 2, "kick_angle")


class Quadrupole(_ThickMultipole):
    """Quadrupole(L, k1; max_order=1, num_int_steps=10, ...) -- src/elements/quad.jl:90-111"""

    def __init__(self, L, k1, *, name="QUAD", polynom_a=None, polynom_b=None, max_order=1,
                 num_int_steps=10, rad=0, fringe_entrance=0, fringe_exit=0, t1=None, t2=None,
                 r1=None, r2=None, r_apertures=None, e_apertures=None, kick_angle=None):
        self.name = str(name)
        self.k1 = float(k1)
        self._init_mp(L, polynom_a, polynom_b, max_order, num_int_steps, rad, fringe_entrance,
                      fringe_exit, t1, t2, r1, r2, r_apertures, e_apertures, kick_angle)


class Sextupole(_ThickMultipole):
    """Sextupole(L, k2; max_order=2, num_int_steps=10, ...) -- src/elements/multipole.jl:26-47"""

    def __init__(self, L, k2, *, name="SEXT", polynom_a=None, polynom_b=None, max_order=2,
                 num_int_steps=10, rad=0, fringe_entrance=0, fringe_exit=0, t1=None, t2=None,
                 r1=None, r2=None, r_apertures=None, e_apertures=None, kick_angle=None):
        self.name = str(name)
        self.k2 = float(k2)
        self._init_mp(L, polynom_a, polynom_b, max_order, num_int_steps, rad, fringe_entrance,
                      fringe_exit, t1, t2, r1, r2, r_apertures, e_apertures, kick_angle)


class Octupole(_ThickMultipole):
    """Octupole(L, k3; max_order=3, num_int_steps=10, ...) -- src/elements/multipole.jl:81-102"""

    def __init__(self, L, k3, *, name="OCT", polynom_a=None, polynom_b=None, max_order=3,
                 num_int_steps=10, rad=0, fringe_entrance=0, fringe_exit=0, t1=None, t2=None,
                 r1=None, r2=None, r_apertures=None, e_apertures=None, kick_angle=None):
        self.name = str(name)
        self.k3 = float(k3)
        self._init_mp(L, polynom_a, polynom_b, max_order, num_int_steps, rad, fringe_entrance,
                      fringe_exit, t1, t2, r1, r2, r_apertures, e_apertures, kick_angle)


class ThinMultipole(AbstractElement, _Misaligned):
    """ThinMultipole(L, polynom_a, polynom_b; max_order=1, ...) -- src/elements/multipole.jl:135-155"""

    def __init__(self, L, polynom_a, polynom_b, *, name="MULTIPOLE", max_order=1,
                 num_int_steps=1, rad=0, fringe_entrance=0, fringe_exit=0, t1=None, t2=None,
                 r1=None, r2=None, r_apertures=None, e_apertures=None, kick_angle=None):
        self.name = str(name)
        self.L = float(L)
        self.polynom_a = _poly(polynom_a, "polynom_a")
        self.polynom_b = _poly(polynom_b, "polynom_b")
        self.max_order = int(max_order)
        self.num_int_steps = int(num_int_steps)
        self.rad = int(rad)
        self.fringe_entrance = int(fringe_entrance)
        self.fringe_exit = int(fringe_exit)
        self._init_misalign(t1, t2, r1, r2, r_apertures, e_apertures)
        self.kick_angle = _vec(kick_angle, 2, "kick_angle")


class _Bend(AbstractMagnet, _Misaligned):
    def _init_bend(self, L, angle, e1, e2, polynom_a, polynom_b, max_order, num_int_steps, rad,
                   fint1, fint2, gap, fringe_bend_entrance, fringe_bend_exit,
                   fringe_quad_entrance, fringe_quad_exit, fringe_int_m0, fringe_int_p0,
                   t1, t2, r1, r2, r_apertures, e_apertures, kick_angle):
        self.L = float(L)
        self.angle = float(angle)
        self.e1 = float(e1)
        self.e2 = float(e2)
        self.polynom_a = _poly(polynom_a, "polynom_a")
        self.polynom_b = _poly(polynom_b, "polynom_b")
        self.max_order = int(max_order)
        self.num_int_steps = int(num_int_steps)
        self.rad = int(rad)
        self.fint1 = float(fint1)
        self.fint2 = float(fint2)
        self.gap = float(gap)
        self.fringe_bend_entrance = int(fringe_bend_entrance)
        self.fringe_bend_exit = int(fringe_bend_exit)
        self.fringe_quad_entrance = int(fringe_quad_entrance)
        self.fringe_quad_exit = int(fringe_quad_exit)
        self.fringe_int_m0 = _vec(fringe_int_m0, 5, "fringe_int_m0")
        self.fringe_int_p0 = _vec(fringe_int_p0, 5, "fringe_int_p0")
        self._init_misalign(t1, t2, r1, r2, r_apertures, e_apertures)
        self.kick_angle = _vec(kick_angle, 2, "kick_angle")


class SBend(_Bend):
    """SBend(L, angle, e1=0, e2=0; max_order=0, num_int_steps=10, fint1, fint2, gap,
    fringe_bend_entrance=1, fringe_bend_exit=1, ...) -- src/elements/bend.jl:197-226"""

    def __init__(self, L, angle, e1=0.0, e2=0.0, *, name="SBEND", polynom_a=None, polynom_b=None,
                 max_order=0, num_int_steps=10, rad=0, fint1=0.0, fint2=0.0, gap=0.0,
                 fringe_bend_entrance=1, fringe_bend_exit=1, fringe_quad_entrance=0,
                 fringe_quad_exit=0, fringe_int_m0=None, fringe_int_p0=None, t1=None, t2=None,
                 r1=None, r2=None, r_apertures=None, e_apertures=None, kick_angle=None):
        self.name = str(name)
        self._init_bend(L, angle, e1, e2, polynom_a, polynom_b, max_order, num_int_steps, rad,
                        fint1, fint2, gap, fringe_bend_entrance, fringe_bend_exit,
                        fringe_quad_entrance, fringe_quad_exit, fringe_int_m0, fringe_int_p0,
                        t1, t2, r1, r2, r_apertures, e_apertures, kick_angle)


def RBend(L, angle, **kwargs) -> SBend:
    """Rectangular bend: SBend with e1 = e2 = angle/2 (src/elements/bend.jl:243)."""
    return SBend(L, angle, angle / 2, angle / 2, **kwargs)


class ExactSBend(_Bend):
    """ExactSBend(L, angle, e1=0, e2=0; ..., gk=0.0) -- src/elements/bend.jl:280-311"""

    def __init__(self, L, angle, e1=0.0, e2=0.0, *, name="ESBEND", polynom_a=None,
                 polynom_b=None, max_order=0, num_int_steps=10, rad=0, fint1=0.0, fint2=0.0,
                 gap=0.0, fringe_bend_entrance=1, fringe_bend_exit=1, fringe_quad_entrance=0,
                 fringe_quad_exit=0, fringe_int_m0=None, fringe_int_p0=None, t1=None, t2=None,
                 r1=None, r2=None, r_apertures=None, e_apertures=None, kick_angle=None, gk=0.0):
        self.name = str(name)
        self._init_bend(L, angle, e1, e2, polynom_a, polynom_b, max_order, num_int_steps, rad,
                        fint1, fint2, gap, fringe_bend_entrance, fringe_bend_exit,
                        fringe_quad_entrance, fringe_quad_exit, fringe_int_m0, fringe_int_p0,
                        t1, t2, r1, r2, r_apertures, e_apertures, kick_angle)
        self.gk = float(gk)


def ERBend(L, angle, **kwargs) -> ExactSBend:
    """Exact rectangular bend: ExactSBend with e1 = e2 = angle/2 (src/elements/SCelements.jl:339)."""
    return ExactSBend(L, angle, angle / 2, angle / 2, **kwargs)


class ExactSBendLinearEdge(_Bend):
    """Exact sector-bend body with linear edge maps (no reference counterpart).

    Tracking order: entrance misalignment; ``edge_fringe_entrance(irho, e1, fint1, gap,
    fringe_bend_entrance)`` if ``fringe_bend_entrance != 0`` (src/tracking.jl:160-186);
    the ExactSBend body (src/tracking.jl:923-953); ``edge_fringe_exit(irho, e2, fint2, gap,
    fringe_bend_exit)`` (src/tracking.jl:194-221); exit misalignment.  Keywords are those of
    :class:`SBend`; the fringe values select the linear edge method (1 Brown, 2 SOLEIL,
    3 THOMX) exactly as for SBend.
    """

    def __init__(self, L, angle, e1=0.0, e2=0.0, *, name="ESBEND_LE", polynom_a=None,
                 polynom_b=None, max_order=0, num_int_steps=10, rad=0, fint1=0.0, fint2=0.0,
                 gap=0.0, fringe_bend_entrance=1, fringe_bend_exit=1, fringe_quad_entrance=0,
                 fringe_quad_exit=0, fringe_int_m0=None, fringe_int_p0=None, t1=None, t2=None,
                 r1=None, r2=None, r_apertures=None, e_apertures=None, kick_angle=None):
        self.name = str(name)
        self._init_bend(L, angle, e1, e2, polynom_a, polynom_b, max_order, num_int_steps, rad,
                        fint1, fint2, gap, fringe_bend_entrance, fringe_bend_exit,
                        fringe_quad_entrance, fringe_quad_exit, fringe_int_m0, fringe_int_p0,
                        t1, t2, r1, r2, r_apertures, e_apertures, kick_angle)


class RFCavity(AbstractCavity):
    """RFCavity(L, volt, freq, lag=0.0; h=1.0, philag=0.0, energy=0.0) -- src/elements/rf.jl:338-342

    As in the reference, ``h`` is stored but not used by the map (src/tracking.jl:510) and
    the cavity gives no kick unless ``energy > 0`` (src/tracking.jl:508).
    """

    def __init__(self, L, volt, freq, lag=0.0, *, name="RFCA", h=1.0, philag=0.0, energy=0.0):
        self.name = str(name)
        self.L = float(L)
        self.volt = float(volt)
        self.freq = float(freq)
        self.h = float(h)
        self.lag = float(lag)
        self.philag = float(philag)
        self.energy = float(energy)


class CrabCavity(AbstractCavity):
    """CrabCavity(L; volt=0, freq=0, phi=0, errors=(0, 0), energy=1e9) -- src/elements/rf.jl:35-54.
    ``k = 2 pi freq / c`` is stored as in the reference."""

    def __init__(self, L, *, name="CRABCAVITY", volt=0.0, freq=0.0, phi=0.0, errors=None, energy=1.0e9):
        self.name = str(name)
        self.L = float(L)
        self.volt = float(volt)
        self.freq = float(freq)
        self.k = 6.283185307179586 * self.freq / C_LIGHT
        self.phi = float(phi)
        self.errors = _vec(errors, 2, "errors")
        self.energy = float(energy)


class AccelCavity(AbstractCavity):
    """AccelCavity(L; volt=0, freq=0, h=1, phis=0, energy=1e9) -- src/elements/rf.jl:69-87."""

    def __init__(self, L, *, name="ACCELCAVITY", volt=0.0, freq=0.0, h=1.0, phis=0.0, energy=1.0e9):
        self.name = str(name)
        self.L = float(L)
        self.volt = float(volt)
        self.freq = float(freq)
        self.k = 6.283185307179586 * self.freq / C_LIGHT
        self.h = float(h)
        self.phis = float(phis)
        self.energy = float(energy)


class LongitudinalRFMap(AbstractElement):
    """LongitudinalRFMap(alphac, rf) -- src/elements/rf.jl:102-105; a zero-length slip map
    z -= 2 pi h eta / k * delta tied to an RF element (src/tracking.jl:1562-1585)."""

    def __init__(self, alphac, rf: AbstractElement, *, name="LONGITUDINALRFMAP"):
        self.name = str(name)
        self.L = 0.0
        self.alphac = float(alphac)
        self.rf = rf


class Solenoid(AbstractMagnet, _Misaligned):
    """Solenoid(L, ks; t1, t2, r1, r2) -- src/elements/solenoid.jl:6-26."""

    def __init__(self, L, ks, *, name="SOLENOID", t1=None, t2=None, r1=None, r2=None):
        self.name = str(name)
        self.L = float(L)
        self.ks = float(ks)
        self._init_misalign(t1, t2, r1, r2, None, None)


class Corrector(AbstractMagnet, _Misaligned):
    """Corrector(L, hk, vk; t1, t2, r1, r2) -- src/elements/corrector.jl:6-27."""

    def __init__(self, L, hk, vk, *, name="CORRECTOR", t1=None, t2=None, r1=None, r2=None):
        self.name = str(name)
        self.L = float(L)
        self.xkick = float(hk)
        self.ykick = float(vk)
        self._init_misalign(t1, t2, r1, r2, None, None)


def HKicker(*, name="HKICKER", L=0.0, xkick=0.0) -> Corrector:
    """src/elements/corrector.jl:39-41"""
    return Corrector(L, xkick, 0.0, name=name)


def VKicker(*, name="VKICKER", L=0.0, ykick=0.0) -> Corrector:
    """src/elements/corrector.jl:48-50"""
    return Corrector(L, 0.0, ykick, name=name)


class _SCMeta:
    def _init_sc(self, a, b, Nl, Nm, Nsteps):
        # space-charge metadata: carried, not used by the single-particle maps
        self.a, self.b = float(a), float(b)
        self.Nl, self.Nm, self.Nsteps = int(Nl), int(Nm), int(Nsteps)


class DriftSC(AbstractDrift, _Misaligned, _SCMeta):
    """DriftSC(L; a, b, Nl, Nm, Nsteps, t1, t2, r1, r2) -- src/elements/SCelements.jl:6-38;
    tracked as a drift (src/tracking.jl:981-991)."""

    def __init__(self, L, *, name="DRIFT_SC", t1=None, t2=None, r1=None, r2=None, r_apertures=None,
                 e_apertures=None, a=1.0, b=1.0, Nl=10, Nm=10, Nsteps=1):
        self.name = str(name)
        self.L = float(L)
        self._init_misalign(t1, t2, r1, r2, r_apertures, e_apertures)
        self._init_sc(a, b, Nl, Nm, Nsteps)


class _MultipoleSC(AbstractMagnet, _Misaligned, _SCMeta):
    _default_order = 1

    def __init__(self, L, *, name=None, k0=0.0, k1=0.0, k2=0.0, k3=0.0, polynom_a=None, max_order=None,
                 num_int_steps=10, rad=0, fringe_entrance=0, fringe_exit=0, t1=None, t2=None, r1=None, r2=None,
                 r_apertures=None, e_apertures=None, kick_angle=None, a=1.0, b=1.0, Nl=10, Nm=10, Nsteps=1):
        self.name = str(name or self._default_name)
        self.L = float(L)
        self.k0, self.k1, self.k2, self.k3 = float(k0), float(k1), float(k2), float(k3)
        self.polynom_a = _vec(polynom_a, 4, "polynom_a")
        self.max_order = int(self._default_order if max_order is None else max_order)
        self.num_int_steps = int(num_int_steps)
        self.rad = int(rad)
        self.fringe_entrance, self.fringe_exit = int(fringe_entrance), int(fringe_exit)
        self.kick_angle = _vec(kick_angle, 2, "kick_angle")
        self._init_misalign(t1, t2, r1, r2, r_apertures, e_apertures)
        self._init_sc(a, b, Nl, Nm, Nsteps)


class QuadrupoleSC(_MultipoleSC):
    """QuadrupoleSC(L; k0, k1, k2, k3, polynom_a, max_order=1, ...) -- src/elements/SCelements.jl:51-101."""
    _default_order = 1
    _default_name = "KQUAD_SC"


class SextupoleSC(_MultipoleSC):
    """SextupoleSC(L; ..., max_order=2) -- src/elements/SCelements.jl:116-166."""
    _default_order = 2
    _default_name = "KSEXT_SC"


class OctupoleSC(_MultipoleSC):
    """OctupoleSC(L; ..., max_order=3) -- src/elements/SCelements.jl:181-231."""
    _default_order = 3
    _default_name = "KOCT_SC"


class SBendSC(AbstractMagnet, _Misaligned, _SCMeta):
    """SBendSC(L, angle, e1=0, e2=0; ...) -- src/elements/SCelements.jl:246-311; tracked as
    SBend's map with fringe methods gated by the fringe flags (src/tracking.jl:1041-1074)."""

    def __init__(self, L, angle, e1=0.0, e2=0.0, *, name="SBEND_SC", polynom_a=None, polynom_b=None,
                 max_order=0, num_int_steps=10, rad=0, fint1=0.0, fint2=0.0, gap=0.0,
                 fringe_bend_entrance=1, fringe_bend_exit=1, fringe_quad_entrance=0, fringe_quad_exit=0,
                 fringe_int_m0=None, fringe_int_p0=None, t1=None, t2=None, r1=None, r2=None,
                 r_apertures=None, e_apertures=None, kick_angle=None, a=1.0, b=1.0, Nl=10, Nm=10, Nsteps=1):
        self.name = str(name)
        self.L, self.angle, self.e1, self.e2 = float(L), float(angle), float(e1), float(e2)
        self.polynom_a = _vec(polynom_a, 4, "polynom_a")
        self.polynom_b = _vec(polynom_b, 4, "polynom_b")
        self.max_order, self.num_int_steps, self.rad = int(max_order), int(num_int_steps), int(rad)
        self.fint1, self.fint2, self.gap = float(fint1), float(fint2), float(gap)
        self.fringe_bend_entrance, self.fringe_bend_exit = int(fringe_bend_entrance), int(fringe_bend_exit)
        self.fringe_quad_entrance, self.fringe_quad_exit = int(fringe_quad_entrance), int(fringe_quad_exit)
        self.fringe_int_m0 = _vec(fringe_int_m0, 5, "fringe_int_m0")
        self.fringe_int_p0 = _vec(fringe_int_p0, 5, "fringe_int_p0")
        self.kick_angle = _vec(kick_angle, 2, "kick_angle")
        self._init_misalign(t1, t2, r1, r2, r_apertures, e_apertures)
        self._init_sc(a, b, Nl, Nm, Nsteps)


def RBendSC(L, angle, **kw) -> SBendSC:
    """src/elements/SCelements.jl:332"""
    return SBendSC(L, angle, angle / 2, angle / 2, **kw)


class LBend(AbstractMagnet, _Misaligned):
    """LBend(L, angle, e1=0, e2=0; K, by_error, fint1, fint2, full_gap, ...) --
    src/elements/SCelements.jl:346-379; linear-matrix bend (src/tracking.jl:1076-1170)."""

    def __init__(self, L, angle, e1=0.0, e2=0.0, *, name="LBEND", K=0.0, by_error=0.0, fint1=0.0, fint2=0.0,
                 full_gap=0.0, t1=None, t2=None, r1=None, r2=None, r_apertures=None, e_apertures=None):
        self.name = str(name)
        self.L, self.angle, self.e1, self.e2 = float(L), float(angle), float(e1), float(e2)
        self.K, self.by_error = float(K), float(by_error)
        self.fint1, self.fint2, self.full_gap = float(fint1), float(fint2), float(full_gap)
        self._init_misalign(t1, t2, r1, r2, r_apertures, e_apertures)


class SpaceCharge(AbstractElement):
    """SpaceCharge(L; effective_len, Nl, Nm, a, b) -- src/elements/SCelements.jl:394-408; the
    single-particle map is the identity (src/tracking.jl:1176-1180)."""

    def __init__(self, L, *, name="SPACECHARGE", effective_len=0.0, Nl=10, Nm=10, a=1.0, b=1.0):
        self.name = str(name)
        self.L, self.effective_len = float(L), float(effective_len)
        self.Nl, self.Nm, self.a, self.b = int(Nl), int(Nm), float(a), float(b)


class Translation(AbstractElement):
    """Translation(L; dx=0, dy=0, ds=0) -- src/elements/SCelements.jl:420-433."""

    def __init__(self, L, *, name="TRANSLATION", dx=0.0, dy=0.0, ds=0.0):
        self.name = str(name)
        self.L, self.dx, self.dy, self.ds = float(L), float(dx), float(dy), float(ds)


class YRotation(AbstractElement):
    """YRotation(L; angle=0) -- src/elements/SCelements.jl:444-455."""

    def __init__(self, L, *, name="YROTATION", angle=0.0):
        self.name = str(name)
        self.L, self.angle = float(L), float(angle)


class LorentzBoost(AbstractElement):
    """LorentzBoost(angle; mode=0) -- src/elements/beambeam.jl:6-19 (cos and tan stored)."""

    def __init__(self, angle, *, name="LORENTZBOOST", mode=0):
        import math

        self.name = str(name)
        self.L = 0.0
        self.angle = float(angle)
        self.cosang, self.tanang = math.cos(self.angle), math.tan(self.angle)
        self.mode = int(mode)


class InvLorentzBoost(AbstractElement):
    """InvLorentzBoost(angle; mode=0) -- src/elements/beambeam.jl:30-43 (sin and cos stored)."""

    def __init__(self, angle, *, name="INVLORENTZBOOST", mode=0):
        import math

        self.name = str(name)
        self.L = 0.0
        self.angle = float(angle)
        self.sinang, self.cosang = math.sin(self.angle), math.cos(self.angle)
        self.mode = int(mode)


class StrongThinGaussianBeam(AbstractElement):
    """StrongThinGaussianBeam(amplitude, rmssizex, rmssizey; zloc, xoffset, yoffset) --
    src/elements/beambeam.jl:54-69."""

    def __init__(self, amplitude, rmssizex, rmssizey, *, name="STRONGTHINGAUSSIANBEAM", zloc=0.0, xoffset=0.0,
                 yoffset=0.0):
        self.name = str(name)
        self.L = 0.0
        self.amplitude, self.rmssizex, self.rmssizey = float(amplitude), float(rmssizex), float(rmssizey)
        self.zloc, self.xoffset, self.yoffset = float(zloc), float(xoffset), float(yoffset)


class StrongGaussianBeam(AbstractElement):
    """StrongGaussianBeam(charge, mass, atomnum, num_particle, total_energy, beamsize; nzslice=1,
    zslice_center, zslice_npar, xoffsets, yoffsets) -- src/elements/beambeam.jl:83-119."""

    def __init__(self, charge, mass, atomnum, num_particle, total_energy, beamsize, *, name="STRONGGAUSSIANBEAM",
                 nzslice=1, zslice_center=None, zslice_npar=None, xoffsets=None, yoffsets=None):
        import math

        self.name = str(name)
        self.L = 0.0
        self.charge, self.mass, self.atomnum = float(charge), float(mass), float(atomnum)
        self.num_particle = int(num_particle)
        self.total_energy = float(total_energy)
        mom2 = max(self.total_energy ** 2 - self.mass ** 2, 0.0)
        self.momentum = math.sqrt(mom2)
        self.gamma = self.total_energy / self.mass
        self.beta = 0.0 if self.total_energy == 0 else self.momentum / self.total_energy
        self.beamsize = np.array([float(beamsize[0]), float(beamsize[1])])
        self.nzslice = int(nzslice)
        n = self.nzslice
        self.zslice_center = (np.linspace(-1.0, 1.0, n) if zslice_center is None
                              else np.array(zslice_center, dtype=np.float64))
        self.zslice_npar = (np.full(n, 1.0 / max(n, 1)) if zslice_npar is None
                            else np.array(zslice_npar, dtype=np.float64))
        self.xoffsets = np.zeros(n) if xoffsets is None else np.array(xoffsets, dtype=np.float64)
        self.yoffsets = np.zeros(n) if yoffsets is None else np.array(yoffsets, dtype=np.float64)


class LongitudinalRLCWake(AbstractElement):
    """LongitudinalRLCWake(; freq=1e9, Rshunt=1e6, Q0=1, scale=0) -- src/elements/wake.jl:6-18."""

    def __init__(self, *, name="LONGITUDINALRLCWAKE", freq=1.0e9, Rshunt=1.0e6, Q0=1.0, scale=0.0):
        self.name = str(name)
        self.L = 0.0
        self.freq, self.Rshunt, self.Q0, self.scale = float(freq), float(Rshunt), float(Q0), float(scale)


class LongitudinalWake(AbstractElement):
    """LongitudinalWake(times, wakefields; fliphalf=-1, scale=0) -- src/elements/wake.jl:44-59."""

    def __init__(self, times, wakefields, *, name="LONGITUDINALWAKE", fliphalf=-1.0, scale=0.0):
        t = np.array(times, dtype=np.float64).reshape(-1)
        w = np.array(wakefields, dtype=np.float64).reshape(-1)
        if t.shape != w.shape:
            raise ValueError("times and wakefields must have the same length")
        if t.shape[0] < 2:
            raise ValueError("LongitudinalWake requires at least two samples")
        self.name = str(name)
        self.L = 0.0
        self.times, self.wakefields = t, w
        self.fliphalf, self.scale = float(fliphalf), float(scale)


class Wiggler(AbstractMagnet, _Misaligned):
    """Wiggler(L; lw=0, Bmax=0, Nsteps=10, By=[1,1,0,1,1,0], Bx=[], energy=1e9, rad=0, t1, t2,
    r1, r2) -- src/elements/wiggler.jl:6-44.  By / Bx hold 6 integers per harmonic."""

    def __init__(self, L, *, name="WIGGLER", lw=0.0, Bmax=0.0, Nsteps=10, By=(1, 1, 0, 1, 1, 0), Bx=(),
                 energy=1.0e9, rad=0, t1=None, t2=None, r1=None, r2=None):
        self.name = str(name)
        self.L, self.lw, self.Bmax = float(L), float(lw), float(Bmax)
        self.Nsteps = int(Nsteps)
        self.By = [int(v) for v in By]
        self.Bx = [int(v) for v in Bx]
        self.energy = float(energy)
        self.NHharm, self.NVharm = len(self.By) // 6, len(self.Bx) // 6
        self.rad = int(rad)
        self._init_misalign(t1, t2, r1, r2, None, None)


class TimeVaryingElement(AbstractElement):
    """Element with time- or turn-dependent parameters (src/time_dependence.jl:127-136).

    ``params`` maps field names to numbers or to callables ``f(time, turn)`` (the
    reference's ``TimeDependentParam`` of a ``TimeContext``).  The device table resolves the
    element once per turn (``compile_table(..., turn=, nturns=, time=, dt_turn=)``), as
    ``linepass!`` does at the start of every call (src/lattice.jl:267).
    """

    def __init__(self, base: AbstractElement, params: dict):
        for k in params:
            if not hasattr(base, k):
                raise ValueError(f"{type(base).__name__} has no field {k!r}")  # src/time_dependence.jl:150
        self.base = base
        self.params = dict(params)
        self.name = getattr(base, "name", "TVE")

    def materialize(self, time: float = 0.0, turn: int = 0) -> AbstractElement:
        """Static copy with every parameter evaluated at (time, turn) and converted to the
        field's type (src/time_dependence.jl:163-174)."""
        import copy

        e = copy.deepcopy(self.base)
        for k, f in self.params.items():
            v = f(time, turn) if callable(f) else f
            old = getattr(self.base, k)
            if isinstance(old, (bool, int)) and not isinstance(old, float):
                v = int(v)
            elif isinstance(old, float):
                v = float(v)
            setattr(e, k, v)
        return e


def timed(elem: AbstractElement, **params) -> TimeVaryingElement:
    """timed(elem; kwargs...) (src/time_dependence.jl:141-155): any field of ``elem`` given as
    a number or as a callable ``f(time, turn)``."""
    return TimeVaryingElement(elem, params)


def get_length(elem: AbstractElement) -> float:
    """get_length (src/lattice.jl:139-169)."""
    if isinstance(elem, Marker):
        return 0.0
    if isinstance(elem, TimeVaryingElement):
        return get_length(elem.base)
    return float(elem.L)
