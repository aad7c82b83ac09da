"""trackpad_hip -- MI355X (gfx950) backend for TrackPad.jl's multi-particle tracking path.

Host-side mirror of the reference interface (src/TrackPad.jl:37-61 exports used by the
path): element types, Beam, Lattice, ``linepass_`` / ``ringpass_`` (Julia's ``linepass!`` /
``ringpass!``) and the device-resident ``track_``; all tracking runs in libtrackpad_hip
(hand-written HIP kernels behind a C ABI, include/trackpad_hip.h).
"""

from .elements import (  # noqa: F401
    MAXORD, C_LIGHT, AbstractCavity, AbstractDrift, AbstractElement, AbstractMagnet, Drift, ERBend,
    ExactSBend, ExactSBendLinearEdge, Marker, Octupole, Quadrupole, RBend, RFCavity, SBend, Sextupole,
    ThinMultipole, TimeVaryingElement, get_length, timed,
    AccelCavity, Corrector, CrabCavity, DriftSC, HKicker, InvLorentzBoost, LBend, LongitudinalRFMap,
    LongitudinalRLCWake, LongitudinalWake, LorentzBoost, OctupoleSC, QuadrupoleSC, RBendSC, SBendSC, SextupoleSC,
    Solenoid, SpaceCharge, StrongGaussianBeam, StrongThinGaussianBeam, Translation, VKicker, Wiggler, YRotation,
)
from .lattice import (  # noqa: F401
    M_ELECTRON, M_PROTON, Beam, Lattice, beti, findelem, materialize_lattice, spos, total_length,
)
from .table import TableInfo, compile_table  # noqa: F401
from ._lib import LIB_PATH, TrackPadHIPError  # noqa: F401
from .tracking import (  # noqa: F401
    DeviceTable, HIPContext, HIPParticles, default_context, device_table, linepass_, ringpass_, track_,
)

from .adjoint import Gradients, ringpass_grad, set_strengths, strength_positions  # noqa: F401
from .batch import BatchGradients, HIPBatch  # noqa: F401
from .madx import MadxError, read_madx, write_madx  # noqa: F401
from .optics import (  # noqa: F401
    fastfindm66, fastfindm66_refpts, find_closed_orbit_4d, find_closed_orbit_6d, findm66, findm66_refpts, getchrom,
    gettune, jacobians, one_turn_map, tune_from_map,
)

__version__ = "0.1.0"


def __getattr__(name):
    if name == "TrackLoss":
        from . import adjoint

        return adjoint.TrackLoss
    raise AttributeError(name)
