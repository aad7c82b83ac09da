"""ctypes binding of libtrackpad_hip.so (include/trackpad_hip.h).

The shared library is built in-tree (``make -C trackpad.jl_amd/csrc``) into
``trackpad.jl_amd/lib/libtrackpad_hip.so``.  There is no fallback: if the library is
missing or cannot be loaded, every device entry point raises.
"""

from __future__ import annotations

import ctypes as C
import os
import threading

__all__ = ["lib", "TrackPadHIPError", "check", "LIB_PATH", "TP_TRACK_FAST", "TP_TRACK_ASYNC",
           "TP_TRACK_TIMED", "TimingT", "LaunchCfg", "TP_TUNE_ADJOINT", "TP_DRIFT_AUTO", "TP_DRIFT_SERIES",
           "TP_DRIFT_RSQRT", "TP_NOALLOC_MAX_N"]

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("TRACKPAD_HIP_LIB", os.path.join(os.path.dirname(_HERE), "lib", "libtrackpad_hip.so"))

TP_TRACK_FAST = 0x1
TP_TRACK_ASYNC = 0x2
TP_TRACK_TIMED = 0x4

TP_DRIFT_AUTO = 0
TP_DRIFT_SERIES = 1
TP_DRIFT_RSQRT = 2

TP_NOALLOC_MAX_N = 1048576

_STATUS = {0: "TP_OK", -1: "TP_EINVAL", -2: "TP_EUNSUPPORTED", -3: "TP_EHIP", -4: "TP_ENOMEM",
           -5: "TP_ETABLE"}


class TrackPadHIPError(RuntimeError):
    """A libtrackpad_hip call failed; ``code`` is the TP_E* status."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"{_STATUS.get(code, code)}: {msg}")
        self.code = code


class LaunchCfg(C.Structure):
    _fields_ = [("block", C.c_int32), ("ppl", C.c_int32), ("turns_per_launch", C.c_int32),
                ("fwd_waves", C.c_int32), ("back_waves", C.c_int32), ("candidates", C.c_int32),
                ("rate", C.c_double)]


TP_TUNE_ADJOINT = 0x100


class TimingT(C.Structure):
    _fields_ = [("kernel_ms", C.c_double), ("launches", C.c_int32), ("block", C.c_int32),
                ("ppl", C.c_int32), ("turns_per_launch", C.c_int32), ("forward_ms", C.c_double),
                ("backward_ms", C.c_double)]


_lock = threading.Lock()
_lib = None

_P = C.c_void_p
_PD = C.POINTER(C.c_double)
_PI64 = C.POINTER(C.c_int64)
_PI32 = C.POINTER(C.c_int32)

_SIGS = {
    "tp_version": ([], C.c_int),
    "tp_last_error": ([], C.c_char_p),
    "tp_device_count": ([C.POINTER(C.c_int)], C.c_int),
    "tp_ctx_create": ([C.c_int, C.POINTER(_P)], C.c_int),
    "tp_ctx_destroy": ([_P], C.c_int),
    "tp_set_launch": ([_P, C.c_int, C.c_int, C.c_int], C.c_int),
    "tp_timing": ([_P, C.POINTER(TimingT)], C.c_int),
    "tp_table_create": ([_P, C.c_void_p, C.c_size_t, C.POINTER(_P)], C.c_int),
    "tp_table_destroy": ([_P], C.c_int),
    "tp_table_check": ([C.c_void_p, C.c_size_t, C.POINTER(C.c_uint32)], C.c_int),
    "tp_particles_copy": ([_P, _P], C.c_int),
    "tp_table_info": ([_P, _PI32, _PI32, _PD], C.c_int),
    "tp_particles_create": ([_P, C.c_int64, C.POINTER(_P)], C.c_int),
    "tp_particles_destroy": ([_P], C.c_int),
    "tp_particles_upload": ([_P, _PD, _PI64, _PI32, _PI32], C.c_int),
    "tp_particles_download": ([_P, _PD, _PI64, _PI32, _PI32], C.c_int),
    "tp_particles_device_ptrs": ([_P, C.POINTER(_PD), C.POINTER(_PI32), C.POINTER(_PI32), _PI64], C.c_int),
    "tp_track": ([_P, _P, _P, C.c_int32, C.c_int32, C.c_uint32], C.c_int),
    "tp_track_noalloc": ([_P, _P, _P, C.c_int32, C.c_int32, C.c_uint32], C.c_int),
    "tp_sync": ([_P], C.c_int),
    "tp_adjoint": ([_P, _P, _P, C.c_int32, C.c_int32, C.c_int32, C.c_uint32, _PD, _PI64, _PD, _PD], C.c_int),
    "tp_jacobian": ([_P, _P, _P, C.c_int32, C.c_int32, C.c_int32, C.c_uint32, _PD], C.c_int),
    "tp_track_host": ([_P, _P, _PD, C.c_int64, _PI64, _PI32, _PI32, C.c_int32, C.c_int32, C.c_uint32], C.c_int),
    "tp_batch_create": ([_P, C.c_int32, C.POINTER(_P), C.POINTER(_P), C.POINTER(_P)], C.c_int),
    "tp_batch_destroy": ([_P], C.c_int),
    "tp_batch_set_launch": ([_P, C.c_int, C.c_int, C.c_int], C.c_int),
    "tp_set_waves": ([_P, C.c_int, C.c_int], C.c_int),
    "tp_set_drift": ([_P, C.c_int], C.c_int),
    "tp_autotune": ([_P, _P, _P, C.c_int32, C.c_uint32, C.c_double, C.c_char_p, C.c_void_p], C.c_int),
    "tp_batch_autotune": ([_P, _P, _PI32, C.c_int32, C.c_uint32, C.c_char_p, C.c_void_p], C.c_int),
    "tp_batch_track": ([_P, _P, _PI32, _PI32, C.c_uint32], C.c_int),
    "tp_batch_adjoint": ([_P, _P, _PI32, _PI32, C.c_int32, C.c_uint32, _PD, _PI64, _PD, _PD], C.c_int),
    "tp_batch_info": ([_P, _PI32, _PI32, _PI32, _PI64], C.c_int),
}


def lib():
    """The loaded library (raises if it is not built)."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise TrackPadHIPError(
                    -2, f"{LIB_PATH} is not built; run `make -C trackpad.jl_amd/csrc` "
                        "(or __graft_entry__.build()) first")
            L = C.CDLL(LIB_PATH)
            for name, (args, res) in _SIGS.items():
                f = getattr(L, name)
                f.argtypes = args
                f.restype = res
            _lib = L
        return _lib


def check(rc: int) -> None:
    if rc != 0:
        raise TrackPadHIPError(rc, lib().tp_last_error().decode(errors="replace"))
