"""C ABI checks that need no GPU: the library loads, exports every entry point that
include/*.h declares, validates tables without a device and reports errors as codes."""

import ctypes as C
import os
import re
import struct

import numpy as np
import pytest

from trackpad_hip import Beam, Drift, Lattice, Quadrupole, compile_table
from trackpad_hip import _lib as L
from trackpad_hip import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include")


def declared_functions():
    names = set()
    for fn in os.listdir(INCLUDE):
        if not fn.endswith(".h"):
            continue
        src = open(os.path.join(INCLUDE, fn)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        names |= set(re.findall(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(tp_\w+)\s*\(", src, flags=re.M))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = L.lib()
    names = declared_functions()
    assert len(names) >= 18, names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert lib.tp_version() == 1


def _check(blob):
    f = C.c_uint32()
    rc = L.lib().tp_table_check(blob, len(blob), C.byref(f))
    return rc, f.value, L.lib().tp_last_error().decode()


def test_table_check_features():
    beam = Beam(3e9)
    rc, f, _ = _check(compile_table(W.fodo_lattice(), beam))
    assert rc == 0 and f == 0  # drift / quad kernel variant
    rc, f, _ = _check(compile_table(W.dba_ring(beam=beam), beam))
    assert rc == 0 and f == (4 | 8)  # exact bend with linear edges + RF
    rc, f, _ = _check(compile_table(W.da_ring(ncell=2, beam=beam), beam))
    assert rc == 0 and f & 32  # general order-10 kicks


def test_table_check_rejects_malformed():
    beam = Beam(3e9)
    good = bytearray(compile_table(Lattice([Drift(1.0), Quadrupole(0.5, 1.0)]), beam))
    assert _check(bytes(good))[0] == 0
    bad = bytearray(good)
    bad[0:4] = b"XXXX"
    assert _check(bytes(bad))[0] == -5
    assert _check(bytes(good[:-8]))[0] == -5            # truncated
    bad = bytearray(good)
    struct.pack_into("<i", bad, 64, 7)                  # seq[0] out of range
    rc, _, msg = _check(bytes(bad))
    assert rc == -5 and "seq" in msg
    n_seq = struct.unpack_from("<I", good, 8)[0]
    rec0 = 64 + ((n_seq * 4 + 7) // 8) * 8
    bad = bytearray(good)
    struct.pack_into("<i", bad, rec0, 42)               # unknown tag
    assert _check(bytes(bad))[0] == -2
    bad = bytearray(good)
    struct.pack_into("<i", bad, rec0 + 12, 10 ** 6)     # parameter block outside the pool
    assert _check(bytes(bad))[0] == -5
    bad = bytearray(good)
    struct.pack_into("<d", bad, 32, 0.5)                # beti < 1
    assert _check(bytes(bad))[0] == -5


def test_errors_are_codes_without_a_device():
    lib = L.lib()
    h = C.c_void_p()
    n = C.c_int(-1)
    rc = lib.tp_device_count(C.byref(n))
    if rc == 0 and n.value > 0:
        pytest.skip("a GPU is visible; the no-device error path is not reachable here")
    rc = lib.tp_ctx_create(0, C.byref(h))
    assert rc in (-1, -3) and h.value is None
    assert lib.tp_last_error()
    assert lib.tp_track(None, None, None, 0, 1, 0) == -1
    assert lib.tp_set_launch(None, 256, 1, 0) == -1


def test_time_varying_element_is_refused():
    from trackpad_hip import TimeVaryingElement, materialize_lattice

    lat = Lattice([Drift(1.0), TimeVaryingElement(Quadrupole(0.5, 1.0), {"k1": lambda t, n: 1.0 + 0.1 * n})])
    with pytest.raises(ValueError, match="materialize"):
        compile_table(lat, Beam(3e9))
    m = materialize_lattice(lat, turn=3)
    assert m[1].k1 == pytest.approx(1.3)
    assert _check(compile_table(m, Beam(3e9)))[0] == 0
