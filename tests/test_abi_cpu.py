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


def test_time_varying_lattice_rows():
    """A TimeVaryingElement is resolved per turn into sequence rows (ringpass! resolves it in
    every linepass! call, src/lattice.jl:267, 305-317); static elements share one record."""
    import struct

    from trackpad_hip import TimeVaryingElement, materialize_lattice, timed

    seen = []

    def k1(t, n):
        seen.append((t, n))
        return 1.0 + 0.1 * n

    q = Quadrupole(0.5, 1.0)
    lat = Lattice([Drift(1.0), TimeVaryingElement(q, {"k1": k1}), Drift(1.0)])
    # no turn range: one row, the lattice at (time, turn), the same bytes as materialising
    one = compile_table(lat, Beam(3e9), time=0.25, turn=3)
    assert one == compile_table(materialize_lattice(lat, time=0.25, turn=3), Beam(3e9))
    # four rows from turn 3, time advanced by accumulation as ringpass! does
    seen.clear()
    blob = compile_table(lat, Beam(3e9), turn=3, nturns=4, time=0.5, dt_turn=0.1)
    assert _check(blob)[0] == 0
    h = struct.unpack_from("<6IIi4d", blob)
    n_seq, n_rec, n_rows, row_turn0 = h[2], h[3], h[6], h[7]
    assert (n_seq, n_rows, row_turn0) == (3, 4, 3)
    t = 0.5
    want = []
    for k in range(4):
        want.append((t, 3 + k))
        t += 0.1
    assert seen == want
    seq = struct.unpack_from("<12i", blob, 64)
    assert len({seq[0], seq[2], seq[3], seq[5], seq[6], seq[8], seq[9], seq[11]}) == 1  # the drifts
    assert len({seq[1], seq[4], seq[7], seq[10]}) == 4  # four different quadrupoles
    assert n_rec == 5
    # the k1 slot of each row's quadrupole record (poly b1 of the Yoshida block)
    recs_at = 64 + 48
    pool_at = recs_at + n_rec * 32
    for k in range(4):
        r = struct.unpack_from("<4iI3i", blob, recs_at + 32 * seq[1 + 3 * k])
        b1 = struct.unpack_from("<d", blob, pool_at + 8 * (r[3] + 6 + 2))[0]
        assert b1 == 1.0 + 0.1 * (3 + k)
    # integer fields keep their type
    tq = timed(Quadrupole(0.4, 0.6), num_int_steps=lambda t, n: 8 if t < 0.5 else 12)
    assert tq.materialize(0.2, 0).num_int_steps == 8 and tq.materialize(1.0, 5).num_int_steps == 12
    with pytest.raises(ValueError, match="no field"):
        timed(Quadrupole(0.4, 0.6), k7=1.0)

def test_wiggler_block_is_validated():
    """The wiggler's block length follows its harmonic counts; counts that run past the pool
    are refused."""
    from trackpad_hip import Wiggler

    beam = Beam(3e9)
    blob = bytearray(compile_table(Lattice([Wiggler(1.2, lw=0.2, Bmax=0.8, Nsteps=8, Bx=[1, 1, 5, 3, 4, 0])]), beam))
    rc, f, _ = _check(bytes(blob))
    assert rc == 0 and f & 128
    n_seq = struct.unpack_from("<I", blob, 8)[0]
    rec0 = 64 + ((n_seq * 4 + 7) // 8) * 8
    n_rec = struct.unpack_from("<I", blob, 12)[0]
    pool0 = rec0 + 32 * n_rec
    off = struct.unpack_from("<i", blob, rec0 + 12)[0]
    struct.pack_into("<d", blob, pool0 + 8 * (off + 3), 3.0)  # three By harmonics: past the pool
    rc, _, msg = _check(bytes(blob))
    assert rc == -5 and "pool" in msg
    struct.pack_into("<d", blob, pool0 + 8 * (off + 3), -1.0)
    rc, _, msg = _check(bytes(blob))
    assert rc == -5 and "harmonics" in msg
