"""The table blob is untrusted input at the C ABI: tp_table_check (and tp_table_create, which
parses the same way) must answer every blob with a code, never crash, and accept only blobs
whose every sequence entry, parameter block and misalignment block lies inside the table --
the bounds the device kernels rely on when they read a record's block.  Checked on seeded
mutations (random 32-bit words, sign flips, huge values, doubles set to NaN / inf) of the
blobs of lattices covering every table tag (no GPU)."""

import ctypes as C
import struct

import numpy as np

from trackpad_hip import (AccelCavity, Beam, Corrector, CrabCavity, Drift, ExactSBend, ExactSBendLinearEdge,
                          InvLorentzBoost, Lattice, LBend, LongitudinalRFMap, LongitudinalRLCWake, LongitudinalWake,
                          LorentzBoost, Quadrupole, RFCavity, SBend, SBendSC, Sextupole, Solenoid, StrongGaussianBeam,
                          StrongThinGaussianBeam, ThinMultipole, Translation, Wiggler, YRotation, compile_table, timed)
from trackpad_hip import _lib as L

HDR = struct.Struct("<6IIi4d")
REC = struct.Struct("<4iI3i")
M_E = 0.51099895069e6

# Parameter block length per tag (include/trackpad_table.h), m = max_order, n = nsteps
POLY = lambda m: 2 * (m + 1)
BLOCK = {0: lambda m, n: 0, 1: lambda m, n: 2, 2: lambda m, n: 6 + POLY(m), 3: lambda m, n: POLY(m),
         4: lambda m, n: 17 + POLY(m), 5: lambda m, n: 24 + POLY(m), 6: lambda m, n: 24 + POLY(m),
         7: lambda m, n: 7, 8: lambda m, n: 2, 9: lambda m, n: 6, 10: lambda m, n: 3, 11: lambda m, n: 3,
         12: lambda m, n: 6, 13: lambda m, n: 4, 14: lambda m, n: 1, 15: lambda m, n: 2, 16: lambda m, n: 2,
         17: lambda m, n: 3 + 3 * n, 18: lambda m, n: 6, 19: lambda m, n: 3 + 2 * n, 20: lambda m, n: 13,
         21: lambda m, n: 0}


def _lattices():
    beam = Beam(3e9)
    rf = RFCavity(0.0, 1e6, 5e8, energy=3e9)
    els = [Drift(1.0), Quadrupole(0.4, 1.2, t1=[1e-4, 0, 0, 0, 0, 0], r1=np.eye(6)), Sextupole(0.2, 10.0),
           ThinMultipole(0.0, [0.0] * 4, [0.0, 0.0, 1.0, 0.0], max_order=2),
           SBend(0.8, 0.05, 0.01, 0.02, fint1=0.5, gap=0.03),
           ExactSBend(1.0, 0.1, 0.03, 0.02, gk=0.01), ExactSBendLinearEdge(1.0, 0.1, 0.01, 0.01), rf,
           Solenoid(0.4, 0.3), Corrector(0.2, 1e-5, -2e-5), Translation(0.0, dx=1e-4, ds=1e-3),
           YRotation(0.0, angle=1e-3), CrabCavity(0.3, volt=1e5, freq=5e8, energy=3e9),
           AccelCavity(0.0, volt=1e6, freq=5e8, h=400.0, energy=3e9),
           LongitudinalRFMap(1e-3, AccelCavity(0.0, volt=1e6, freq=5e8, h=400.0)), LorentzBoost(0.01),
           InvLorentzBoost(0.01), StrongThinGaussianBeam(1e-7, 1e-3, 1e-3),
           StrongGaussianBeam(3e-3, M_E, 1.0, 10 ** 6, 3e9, [1e-3, 1.2e-3], nzslice=3),
           LongitudinalRLCWake(freq=1e9, Rshunt=1e4, Q0=2.0, scale=1e-20),
           LongitudinalWake([0.0, 1e-12, 2e-12], [1.0, 2.0, 0.5], scale=1e-7), LBend(0.8, 0.05, K=0.1),
           SBendSC(0.0, 0.1), Wiggler(0.4, lw=0.2, Bmax=0.8, Nsteps=4, energy=3e9, rad=1)]
    yield compile_table(Lattice(els), beam)
    q = Quadrupole(0.4, 1.0)
    tv = Lattice([Drift(1.0), timed(q, k1=lambda t, n: 1.0 + 0.1 * n), Drift(1.0)])
    yield compile_table(tv, beam, turn=3, nturns=4)


def _check(blob):
    # an allocation of exactly the blob's size, so that a read past its end leaves the
    # allocation (the sanitizer run of these tests sees it: PYTHONMALLOC=malloc there)
    buf = (C.c_char * len(blob)).from_buffer_copy(bytes(blob))
    f = C.c_uint32()
    return L.lib().tp_table_check(buf, len(blob), C.byref(f))


def _accepted_blob_in_bounds(blob):
    h = HDR.unpack_from(blob, 0)
    n_seq, n_rec, pool_len, flags, n_rows = h[2], h[3], h[4], h[5], h[6]
    rows = n_rows if n_rows > 1 else 1
    seq = np.frombuffer(blob, np.int32, rows * n_seq, HDR.size)
    assert np.all((seq >= 0) & (seq < n_rec))
    seq_bytes = (rows * n_seq * 4 + 7) // 8 * 8
    pool = np.frombuffer(blob, np.float64, pool_len, HDR.size + seq_bytes + n_rec * REC.size)
    for i in range(n_rec):
        tag, n, m, off, fl, moff, gorder, _ = REC.unpack_from(blob, HDR.size + seq_bytes + i * REC.size)
        assert 0 <= m <= 10 and 0 <= n and -1 <= gorder <= m
        if tag == 22:  # wiggler: 9 + 11 per harmonic, counts from the block
            assert 0 <= off and off + 9 <= pool_len
            length = 9 + 11 * (int(pool[off + 3]) + int(pool[off + 4]))
        else:
            length = BLOCK[tag](m, n)
        assert 0 <= off and off + length <= pool_len, (tag, off, length, pool_len)
        if fl & 0xF:
            assert 0 <= moff and moff + 84 <= pool_len


def test_table_check_survives_mutations():
    rng = np.random.default_rng(2024)
    accepted = rejected = 0
    for good in _lattices():
        assert _check(good) == 0
        _accepted_blob_in_bounds(good)
        words = len(good) // 4
        for trial in range(2000):
            blob = bytearray(good)
            for _ in range(int(rng.integers(1, 4))):
                w = int(rng.integers(0, words))
                kind = int(rng.integers(0, 5))
                if kind == 0:
                    v = int(rng.integers(-2 ** 31, 2 ** 31))
                elif kind == 1:
                    v = -struct.unpack_from("<i", blob, 4 * w)[0]
                elif kind == 2:
                    v = int(rng.choice([0x7fffffff, -0x80000000, -1, 0, 1, 2 ** 20]))
                elif kind == 3:
                    v = struct.unpack_from("<i", blob, 4 * w)[0] + int(rng.integers(-3, 4))
                else:  # a double set to NaN / +-inf / huge, when the word starts one
                    w -= w % 2
                    d = float(rng.choice([np.nan, np.inf, -np.inf, 1e300, -1.0, 64.0]))
                    struct.pack_into("<d", blob, 4 * w, d)
                    continue
                struct.pack_into("<i", blob, 4 * w, (v + 2 ** 31) % 2 ** 32 - 2 ** 31)
            rc = _check(blob)
            assert rc <= 0
            if rc == 0:
                accepted += 1
                _accepted_blob_in_bounds(blob)
            else:
                rejected += 1
    assert accepted > 100 and rejected > 100, (accepted, rejected)


def test_table_check_truncations_and_garbage():
    rng = np.random.default_rng(7)
    for good in _lattices():
        for cut in list(range(0, HDR.size + 16)) + list(rng.integers(0, len(good), 50)):
            assert _check(good[:int(cut)]) < 0
        for _ in range(50):
            junk = bytearray(rng.integers(0, 256, len(good), dtype=np.uint8).tobytes())
            junk[:HDR.size] = good[:HDR.size]  # a valid header over random contents
            rc = _check(junk)
            assert rc <= 0
            if rc == 0:
                _accepted_blob_in_bounds(junk)
