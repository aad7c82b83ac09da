"""The fast build's two derived rotations of the exact bend body (src/tracking.jl:858-894),
replayed on the CPU with every multiply, add and FMA rounded once to double:

* rot_body (trackpad.jl_amd/csrc/tp_maps.h): a pure-dipole bend's 4n sub-body rotations
  (the record's cos/sin of the sub-step angles L1 irho, L2 irho) composed into one, the step's
  rotation raised to the n-th power by repeated squaring.  Held to a few ulp of the exact
  product of the same 4n rotation matrices, computed to 50 digits.
* the exact-body pair of bends with kicks (esbend_body in tp_track_kernels.hip): the doubled
  sub-step angle's cosine 1 - 2 s1^2 and sine 2 c1 s1, held to 1 ulp of cos / sin of twice
  the record's angle.  The form c1^2 - s1^2, which carries c1's rounding twice, is shown to
  miss that bound for some angles.

The tests read the operation sequences from the sources, so a change of form there fails here
until it is replayed.  No GPU."""

import math
import struct
from fractions import Fraction as F
from pathlib import Path

import mpmath as mp

from trackpad_hip import Beam, compile_table
from trackpad_hip import workloads as W

CSRC = Path(__file__).resolve().parents[1] / "trackpad.jl_amd" / "csrc"

mp.mp.dps = 50


def rnd(x: F) -> F:
    return F(float(x))  # one rounding to double (round to nearest even)


def fma(a, b, c) -> F:
    return rnd(F(a) * F(b) + F(c))


def mul(a, b) -> F:
    return rnd(F(a) * F(b))


def rot_mul(u, v):
    """tp_maps.h rot_mul: (fma(u.c, v.c, -(u.s v.s)), fma(u.s, v.c, u.c v.s))."""
    return fma(u[0], v[0], -mul(u[1], v[1])), fma(u[1], v[0], mul(u[0], v[1]))


def rot_mul_slip(u, v):
    """A plausible slip: the sines' product added in the cosine."""
    return fma(u[0], v[0], mul(u[1], v[1])), fma(u[1], v[0], mul(u[0], v[1]))


def rot_body(c1, s1, c2, s2, ns, mulf=rot_mul):
    h = mulf((c1, s1), (c2, s2))
    st = mulf(h, h)
    r = (F(1), F(0))
    n = ns
    while n > 0:
        if n & 1:
            r = mulf(r, st)
        if n > 1:
            st = mulf(st, st)
        n >>= 1
    return r


def _sources_match():
    maps = (CSRC / "tp_maps.h").read_text()
    assert "return Rot{__builtin_fma(u.c, v.c, -(u.s * v.s)), __builtin_fma(u.s, v.c, u.c * v.s)};" in maps
    body = maps[maps.index("TP_DEV Rot rot_body("):]
    body = body[:body.index("\n}\n")]
    for line in ("const Rot h = rot_mul(Rot{c1, s1}, Rot{c2, s2});", "Rot st = rot_mul(h, h);",
                 "Rot r{1.0, 0.0};", "for (int n = ns; n > 0; n >>= 1) {", "if (n & 1) r = rot_mul(r, st);",
                 "if (n > 1) st = rot_mul(st, st);"):
        assert line in body, line
    trk = (CSRC / "tp_track_kernels.hip").read_text()
    assert "const double c11 = __builtin_fma(-2.0 * s1, s1, 1.0), s11 = 2.0 * (c1 * s1);" in trk
    assert "const Rot R = rot_body(c1, s1, c2, s2, nsteps);" in trk


def _bend_record(lat, beam):
    """irho, L1, L2, c1, s1, c2, s2 and nsteps of the first exact bend of a lattice's table."""
    blob = compile_table(lat, beam)
    h = struct.unpack_from("<6IIi4d", blob)
    n_seq, n_rec, pool_len, n_rows = h[2], h[3], h[4], max(h[6], 1)
    pool_at = len(blob) - 8 * pool_len
    recs_at = pool_at - 32 * n_rec
    assert recs_at >= 64 + 4 * n_seq * n_rows
    for k in range(n_rec):
        tag, nsteps, _m, off, _fl, _mo, _g, _kick = struct.unpack_from("<4iI3i", blob, recs_at + 32 * k)
        if tag == 6:  # TP_ESBEND_LE
            p = struct.unpack_from("<11d", blob, pool_at + 8 * off)
            return p[0], p[1], p[2], p[7], p[8], p[9], p[10], nsteps
    raise AssertionError("no exact bend in the table")


def _exact_product(c1, s1, c2, s2, ns):
    """The product of the 4 ns sub-body rotation matrices of the given doubles, to 50 digits."""
    r = mp.matrix([[1, 0], [0, 1]])
    m1 = mp.matrix([[mp.mpf(c1), mp.mpf(s1)], [-mp.mpf(s1), mp.mpf(c1)]])
    m2 = mp.matrix([[mp.mpf(c2), mp.mpf(s2)], [-mp.mpf(s2), mp.mpf(c2)]])
    step = m1 * m2 * m2 * m1
    for _ in range(ns):
        r = step * r
    return r[0, 0], r[0, 1]


def test_rot_body_is_the_product_of_the_sub_body_rotations():
    _sources_match()
    beam = Beam(W.ENERGY)
    cases = [_bend_record(W.dba_ring(beam=beam), beam), _bend_record(W.da_ring(ncell=80, beam=beam), beam)]
    for irho, L1, L2, c1, s1, c2, s2, ns in cases:
        assert (c1, s1, c2, s2) == (math.cos(irho * L1), math.sin(irho * L1), math.cos(irho * L2),
                                    math.sin(irho * L2))
        for n in (1, 2, 3, 7, ns):
            cr, sr = rot_body(c1, s1, c2, s2, n)
            ce, se = _exact_product(c1, s1, c2, s2, n)
            # each composition rounds both components once: 2 for the step, then at most two
            # per bit of n (a square and a product), each about half an ulp
            k = 3 + 2 * n.bit_length()
            ulp_c, ulp_s = math.ulp(float(ce)), math.ulp(float(se))
            assert abs(mp.mpf(cr.numerator) / cr.denominator - ce) <= k * ulp_c, (n, float(ce))
            assert abs(mp.mpf(sr.numerator) / sr.denominator - se) <= k * ulp_s, (n, float(se))
        # the whole body's angle: the product is the rotation by n (2 L1 + 2 L2) irho
        ce, se = _exact_product(c1, s1, c2, s2, ns)
        theta = mp.mpf(irho) * ns * (2 * mp.mpf(L1) + 2 * mp.mpf(L2))
        assert abs(ce - mp.cos(theta)) < 1e-14 and abs(se - mp.sin(theta)) < 1e-14
        # a plausible slip in the composition fails the same bound
        cb, _ = rot_body(c1, s1, c2, s2, ns, rot_mul_slip)
        assert abs(mp.mpf(cb.numerator) / cb.denominator - ce) > 1e3 * math.ulp(float(ce))


def test_doubled_sub_step_angle_of_the_body_pairs():
    _sources_match()
    misses_of_the_other_form = 0
    for k in range(400):
        phi = 1e-4 * (1.0 + 0.37 * k) ** 1.5  # 1e-4 .. ~0.3 rad
        c1, s1 = math.cos(phi), math.sin(phi)
        c11 = fma(-2.0 * s1, s1, 1.0)
        s11 = rnd(2 * mul(c1, s1))
        want_c, want_s = mp.cos(2 * mp.mpf(phi)), mp.sin(2 * mp.mpf(phi))
        assert abs(mp.mpf(float(c11)) - want_c) <= math.ulp(float(want_c)), phi
        assert abs(mp.mpf(float(s11)) - want_s) <= 2 * math.ulp(float(want_s)), phi
        other = fma(c1, c1, -mul(s1, s1))
        if abs(mp.mpf(float(other)) - want_c) > math.ulp(float(want_c)):
            misses_of_the_other_form += 1
    assert misses_of_the_other_form > 0
