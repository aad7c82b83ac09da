"""The fast build's drift at small angles (inv_pz in trackpad.jl_amd/csrc/tp_maps.h) forms
L / pz, pz = sqrt(A - w), w = px^2 + py^2, as the quadratic Lr + w (c1 + w c2) of the
binomial series of (1 - w/A)^(-1/2) about the hoisted R = 1/sqrt(A), with Lr = L R,
c1 = Lr H, c2 = c1 H3, H = 1/(2A), H3 = 3/(4A), for w below WB = 2^-18 A.  This test reads
the constants from the source, replays the kernel's operation sequence with every multiply,
add and FMA rounded once to double (exact rationals, then one rounding), and holds the
result to 3 ulp of L / pz computed to 50 digits, for angles up to the switch bound and delta
in [-0.3, 0.3].  It also checks the constants against the series' coefficients and that the
dropped terms stay below 2^-55 at the bound.  No GPU."""

import re
from fractions import Fraction as F
from pathlib import Path

import mpmath as mp
import numpy as np

SRC = Path(__file__).resolve().parents[1] / "trackpad.jl_amd" / "csrc" / "tp_maps.h"


def _consts():
    text = SRC.read_text()
    h = float(re.search(r"c\.H = ([0-9.e+-]+) / c\.A;", text).group(1))
    h3 = float(re.search(r"c\.H3 = ([0-9.e+-]+) \* c\.H;", text).group(1))
    wb = float.fromhex(re.search(r"c\.WB = c\.A \* (0x[0-9a-fp+-]+);", text).group(1))
    body = text[text.index("TP_DEV double inv_pz("):]
    body = body[:body.index("\n}\n")]
    # the operation sequence this test replays
    assert "const double w = __builtin_fma(s.px, s.px, s.py * s.py);" in body
    assert "const double Lr = L * c.R;" in body and "const double c1 = Lr * c.H;" in body
    assert "const double c2 = c1 * c.H3;" in body
    assert "NormL = __builtin_fma(w, __builtin_fma(w, c2, c1), Lr);" in body
    assert "if (__builtin_expect(!(w < c.WB), 0)) NormL = L * rsqrt_nr(c.A - w);" in body
    return h, h3, wb


def rnd(x: F) -> F:
    return F(float(x))  # one rounding to double (round to nearest even)


def test_constants_and_truncation():
    h, h3, wb = _consts()
    # (1 - u)^(-1/2) = 1 + u/2 + 3u^2/8 + ...: c1 = L R / (2A), c2 = c1 (3 / (4A)) = 3 L R / (8 A^2)
    assert h == 0.5 and h3 == 1.5 and wb == 2.0 ** -18
    mp.mp.dps = 50
    u = mp.mpf(wb)
    tail = mp.nsum(lambda n: mp.binomial(2 * n, n) / mp.mpf(4) ** n * u ** n, [3, mp.inf])
    assert tail < mp.mpf(2) ** -55.6


def test_series_within_three_ulp():
    h, h3, wb = _consts()
    mp.mp.dps = 50
    rng = np.random.default_rng(11)
    beti = 1.0 / 0.999  # a slightly sub-relativistic beam
    worst = 0.0
    used = 0
    for i in range(3000):
        delta = float(rng.uniform(-0.3, 0.3))
        L = float(rng.choice([0.05, 0.37, 1.0, 2.5]))
        A = rnd(rnd(F(1) + rnd(rnd(F(2) * F(delta)) * F(beti))) + rnd(F(delta) * F(delta)))
        H = rnd(F(h) / A)
        H3 = rnd(F(h3) * H)
        WB = A * F(wb)  # exact: a power-of-two scaling
        R = rnd(F(1) / F(float(mp.sqrt(mp.mpf(float(A))))))  # 1/sqrt(A) with both steps rounded
        # angles up to the switch, half of them within 5 % of it
        frac = rng.uniform(0.95, 1.0) if i % 2 else rng.uniform(0.0, 1.0) ** 4
        wmax = float(WB) * frac
        phi = rng.uniform(0, 2 * np.pi)
        px, py = np.sqrt(wmax) * np.cos(phi), np.sqrt(wmax) * np.sin(phi)
        Px, Py = F(float(px)), F(float(py))
        w = rnd(Px * Px + rnd(Py * Py))  # fma(px, px, py * py)
        if not w < WB:
            continue
        used += 1
        Lr = rnd(F(L) * R)
        c1 = rnd(Lr * H)
        c2 = rnd(c1 * H3)
        normL = rnd(w * rnd(w * c2 + c1) + Lr)
        exact = mp.mpf(L) / mp.sqrt(mp.mpf(float(A)) - mp.mpf(float(Px)) ** 2 - mp.mpf(float(Py)) ** 2)
        err = abs((mp.mpf(float(normL)) - exact) / exact)
        worst = max(worst, float(err))
    assert used > 2500
    assert worst <= 3 * 2.0 ** -53, worst
    assert worst > 0.0


def test_a_plausible_mistake_fails():
    """The same replay with the w^2 coefficient 3/8 taken as 1/2 (H3 = 2/A), or with the bound
    four times wider, leaves the 3 ulp band: the test above resolves both."""
    mp.mp.dps = 50
    A, L = F(1), 1.0
    for h3, wbf in ((2.0, 2.0 ** -18), (1.5, 2.0 ** -16)):
        H = rnd(F(1, 2) / A)
        H3 = rnd(F(h3) * H)
        w = rnd(A * F(wbf) * F(99, 100))
        R = F(1)
        Lr = rnd(F(L) * R)
        c1 = rnd(Lr * H)
        c2 = rnd(c1 * H3)
        normL = rnd(w * rnd(w * c2 + c1) + Lr)
        exact = mp.mpf(L) / mp.sqrt(1 - mp.mpf(float(w)))
        assert abs((mp.mpf(float(normL)) - exact) / exact) > 3 * 2.0 ** -53
