"""The fast build's drift at small angles (inv_pz in trackpad.jl_amd/csrc/tp_maps.h) forms
L / pz, pz = sqrt(A - px^2 - py^2), as a cubic in v = k (px^2 + py^2) / A about the hoisted
R = 1/sqrt(A), k = sqrt(3)/2.  This test reads the constants from the source, replays the
kernel's operation sequence with every multiply, add and FMA rounded once to double (exact
rationals, then one rounding), and holds the result to 3 ulp of L / pz computed to 50 digits,
for angles up to the switch bound and delta in [-0.3, 0.3].  It also checks the constants
against their closed forms and that the bound lies below k 2^-14 (where the dropped terms of
(1 - w)^(-1/2) stay below 2^-57).  No GPU."""

import re
from fractions import Fraction as F
from pathlib import Path

import mpmath as mp
import numpy as np

SRC = Path(__file__).resolve().parents[1] / "trackpad.jl_amd" / "csrc" / "tp_maps.h"


def _consts():
    text = SRC.read_text()
    body = text[text.index("TP_DEV double inv_pz("):]
    body = body[:body.index("\n}\n")]
    t = re.search(r"__builtin_fma\(v, ([0-9.e+-]+), 0\.5\)", body).group(1)
    q1 = re.search(r"Lr \* ([0-9.e+-]+)\), Lr\)", body).group(1)
    bound = re.search(r"!\(v < ([0-9a-fx.p+-]+)\)", body).group(1)
    k = re.search(r"c\.IA = ([0-9.e+-]+) / c\.A;", text).group(1)
    return float(k), float(t), float(q1), float.fromhex(bound)


def rnd(x: F) -> F:
    return F(float(x))  # one rounding to double (round to nearest even)


def test_constants_closed_forms():
    k, q3, q1, bound = _consts()
    mp.mp.dps = 50
    kk = mp.sqrt(3) / 2
    assert k == float(kk)
    assert q1 == float(1 / (2 * kk))
    assert q3 == float(mp.mpf(5) / (16 * kk ** 3))
    assert mp.mpf(3) / (8 * kk ** 2) == mp.mpf("0.5")  # the w^2 coefficient is the inline 1/2
    assert bound <= float(kk * mp.mpf(2) ** -14) and bound > 0.99 * float(kk * mp.mpf(2) ** -14)
    # the tail after w^3 for w below the bound: 35/128 w^4 / (1 - w) < 2^-57
    w = mp.mpf(bound) / kk
    assert mp.mpf(35) / 128 * w ** 4 / (1 - w) < mp.mpf(2) ** -57


def test_series_within_three_ulp():
    k, q3, q1, bound = _consts()
    mp.mp.dps = 50
    rng = np.random.default_rng(11)
    beti = 1.0 / 0.999  # a slightly sub-relativistic beam
    worst = 0.0
    for i in range(3000):
        delta = float(rng.uniform(-0.3, 0.3))
        L = float(rng.choice([0.05, 0.37, 1.0, 2.5]))
        A = rnd(rnd(F(1) + rnd(rnd(F(2) * F(delta)) * F(beti))) + rnd(F(delta) * F(delta)))
        IA = rnd(F(k) / A)
        R = rnd(F(1) / F(float(mp.sqrt(mp.mpf(float(A))))))  # 1/sqrt(A) with both steps rounded
        # angles up to the switch, half of them within 5 % of it
        frac = rng.uniform(0.95, 1.0) if i % 2 else rng.uniform(0.0, 1.0) ** 4
        wmax = float(bound / k) * frac * float(A)
        phi = rng.uniform(0, 2 * np.pi)
        px, py = np.sqrt(wmax) * np.cos(phi), np.sqrt(wmax) * np.sin(phi)
        Px, Py = F(float(px)), F(float(py))
        v = rnd(rnd(Px * Px + rnd(Py * Py)) * IA)  # fma(px, px, py*py) * IA
        if not v < F(bound):
            continue
        Lr = rnd(F(L) * R)
        t = rnd(v * F(q3) + F(1, 2))
        inner = rnd(rnd(Lr * v) * t + rnd(Lr * F(q1)))
        normL = rnd(v * inner + Lr)
        exact = mp.mpf(L) / mp.sqrt(mp.mpf(float(A)) - mp.mpf(float(Px)) ** 2 - mp.mpf(float(Py)) ** 2)
        err = abs((mp.mpf(float(normL)) - exact) / exact)
        worst = max(worst, float(err))
    assert worst <= 3 * 2.0 ** -53, worst
    assert worst > 0.0
