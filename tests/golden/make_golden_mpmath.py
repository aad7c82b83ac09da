"""Write tests/golden/mpmath_golden.json: 50-digit values that pin the CPU oracle.

Run from the repository root:  python tests/golden/make_golden_mpmath.py

Each case maps the reference tests' 10 x 6 input (test/verify_fodo.jl:7-18) through one
element or one of the reference test lines, computed in 50-digit arithmetic from
formulations that are independent of the oracle's code where the mathematics allows:

* thin multipole kicks as Re/Im of  sum_n (b_n + i a_n) (x + i y)^n  with complex powers
  (the oracle and the reference use the Horner recurrence, src/tracking.jl:290-298);
* the exact sector-bend body by circle-arc geometry: intersect the particle's circle of
  radius rho*d (d = sqrt(dp1^2 - py^2)) with the exit plane, read off the exit position,
  angle and arc length (the reference uses Forest's closed form, src/tracking.jl:858-894);
* the drift as a straight line x + L tan(theta) (src/tracking.jl:87-99);
* the Yoshida coefficients' closed form (Yoshida 1990), d1 = 1/(2(2 - 2^(1/3))),
  d2 = (1 - 2^(1/3))/(2(2 - 2^(1/3))), k1 = 1/(2 - 2^(1/3)), k2 = -2^(1/3)/(2 - 2^(1/3)),
  is recorded for the literal check; the step lengths use the reference's literals;
* the linear edge maps and the RF kick with 50-digit tan / sin.

Element constants that the reference computes in double before tracking (SL = L/n and
the step lengths, irho = angle/L, edge tangents) are formed in double here too, so the
fixture restates what the reference computes, not a different rounding of the inputs.
Forest's fringe (bend_fringe) has no independent form here: it is covered by the 50-digit
restatement of the same formula in ``forest_*`` cases (and by symplecticity and its
derivative identity in tests/test_oracle_pinning.py).  yrot and bend_edge, with the fringe
off, are pinned by the plane geometry of a hard-edge magnet with rotated faces
(``wedge_geometry``, the ``ExactSBendWedge*`` cases).
"""

import json
import math
import os

import mpmath as mp

mp.mp.dps = 50
HERE = os.path.dirname(os.path.abspath(__file__))

P10 = [
    [1.0e-4, 2.0e-4, 3.0e-4, -1.0e-4, 5.0e-5, 2.0e-4],
    [-2.2e-4, 1.7e-4, -1.1e-4, 2.3e-4, -7.0e-5, 1.0e-4],
    [3.5e-4, -2.1e-4, 8.0e-5, -1.8e-4, 1.4e-4, -2.6e-4],
    [-4.0e-4, 2.9e-4, 1.6e-4, 9.0e-5, -1.2e-4, 3.1e-4],
    [5.2e-4, -3.3e-4, -2.4e-4, 1.1e-4, 2.6e-4, -3.7e-4],
    [-6.1e-4, 4.4e-4, 2.7e-4, -2.5e-4, -3.0e-4, 4.5e-4],
    [7.0e-4, -5.2e-4, -3.1e-4, 3.4e-4, 3.3e-4, -5.0e-4],
    [-8.0e-4, 6.0e-4, 3.9e-4, -4.2e-4, -3.8e-4, 5.8e-4],
    [9.1e-4, -6.7e-4, -4.6e-4, 5.1e-4, 4.4e-4, -6.3e-4],
    [-9.8e-4, 7.3e-4, 5.3e-4, -5.9e-4, -4.9e-4, 7.1e-4],
]
E_REF = 3.5e9
M_E = 0.51099895069e6
C_LIGHT = 2.99792458e8


def beam_beti(E, m=M_E):
    """Beam(E): gamma = (E + m)/m, beta = sqrt(1 - 1/gamma^2), in double (src/lattice.jl:50-61)."""
    g = (E + m) / m
    b = math.sqrt(1.0 - 1.0 / (g * g))
    return 1.0 / b


def yoshida_closed():
    c = mp.cbrt(2)
    d1 = 1 / (2 * (2 - c))
    d2 = (1 - c) / (2 * (2 - c))
    k1 = 1 / (2 - c)
    k2 = -c / (2 - c)
    return [float(d1), float(d2), float(k1), float(k2)]


# The reference's decimal literals (src/tracking.jl:29-32) as doubles; they lie 2-6 ulp
# from the closed form above (checked in tests/test_oracle_pinning.py), and the step
# lengths of the reference are formed from the literals, so the fixture uses them too.
D1, D2, K1, K2 = 0.6756035959798286638, -0.1756035959798286639, 1.351207191959657328, -1.702414383919314656


def mpv(r):
    return [mp.mpf(v) for v in r]


def drift(r, L, bt):
    x, px, y, py, z, d = r
    P2 = 1 + 2 * d * bt + d * d           # total momentum^2 in units of p0
    pz = mp.sqrt(P2 - px * px - py * py)
    tx, ty = px / pz, py / pz             # straight line: dx/ds = tan(theta_x)
    return [x + L * tx, px, y + L * ty, py, z + L * (bt + d) / pz - L * bt, d]


def field_sum(a, b, m, x, y):
    """sum_{n<=m} (b_n + i a_n) (x + i y)^n by complex powers."""
    w = mp.mpc(x, y)
    s = mp.mpc(0)
    for n in range(m + 1):
        s += mp.mpc(b[n], a[n]) * w ** n
    return s.real, s.imag


def kick(r, a, b, m, K):
    x, px, y, py, z, d = r
    re, im = field_sum(a, b, m, x, y)
    return [x, px - K * re, y, py + K * im, z, d]


def bkick(r, a, b, m, K, irho, bt):
    x, px, y, py, z, d = r
    re, im = field_sum(a, b, m, x, y)
    return [x, px - K * (re - (d - x * irho) * irho), y, py + K * im, z + K * irho * x * bt, d]


def steps(L, n):
    SL = L / n
    return SL * D1, SL * D2, SL * K1, SL * K2


def yoshida(r, L, n, bt, kfun):
    L1, L2, k1, k2 = steps(L, n)
    for _ in range(n):
        r = drift(r, L1, bt); r = kfun(r, k1)
        r = drift(r, L2, bt); r = kfun(r, k2)
        r = drift(r, L2, bt); r = kfun(r, k1)
        r = drift(r, L1, bt)
    return r


def body_geometry(r, irho, L, bt):
    """Exact dipole body by circle geometry (pure dipole, |irho| >= 1e-6)."""
    x, px, y, py, z, d = r
    rho = 1 / mp.mpf(irho)
    dp1 = bt + d
    dd = mp.sqrt(dp1 * dp1 - py * py)      # horizontal momentum
    pz = mp.sqrt(dd * dd - px * px)
    R = dd * rho                            # particle's orbit radius
    th = mp.mpf(irho) * mp.mpf(L)           # bend angle of this body piece
    # entrance point and unit direction in the (radial X, tangential Y) frame of the entrance
    A = (rho + x, mp.mpf(0))
    u0 = (px / dd, pz / dd)
    C = (A[0] - R * u0[1], A[1] + R * u0[0])  # centre: left normal of the motion
    e = (mp.cos(th), mp.sin(th))
    ce = C[0] * e[0] + C[1] * e[1]
    cc = C[0] * C[0] + C[1] * C[1]
    t = ce + mp.sqrt(ce * ce - cc + R * R)
    Q = (t * e[0], t * e[1])
    u1 = (-(Q[1] - C[1]) / R, (Q[0] - C[0]) / R)  # counter-clockwise motion about C
    px1 = dd * (u1[0] * e[0] + u1[1] * e[1])
    # swept angle about C from A to Q (counter-clockwise)
    a0 = mp.atan2(A[1] - C[1], A[0] - C[0])
    a1 = mp.atan2(Q[1] - C[1], Q[0] - C[0])
    dphi = a1 - a0
    while dphi < -mp.pi:
        dphi += 2 * mp.pi
    while dphi > mp.pi:
        dphi -= 2 * mp.pi
    s_over_d = R * dphi / dd                # horizontal arc length / d
    return [t - rho, px1, y + py * s_over_d, py, z + dp1 * s_over_d - mp.mpf(L) * bt, d]


def _cross(a, b):
    return a[0] * b[1] - a[1] * b[0]


def wedge_geometry(r, irho, L, e1, e2, bt):
    """Hard-edge sector dipole whose entrance and exit faces are rotated by e1 and e2 (pure
    dipole, irho > 0), by plane geometry, from the entrance reference plane to the exit
    reference plane: a straight line to the entrance face, the particle's circle from there to
    the exit face, a straight line to the exit reference plane.  The frame is the one of
    ``body_geometry``: bend centre at the origin, reference circle of radius rho, entrance
    reference point (rho, 0) moving along +Y, x outward.  The entrance face passes through
    the reference point at angle e1 to the radial line (outer particles reach it later for
    e1 > 0, the rectangular-bend sense: e1 = e2 = angle/2 makes the faces parallel), the exit
    face at angle angle - e2.  The reference computes the same hard-edge magnet as Forest's
    yrot(e1), bend_edge(-e1), body, bend_edge(-e2), yrot(e2) (src/tracking.jl:901-973 with
    the fringe off)."""
    x, px, y, py, z, d = r
    rho = 1 / mp.mpf(irho)
    th = mp.mpf(irho) * mp.mpf(L)
    dp1 = bt + d
    dd = mp.sqrt(dp1 * dp1 - py * py)
    pz = mp.sqrt(dd * dd - px * px)
    R = dd * rho
    P0 = (rho, mp.mpf(0))
    P1 = (rho * mp.cos(th), rho * mp.sin(th))
    er1 = (mp.cos(th), mp.sin(th))
    t1 = (-mp.sin(th), mp.cos(th))
    f0 = (mp.cos(mp.mpf(e1)), mp.sin(mp.mpf(e1)))
    f1 = (mp.cos(th - mp.mpf(e2)), mp.sin(th - mp.mpf(e2)))
    # straight line from the entrance reference plane to the entrance face
    A = (rho + x, mp.mpf(0))
    u0 = (px / dd, pz / dd)
    s0 = _cross((P0[0] - A[0], P0[1] - A[1]), f0) / _cross(u0, f0)
    Q0 = (A[0] + s0 * u0[0], A[1] + s0 * u0[1])
    # circle through the field to the exit face (the crossing nearest the reference exit)
    C = (Q0[0] - R * u0[1], Q0[1] + R * u0[0])
    b = f1[0] * (P1[0] - C[0]) + f1[1] * (P1[1] - C[1])
    c = (P1[0] - C[0]) ** 2 + (P1[1] - C[1]) ** 2 - R * R
    disc = mp.sqrt(b * b - c)
    t = min((-b + disc, -b - disc), key=abs)
    Q1 = (P1[0] + t * f1[0], P1[1] + t * f1[1])
    a0 = mp.atan2(Q0[1] - C[1], Q0[0] - C[0])
    a1 = mp.atan2(Q1[1] - C[1], Q1[0] - C[0])
    dphi = a1 - a0
    while dphi < -mp.pi:
        dphi += 2 * mp.pi
    while dphi > mp.pi:
        dphi -= 2 * mp.pi
    u1 = (-(Q1[1] - C[1]) / R, (Q1[0] - C[0]) / R)
    # straight line from the exit face to the exit reference plane
    s2 = -((Q1[0] - P1[0]) * t1[0] + (Q1[1] - P1[1]) * t1[1]) / (u1[0] * t1[0] + u1[1] * t1[1])
    Q2 = (Q1[0] + s2 * u1[0], Q1[1] + s2 * u1[1])
    S = s0 + R * dphi + s2                  # horizontal path length
    x1 = (Q2[0] - P1[0]) * er1[0] + (Q2[1] - P1[1]) * er1[1]
    px1 = dd * (u1[0] * er1[0] + u1[1] * er1[1])
    return [x1, px1, y + py * S / dd, py, z + dp1 * S / dd - mp.mpf(L) * bt, d]


def edge_lin(r, irho, e, fint, gap, method, is_exit):
    x, px, y, py, z, d = r
    if fint == 0.0 or gap == 0.0 or method == 0:
        fc = 0.0
    else:
        s, c = math.sin(e), math.cos(e)
        fc = irho * gap * fint * (1.0 + s * s) / c
    fx = irho * math.tan(e)
    if method == 2:
        fy = irho * mp.tan(e - fc / (1 + d)) / (1 + d)
    elif method == 3:
        fy = irho * mp.tan(e - fc - px / (1 + d)) if is_exit else irho * mp.tan(e - fc + px / (1 + d))
    else:
        fy = irho * mp.tan(e - fc / (1 + d))
    return [x, px + x * fx, y, py - y * fy, z, d]


def pxyz(dp1, px, py):
    v = dp1 * dp1 - px * px - py * py
    return mp.sqrt(v) if v > 0 else mp.mpf(0)


def yrot(r, phi, bt):
    if phi == 0:
        return r
    x, px, y, py, z, d = r
    dp1 = bt + d
    c, s = mp.cos(phi), mp.sin(phi)
    pz = pxyz(dp1, px, py)
    p = c * pz - s * px
    return [x * pz / p, s * pz + c * px, y + x * py * s / p, py, z + dp1 * x * s / p, d]


def bend_edge(r, rhoinv, theta, bt):
    x, px, y, py, z, d = r
    dp1 = bt + d
    c, s = mp.cos(theta), mp.sin(theta)
    pz = pxyz(dp1, px, py)
    d2 = pxyz(dp1, 0, py)
    pxn = px * c + (pz - rhoinv * x) * s
    das = mp.asin(px / d2) - mp.asin(pxn / d2)
    num = x * (px * mp.sin(2 * theta) + s * s * (2 * pz - rhoinv * x))
    den = pxyz(dp1, pxn, py) + pz * c - px * s
    return [x * c + num / den, pxn, y + py * (theta / rhoinv + das / rhoinv), py,
            z + dp1 / rhoinv * (theta + das), d]


def bend_fringe(r, irho, gK, bt):
    """Forest 13.13 as written in src/tracking.jl:752-815 (same formula, 50 digits)."""
    x, px, y, py, z, d = r
    b0 = mp.mpf(irho)
    gK = mp.mpf(gK)
    dp1 = bt + d
    pz = pxyz(dp1, px, py)
    xp, yp = px / pz, py / pz
    term1 = 1 + xp ** 2 * (2 + yp ** 2)
    term2 = xp / (1 + yp ** 2)
    phi = -b0 * mp.tan(b0 * gK * term1 * pz - mp.atan(term2))
    px2, py2, pz2 = px ** 2, py ** 2, pz ** 2
    pz4 = pz2 ** 2
    py2z2 = py2 + pz2
    arg_sec = (b0 * gK * (pz4 + px2 * (py2 + 2 * pz2))) / (pz2 * pz) - mp.atan(px * pz / (py2 + pz2))
    powsec = mp.sec(arg_sec) ** 2
    py4 = py2 ** 2
    py6 = py4 * py2
    pz5, pz6 = pz4 * pz, pz4 * pz2
    sq = py2z2 ** 2
    denom = pz5 * (py4 + px2 * pz2 + 2 * py2 * pz2 + pz4)
    dpx = -(b0 * (px2 * pz4 * (py2 - pz2) - pz6 * (py2 + pz2) + b0 * gK * px * (
        pz2 * sq * (2 * py2 + 3 * pz2) + px2 ** 2 * (3 * py2 * pz2 + 2 * pz4) +
        px2 * (3 * py6 + 8 * py4 * pz2 + 9 * py2 * pz4 + 5 * pz6)))) * powsec / denom
    dpy = -(b0 * py * (px * pz4 * (py2 + pz2) + b0 * gK * (-(pz4 * sq) + px2 ** 2 * (3 * py2 * pz2 + 4 * pz4) +
                                                      px2 * (3 * py6 + 10 * py4 * pz2 + 11 * py2 * pz4 + 3 * pz6)))) * powsec / denom
    dd = (b0 * dp1 * (px * pz4 * (py2 - pz2) + b0 * gK * (-(pz4 * sq) + px2 ** 2 * (3 * py2 * pz2 + 2 * pz4) +
                                                        px2 * (3 * py6 + 8 * py4 * pz2 + 7 * py2 * pz4 + pz6)))) * powsec / denom
    yf = (2 * y) / (1 + mp.sqrt(1 - 2 * dpy * y))
    return [x + dpx * yf ** 2 / 2, px, yf, py - phi * yf, z - dd * yf ** 2 / 2, d]


def poly(e_a, e_b):
    a = [mp.mpf(v) for v in e_a] + [mp.mpf(0)] * (11 - len(e_a))
    b = [mp.mpf(v) for v in e_b] + [mp.mpf(0)] * (11 - len(e_b))
    return a, b


def run_case(kind, p, r, bt):
    if kind == "drift":
        return drift(r, p["L"], bt)
    if kind == "quad":
        a, b = poly([0] * 4, [0, p["k1"], 0, 0])
        return yoshida(r, p["L"], p["n"], bt, lambda q, K: kick(q, a, b, 1, K))
    if kind == "sext":
        a, b = poly([0] * 4, [0, 0, p["k2"] / 2, 0])
        return yoshida(r, p["L"], p["n"], bt, lambda q, K: kick(q, a, b, 2, K))
    if kind == "oct":
        a, b = poly([0] * 4, [0, 0, 0, p["k3"] / 6])
        return yoshida(r, p["L"], p["n"], bt, lambda q, K: kick(q, a, b, 3, K))
    if kind == "thinmp":
        a, b = poly(p["a"], p["b"])
        return kick(r, a, b, p["m"], 1)
    if kind == "sbend":
        L, ang = p["L"], p["angle"]
        irho = ang / L
        a, b = poly([0] * 4, [0] * 4)
        if p.get("fbe", 1):
            r = edge_lin(r, irho, p["e1"], p.get("fint1", 0.0), p.get("gap", 0.0), p.get("fbe", 1), False)
        r = yoshida(r, L, p["n"], bt, lambda q, K: bkick(q, a, b, 0, K, irho, bt))
        if p.get("fbx", 1):
            r = edge_lin(r, irho, p["e2"], p.get("fint2", 0.0), p.get("gap", 0.0), p.get("fbx", 1), True)
        return r
    if kind in ("esbend_le", "esbend_body"):
        L, ang = p["L"], p["angle"]
        irho = ang / L
        if kind == "esbend_le":
            r = edge_lin(r, irho, p["e1"], p["fint1"], p["gap"], 1, False)
        n = p["n"]
        if n == 0:
            r = body_geometry(r, irho, L, bt)
        else:
            L1, L2, _, _ = steps(L, n)
            for _ in range(n):  # pure dipole: the sub-kicks are identities
                for Ls in (L1, L2, L2, L1):
                    r = body_geometry(r, irho, Ls, bt)
        if kind == "esbend_le":
            r = edge_lin(r, irho, p["e2"], p["fint2"], p["gap"], 1, True)
        return r
    if kind == "forest_esbend":
        L, ang, e1, e2 = p["L"], p["angle"], p["e1"], p["e2"]
        irho = ang / L
        r = yrot(r, e1, bt)
        r = bend_fringe(r, irho, p["gk"], bt)
        r = bend_edge(r, irho, -e1, bt)
        L1, L2, _, _ = steps(L, p["n"])
        for _ in range(p["n"]):
            for Ls in (L1, L2, L2, L1):
                r = body_geometry(r, irho, Ls, bt)
        r = bend_edge(r, irho, -e2, bt)
        r = bend_fringe(r, -irho, p["gk"], bt)
        return yrot(r, e2, bt)
    if kind == "wedge":
        return wedge_geometry(r, p["angle"] / p["L"], p["L"], p["e1"], p["e2"], bt)
    if kind == "rf":
        L = p["L"]
        beta = 1 / mp.mpf(bt)
        if L > 0:
            r = drift(r, L / 2, bt)
        nv = mp.mpf(p["volt"]) / p["energy"]
        phase = mp.mpf(2 * math.pi) * p["freq"] * ((r[4] - p["lag"]) / C_LIGHT) - p["philag"]
        r = r[:5] + [r[5] - nv * mp.sin(phase) / (beta * beta)]
        if L > 0:
            r = drift(r, L / 2, bt)
        return r
    raise ValueError(kind)


CASES = [
    # name, kind, params, energy  (element cases follow test/verify_elements.jl:128-147)
    ("Drift", "drift", {"L": 0.7}),
    ("Quadrupole", "quad", {"L": 0.4, "k1": 1.3, "n": 10}),
    ("Sextupole", "sext", {"L": 0.4, "k2": 2.0, "n": 10}),
    ("Octupole", "oct", {"L": 0.4, "k3": -3.0, "n": 10}),
    ("ThinMultipole", "thinmp", {"a": [0.0] * 4, "b": [0.0, 0.2, -0.1, 0.05], "m": 3}),
    ("ThinMultipole10", "thinmp", {"a": [1e-3, -2e-2, 0.3, 5.0, -40.0, 300.0, 2e3, -3e4, 1e5, 2e6, -1e7],
                                   "b": [-1e-3, 1e-2, 0.7, -4.0, 60.0, -500.0, 4e3, 2e4, -3e5, 1e6, 3e7], "m": 10}),
    ("SBend", "sbend", {"L": 0.9, "angle": 0.15, "e1": 0.03, "e2": 0.02, "n": 10}),
    ("RBend", "sbend", {"L": 0.9, "angle": 0.15, "e1": 0.075, "e2": 0.075, "n": 10}),
    ("SBendFringe", "sbend", {"L": 1.5, "angle": 0.157, "e1": 0.01, "e2": 0.0, "n": 10, "fint1": 0.5,
                              "fint2": 0.5, "gap": 0.03}),
    ("ExactBodyN0", "esbend_body", {"L": 0.9, "angle": 0.15, "n": 0}),
    ("ExactBodyN10", "esbend_body", {"L": 1.5, "angle": 2 * math.pi / 40, "n": 10}),
    ("ExactSBendLinearEdge", "esbend_le", {"L": 1.5, "angle": 2 * math.pi / 40, "e1": 0.0, "e2": 0.0,
                                           "n": 10, "fint1": 0.5, "fint2": 0.5, "gap": 0.03}),
    ("ExactSBend", "forest_esbend", {"L": 0.9, "angle": 0.15, "e1": 0.03, "e2": 0.02, "n": 10, "gk": 0.0}),
    ("ExactSBendGK", "forest_esbend", {"L": 1.0, "angle": math.pi / 9, "e1": math.pi / 18,
                                       "e2": math.pi / 18, "n": 10, "gk": 0.02}),
    ("RFCavity", "rf", {"L": 0.3, "volt": 2.0e6, "freq": 500.0e6, "lag": 0.002, "philag": 0.3,
                        "energy": 3.5e9}),
    # linear edge methods 2 (SOLEIL) and 3 (THOMX), entrance and exit, with fringe
    # (src/tracking.jl:174-178, 207-213; the 50-digit values restate the edge formula)
    ("SBendSOLEIL", "sbend", {"L": 1.5, "angle": 0.157, "e1": 0.02, "e2": 0.03, "n": 10, "fint1": 0.5,
                              "fint2": 0.4, "gap": 0.03, "fbe": 2, "fbx": 2}),
    ("SBendTHOMX", "sbend", {"L": 1.5, "angle": 0.157, "e1": 0.02, "e2": 0.03, "n": 10, "fint1": 0.5,
                             "fint2": 0.4, "gap": 0.03, "fbe": 3, "fbx": 3}),
    ("SBendMixedMethods", "sbend", {"L": 0.9, "angle": -0.12, "e1": -0.01, "e2": 0.04, "n": 8, "fint1": 0.6,
                                    "fint2": 0.5, "gap": 0.05, "fbe": 3, "fbx": 2}),
    # Forest's rotated faces with the fringe off, against the plane geometry of a hard-edge
    # wedge magnet (an independent construction: wedge_geometry)
    ("ExactSBendWedge", "wedge", {"L": 0.9, "angle": 0.15, "e1": 0.03, "e2": 0.02, "n": 0}),
    ("ExactSBendWedgeRect", "wedge", {"L": 1.0, "angle": 0.3, "e1": 0.15, "e2": 0.15, "n": 10}),
    ("ExactSBendWedgeMixed", "wedge", {"L": 1.2, "angle": 0.2, "e1": -0.05, "e2": 0.08, "n": 4}),
]


def main():
    bt = beam_beti(E_REF)
    out = {"energy": E_REF, "beti": bt, "particles": P10, "yoshida_closed_form": yoshida_closed(),
           "beti_mp": mp.nstr(1 / mp.sqrt(1 - 1 / ((E_REF + mp.mpf(M_E)) / M_E) ** 2), 40), "cases": []}
    for name, kind, p in CASES:
        rows = []
        for r in P10:
            q = run_case(kind, p, mpv(r), bt)
            rows.append([mp.nstr(v, 30) for v in q])
        out["cases"].append({"name": name, "kind": kind, "params": p, "out": rows})
    with open(os.path.join(HERE, "mpmath_golden.json"), "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", len(out["cases"]), "cases")


if __name__ == "__main__":
    main()
