// tp_adjoint.h -- inverse maps and vector-Jacobian products of the element maps.
//
// The reverse pass walks a turn backwards from the state at its end.  Every sub-step of
// an element (drift, kick, exact body, edge, RF kick) is undone by its inverse map --
// the same map run with -L / -K, which is exact in exact arithmetic because the drift
// leaves the momenta unchanged, the kick leaves x and y unchanged, and the exact body is
// the flow of a Hamiltonian for a signed length -- and its VJP is applied at the
// reconstructed input.  Derivatives are those of the reference formulas
// (src/tracking.jl:87-108, 160-221, 243-305, 858-894, 502-518) taken analytically.
//
// St holds the particle state, Ad the adjoint (dL/dstate) of the same point.
#ifndef TP_ADJOINT_H
#define TP_ADJOINT_H

#include "tp_extra.h"

namespace tpk {

struct Ad {
  double x, px, y, py, z, d;
};

// ---------------------------------------------------------------------------------------
// drift6 (exact Hamiltonian): inverse + VJP.  s is the drift's output on entry and its
// input on return.  NormL depends only on px, py, delta, which the drift leaves alone.
// ---------------------------------------------------------------------------------------
// drift6 with USE_EXACT_HAMILTONIAN = false (src/tracking.jl:100-107): NormL = L/(1+delta),
// x' = x + NormL px, y' = y + NormL py, z' = z + NormL (px^2 + py^2) / (2 (1+delta)).
TP_DEV void drift_lin_back(St& s, Ad& a, double L) {
  const double u = 1.0 + s.d;
  const double N = L / u;
  s.x = s.x - N * s.px;
  s.y = s.y - N * s.py;
  const double q = N * (s.px * s.px + s.py * s.py) / (2.0 * u);
  s.z = s.z - q;
  // dN/d delta = -N/u, dq/d delta = -2q/u, dq/d px = N px / u
  const double g = a.z * N / u;
  const double ad = -(a.x * s.px + a.y * s.py) * N / u - a.z * 2.0 * q / u;
  a.px = a.px + a.x * N + g * s.px;
  a.py = a.py + a.y * N + g * s.py;
  a.d = a.d + ad;
}

// SER: 1/pz by the series form (inv_pz) -- the batched reverse kernel, whose launches fill
// the SIMDs; otherwise the refined rsqrt.
template <bool SER = false>
TP_DEV void drift_back(St& s, Ad& a, const Cx& c, double L, double Lb, bool lin = false) {
  if (lin) {
    drift_lin_back(s, a, L);
    return;
  }
#if TP_FAST
  const double rz = SER ? inv_pz(s, c, 1.0) : rsqrt_nr(c.A - s.px * s.px - s.py * s.py);
  const double NormL = L * rz;
  const double ipz2 = rz * rz;
#else
  const double pz2 = c.A - s.px * s.px - s.py * s.py;
  const double NormL = L / sqrt(pz2);
  const double ipz2 = 1.0 / pz2;
#endif
  s.x = s.x - NormL * s.px;
  s.y = s.y - NormL * s.py;
#if TP_FAST
  s.z = s.z - __builtin_fma(NormL, c.B, -Lb);
#else
  s.z = s.z - NormL * c.B + Lb;
#endif
  // d NormL / d px = NormL px / pz2, / d py = NormL py / pz2, / d delta = -NormL B / pz2
  const double f = (a.x * s.px + a.y * s.py + a.z * c.B) * NormL * ipz2;
  a.px = a.px + a.x * NormL + f * s.px;
  a.py = a.py + a.y * NormL + f * s.py;
  a.d = a.d + a.z * NormL - f * c.B;
}

// F(w) = sum c_n w^n (c_n = b_n + i a_n) and F'(w), for the kick specialisations.
template <int KK>
TP_DEV void field_and_slope(const St& s, const double* __restrict__ poly, int m, double& re, double& im, double& U,
                            double& V) {
  if (KK == TP_KICK_NONE) {
    re = im = U = V = 0.0;
  } else if (KK == TP_KICK_QUAD) {
    const double b1 = poly[2];
    re = b1 * s.x;
    im = b1 * s.y;
    U = b1;
    V = 0.0;
  } else if (KK == TP_KICK_SEXT) {
    const double b2 = poly[4];
    const double rx = b2 * s.x, iy = b2 * s.y;
    re = rx * s.x - iy * s.y;
    im = iy * s.x + rx * s.y;
    U = 2.0 * rx;
    V = 2.0 * iy;
  } else {
    double ReSum = poly[2 * m], ImSum = poly[2 * m + 1];
    double DR = m * poly[2 * m], DI = m * poly[2 * m + 1];
    for (int i = m - 1; i >= 0; --i) {
      const double t = ReSum * s.x - ImSum * s.y + poly[2 * i];
      ImSum = ImSum * s.x + ReSum * s.y + poly[2 * i + 1];
      ReSum = t;
      if (i >= 1) {
        const double u = DR * s.x - DI * s.y + i * poly[2 * i];
        DI = DI * s.x + DR * s.y + i * poly[2 * i + 1];
        DR = u;
      }
    }
    re = ReSum;
    im = ImSum;
    if (m == 0) {
      DR = 0.0;
      DI = 0.0;
    }
    U = DR;
    V = DI;
  }
}

// Re and Im of w^g for the strength gradient (g = 0..3 covers dipole..octupole).
TP_DEV void wpow(const St& s, int g, double& wr, double& wi) {
  wr = 1.0;
  wi = 0.0;
  for (int i = 0; i < g; ++i) {
    const double t = wr * s.x - wi * s.y;
    wi = wr * s.y + wi * s.x;
    wr = t;
  }
}

// strthinkick: inverse + VJP; gacc += dL/d b_g (main strength of the element).
template <int KK>
TP_DEV void kick_back(St& s, Ad& a, const double* __restrict__ poly, int m, double K, int g, double& gacc) {
  if (KK == TP_KICK_NONE) {
    // the identity map, but a magnet at zero strength still has a strength gradient:
    // d px' / d b_g = -K Re(w^g), d py' / d b_g = K Im(w^g)
    if (g >= 0) {
      double wr, wi;
      wpow(s, g, wr, wi);
      gacc = gacc + K * (a.py * wi - a.px * wr);
    }
    return;
  }
  double re, im, U, V;
  field_and_slope<KK>(s, poly, m, re, im, U, V);
  s.px = s.px + K * re;
  s.py = s.py - K * im;
  if (g >= 0) {
    double wr, wi;
    wpow(s, g, wr, wi);
    gacc = gacc + K * (a.py * wi - a.px * wr);
  }
  const double ax = a.x - K * (a.px * U - a.py * V);
  const double ay = a.y + K * (a.px * V + a.py * U);
  a.x = ax;
  a.y = ay;
}

#if TP_FAST
// ---------------------------------------------------------------------------------------
// A Yoshida step of a thick multipole backwards, in the forms the fast forward kernel uses
// (mpole_steps): drifts without their path-length terms (drift_xy), the step's path length
// added to z at once, the sextupole kick with K b2 folded (kick_sext_kb).  Inside a step
// neither z nor its adjoint is read or changed by a drift or a kick, so a.z B is formed once
// per element and the a.z NormL terms of delta's adjoint once per step.
// ---------------------------------------------------------------------------------------
// drift_xy backwards: x and y undone; the adjoints of px, py and delta updated except the
// a.z NormL term of delta.  azB = a.z B.  Returns NormL.  PAIR: two drifts of length L with
// nothing between them (drift_xy_pair) undone at once -- neither the momenta nor a.x, a.y
// change in a drift, so the second sees the first's NormL and f, and the pair is one drift
// with NormL and its derivative doubled.
template <bool SER, bool PAIR = false>
TP_DEV double drift_xy_back(St& s, Ad& a, const Cx& c, double L, double azB) {
  const double rz = SER ? inv_pz(s, c, 1.0) : rsqrt_nr(c.A - s.px * s.px - s.py * s.py);
  const double NormL = L * rz;
  const double Nm = PAIR ? NormL + NormL : NormL;
  s.x = s.x - Nm * s.px;
  s.y = s.y - Nm * s.py;
  // d NormL / d px = NormL px / pz^2, / d py likewise, / d delta = -NormL B / pz^2
  const double f = __builtin_fma(a.x, s.px, __builtin_fma(a.y, s.py, azB)) * (Nm * (rz * rz));
  a.px = a.px + a.x * Nm + f * s.px;
  a.py = a.py + a.y * Nm + f * s.py;
  a.d = a.d - f * c.B;
  return NormL;
}

// The series drift of one length, as the forward's inv_pz forms it per element and particle:
// NormL = Lr + w (c1 + w c2), Lr = L R, c1 = Lr H, c2 = c1 H3; and d1 = 2 c1, d2 = 4 c2, so that
// NormL / pz^2 = 2 dNormL/dw = d1 + w d2 -- the exact derivative of the map the forward
// applied (lanes past the series bound take the refined rsqrt and its own derivative).
struct DSer {
  double Lr, c1, c2, d1, d2;
};
TP_DEV DSer dser(const Cx& c, double L) {
  DSer k;
  k.Lr = L * c.R;
  k.c1 = k.Lr * c.H;
  k.c2 = k.c1 * c.H3;
  k.d1 = k.c1 + k.c1;
  k.d2 = k.c2 * 4.0;
  return k;
}
// PAIR: two drifts of length L undone at once, as in drift_xy_back.
template <bool PAIR = false>
TP_DEV double drift_xy_back_ser(St& s, Ad& a, const Cx& c, double L, const DSer& k, double azB) {
  const double w = __builtin_fma(s.px, s.px, s.py * s.py);
  double NormL = __builtin_fma(w, __builtin_fma(w, k.c2, k.c1), k.Lr);
  double g = __builtin_fma(w, k.d2, k.d1);
  if (__builtin_expect(!(w < c.WB), 0)) {
    const double rz = rsqrt_nr(c.A - w);
    NormL = L * rz;
    g = NormL * (rz * rz);
  }
  const double Nm = PAIR ? NormL + NormL : NormL;
  const double gm = PAIR ? g + g : g;
  s.x = s.x - Nm * s.px;
  s.y = s.y - Nm * s.py;
  const double f = __builtin_fma(a.x, s.px, __builtin_fma(a.y, s.py, azB)) * gm;
  a.px = a.px + a.x * Nm + f * s.px;
  a.py = a.py + a.y * Nm + f * s.py;
  a.d = a.d - f * c.B;
  return NormL;
}

// Quadrupole kick (px -= K b1 x, py += K b1 y) backwards with Kb = K b1 (one rounding of
// the strength per element, below the rounding of the inverse drifts around it).  gq
// accumulates a.py y - a.px x, so that dL/d b1 = sum over the kicks of K gq.
TP_DEV void kick_quad_back(St& s, Ad& a, double Kb, double& gq) {
  s.px = __builtin_fma(Kb, s.x, s.px);
  s.py = __builtin_fma(-Kb, s.y, s.py);
  gq = __builtin_fma(a.py, s.y, __builtin_fma(-a.px, s.x, gq));
  a.x = __builtin_fma(-Kb, a.px, a.x);
  a.y = __builtin_fma(Kb, a.py, a.y);
}

// Sextupole kick in the forward's folded form (kick_sext_kb, Kb = K b2) backwards: its four
// FMAs undone in reverse order.  With rx = Kb x, iy = Kb y the Jacobian's kick block is
// [[-2 rx, 2 iy], [2 iy, 2 rx]], so a.x += 2 tx, a.y += 2 ty with tx = a.py iy - a.px rx,
// ty = a.px iy + a.py rx; and x tx + y ty = Kb dL/dKb, so dL/d b2 = (sum of x tx + y ty) / b2
// over the element's kicks (gs accumulates the sum).
TP_DEV void kick_sext_back(St& s, Ad& a, double Kb, double& gs) {
  const double rx = Kb * s.x, iy = Kb * s.y;
  s.py = __builtin_fma(-rx, s.y, s.py);
  s.py = __builtin_fma(-iy, s.x, s.py);
  s.px = __builtin_fma(-iy, s.y, s.px);
  s.px = __builtin_fma(rx, s.x, s.px);
  const double tx = __builtin_fma(-a.px, rx, a.py * iy);
  const double ty = __builtin_fma(a.px, iy, a.py * rx);
  gs = __builtin_fma(s.x, tx, __builtin_fma(s.y, ty, gs));
  a.x = __builtin_fma(2.0, tx, a.x);
  a.y = __builtin_fma(2.0, ty, a.y);
}

// The steps of a Quadrupole / Sextupole / Octupole record whose kick is the identity, a pure
// quadrupole or a pure sextupole, exact Hamiltonian, undone; grad: the element has a gradient
// slot and its main strength is the kick's own coefficient (b1 of QUAD, b2 of SEXT).
template <int KK, bool SER>
TP_DEV void mpole_back_fast(St& s, Ad& a, const Cx& c, const double* __restrict__ p, int ns, bool grad,
                            double& gacc) {
  const double L1 = p[0], L2 = p[1], K1 = p[2], K2 = p[3];
  const double Lsb = 2.0 * p[4] + 2.0 * p[5];  // the forward's 2 L1b + 2 L2b
  const double azB = a.z * c.B;
  const double b = KK == TP_KICK_QUAD ? p[6 + 2] : (KK == TP_KICK_SEXT ? p[6 + 4] : 0.0);
  const double Kb1 = K1 * b, Kb2 = K2 * b;
  double g1 = 0.0, g2 = 0.0;
  // The forward's steps (mpole_steps) undone from the exit: the half-drift pairs where
  // steps meet undone as one (PAIR), each step's path sum formed as the forward formed it.
#define TP_KICKS_BACK(D2)                                                 \
  if constexpr (KK == TP_KICK_QUAD) kick_quad_back(s, a, Kb1, g1);        \
  if constexpr (KK == TP_KICK_SEXT) kick_sext_back(s, a, Kb1, g1);        \
  S = S + D2;                                                             \
  if constexpr (KK == TP_KICK_QUAD) kick_quad_back(s, a, Kb2, g2);        \
  if constexpr (KK == TP_KICK_SEXT) kick_sext_back(s, a, Kb2, g1);        \
  S = S + D2;                                                             \
  if constexpr (KK == TP_KICK_QUAD) kick_quad_back(s, a, Kb1, g1);        \
  if constexpr (KK == TP_KICK_SEXT) kick_sext_back(s, a, Kb1, g1);
#define TP_STEP_BACK(D1, D1PAIR, D2)                                      \
  {                                                                       \
    double S = D1;                                                        \
    for (int st = 1; st < ns; ++st) {                                     \
      TP_KICKS_BACK(D2)                                                   \
      const double N = D1PAIR;                                            \
      S = S + N;                                                          \
      s.z = s.z - __builtin_fma(S, c.B, -Lsb);                            \
      a.d = __builtin_fma(a.z, S, a.d);                                   \
      S = N;                                                              \
    }                                                                     \
    TP_KICKS_BACK(D2)                                                     \
    S = S + D1;                                                           \
    s.z = s.z - __builtin_fma(S, c.B, -Lsb);                              \
    a.d = __builtin_fma(a.z, S, a.d);                                     \
  }
  if (ns > 0) {
    if constexpr (SER) {
      const DSer k1 = dser(c, L1), k2 = dser(c, L2);
      TP_STEP_BACK(drift_xy_back_ser(s, a, c, L1, k1, azB), drift_xy_back_ser<true>(s, a, c, L1, k1, azB),
                   drift_xy_back_ser(s, a, c, L2, k2, azB))
    } else {
      TP_STEP_BACK(drift_xy_back<false>(s, a, c, L1, azB), (drift_xy_back<false, true>(s, a, c, L1, azB)),
                   drift_xy_back<false>(s, a, c, L2, azB))
    }
  }
#undef TP_STEP_BACK
#undef TP_KICKS_BACK
  if (grad) {
    if constexpr (KK == TP_KICK_QUAD) gacc = __builtin_fma(K1, g1, __builtin_fma(K2, g2, gacc));
    if constexpr (KK == TP_KICK_SEXT) gacc = gacc + g1 / b;
  }
}
#endif

// bndthinkick: inverse + VJP (no strength gradient for bends).
template <int KK>
TP_DEV void bkick_back(St& s, Ad& a, const double* __restrict__ poly, int m, double K, double Ki, double irho,
                       double beti) {
  double re, im, U, V;
  field_and_slope<KK>(s, poly, m, re, im, U, V);
  s.px = s.px + K * (re - (s.d - s.x * irho) * irho);
  if (KK != TP_KICK_NONE) s.py = s.py - K * im;
  s.z = s.z - Ki * s.x * beti;
  const double ax = a.x - K * (a.px * (U + irho * irho) - a.py * V) + a.z * Ki * beti;
  const double ay = a.y + K * (a.px * V + a.py * U);
  a.d = a.d + a.px * K * irho;
  a.x = ax;
  a.y = ay;
}

// ---------------------------------------------------------------------------------------
// pxyz(dp1, px, py) and its reciprocal (src/tracking.jl:717-720).  The fast build takes
// both from one refined hardware estimate: y0 = rsq(v), e = 1 - v y0^2, then 1/sqrt(v) =
// y0 (1 + e/2 + 3e^2/8) and sqrt(v) = v y0 (1 + e/2 + 3e^2/8), as rsqrt_nr / sqrt_nr.
// ---------------------------------------------------------------------------------------
TP_DEV double pxyz_r(const Cx& c, double px, double py, double& ipz) {
  const double v = c.BB - px * px - py * py;
#if TP_FAST
  const double y = __builtin_amdgcn_rsq(v);
  const double t = v * y;
  const double e = __builtin_fma(-t, y, 1.0);
  const double p = __builtin_fma(e, 0.375, 0.5);
  ipz = __builtin_fma(y * e, p, y);
  return __builtin_fmax(__builtin_fma(t * e, p, t), 0.0);
#else
  const double pz = v > 0.0 ? sqrt(v) : 0.0;
  ipz = 1.0 / pz;
  return pz;
#endif
}

// ---------------------------------------------------------------------------------------
// Exact sector-bend body (Forest 12.18, src/tracking.jl:858-894) backwards.  s enters as
// the body's output and leaves as its input.  The inverse is the body of length -L (sn ->
// -sn); its intermediates are the forward map's: the inverse's pz is the forward output's
// pzn, its px_new is the input px and its pzn the input pz, and its path term is minus the
// forward dasin.  The VJP is then the reverse sweep over the forward body's operations
//   pz = sqrt(B^2 - px^2 - py^2), pzmx = pz - 1 - x irho, pxn = px cs + pzmx sn,
//   pzn = sqrt(B^2 - pxn^2 - py^2), d2 = sqrt(B^2 - py^2),
//   dasin = L + (asin(px/d2) - asin(pxn/d2)) / irho  (d asin(px/d2)/d px = 1/pz),
//   x' = (pzn - pzmx cs + px sn - 1) / irho, y' = y + py dasin, z' = z + B dasin - L beti,
// B = beti + delta, each adjoint accumulated as its operation is undone.
// rho = 1/irho and id2 = 1/d2 are formed by the caller once per element (d2 = pxyz(dp1, 0,
// py) is constant through the body).  pz_io / ipz_io enter as pxyz(dp1, px, py) of the
// output and its reciprocal and leave as those of the input, which are the output's of the
// body before it when the kick between them is the identity.
// ---------------------------------------------------------------------------------------
TP_DEV void ebody_back(St& s, Ad& a, const Cx& c, double irho, double rho, double L, double Lb, double cs, double sn,
                       double d2, double id2, double& pz_io, double& ipz_io) {
  const double B = c.B;
  const double py = s.py;
  const double pxn = s.px;                        // output px
  const double pzn = pz_io, ipzn = ipz_io;        // output pz and 1/pz
  const double pzmx_o = pzn - (1.0 + s.x * irho);
  const double px = pxn * cs - pzmx_o * sn;       // input px
  double ipz;
  const double pz = pxyz_r(c, px, py, ipz);       // input pz and 1/pz
#if TP_FAST
  // asin(px/d2) - asin(pxn/d2) as one series asin, as in the forward body
  const double dasin = L + asin_diff(px, pz, pxn, pzn, d2, id2) * rho;
  const double x_in = (pz - pzmx_o * cs - pxn * sn - 1.0) * rho;
#else
  const double dasin = L + (asin_s(clamp1(px / d2)) - asin_s(clamp1(pxn / d2))) / irho;
  const double x_in = (pz - pzmx_o * cs - pxn * sn - 1.0) / irho;
#endif
  // reverse sweep; y and z pass through with coefficient 1
  const double a_dasin = a.y * py + a.z * B;
  double a_py = a.py + a.y * dasin;
  double a_B = a.z * dasin;
  const double g = a.x * rho;                     // x' = (...) rho
  const double a_pzn0 = g;
  double a_pzmx = -g * cs;
  double a_px = g * sn;
  const double h = a_dasin * rho;                 // dasin = L + dif rho
  a_px = a_px + h * ipz;
  double a_pxn = a.px - h * ipzn;
  // d2 enters dif through both arguments: d(dif)/d(d2) = (pxn/pzn - px/pz) / d2
  const double a_d2 = h * (pxn * ipzn - px * ipz) * id2;
  a_py = a_py - a_d2 * py * id2;
  a_B = a_B + a_d2 * B * id2;
  // pzn = sqrt(B^2 - pxn^2 - py^2)
  const double tn = a_pzn0 * ipzn;
  a_pxn = a_pxn - tn * pxn;
  a_py = a_py - tn * py;
  a_B = a_B + tn * B;
  // pxn = px cs + pzmx sn
  a_px = a_px + a_pxn * cs;
  a_pzmx = a_pzmx + a_pxn * sn;
  // pzmx = pz - 1 - x irho
  const double a_x = -a_pzmx * irho;
  // pz = sqrt(B^2 - px^2 - py^2)
  const double t0 = a_pzmx * ipz;
  a_px = a_px - t0 * px;
  a_py = a_py - t0 * py;
  a_B = a_B + t0 * B;
  a.x = a_x;
  a.px = a_px;
  a.py = a_py;
  a.d = a.d + a_B;
  s.y = s.y - py * dasin;
#if TP_FAST
  s.z = s.z - __builtin_fma(B, dasin, -Lb);
#else
  s.z = s.z - B * dasin + Lb;
#endif
  s.x = x_in;
  s.px = px;
  pz_io = pz;
  ipz_io = ipz;
}

// ---------------------------------------------------------------------------------------
// Linear dipole edge (src/tracking.jl:160-221): inverse + VJP.  blk = [fx, fyc, e, fc].
// ---------------------------------------------------------------------------------------
// MEMO: the fringe term tan(e - fc/(1 + delta)) from the particle's edge memo (edge_tan: the
// same bits as the call).
template <bool MEMO = false>
TP_DEV void edge_fy(const St& s, int method, const double* __restrict__ blk, double irho, bool is_exit, double& fy,
                    double& dfy_px, double& dfy_d, EdgeMemo* mm = nullptr) {
  const double fyc = blk[1], e = blk[2], fc = blk[3];
  const double opd = 1.0 + s.d;
  dfy_px = 0.0;
  dfy_d = 0.0;
  if (method == 3) {
    const double sg = is_exit ? -1.0 : 1.0;
    const double arg = e - fc + sg * s.px / opd;
    const double t = tan_lib(arg);
    fy = irho * t;
    const double sec2 = 1.0 + t * t;
    dfy_px = irho * sec2 * sg / opd;
    dfy_d = -irho * sec2 * sg * s.px / (opd * opd);
    return;
  }
  double t, dt_d;
  if (fc == 0.0) {
    t = fyc;
    dt_d = 0.0;
  } else {
    const double tt = MEMO ? edge_tan(*mm, s, e, fc) : tan_lib(e - fc / opd);
    t = irho * tt;
    dt_d = irho * (1.0 + tt * tt) * fc / (opd * opd);
  }
  if (method == 2) {
    fy = t / opd;
    dfy_d = dt_d / opd - t / (opd * opd);
  } else {
    fy = t;
    dfy_d = dt_d;
  }
}

template <bool MEMO = false>
TP_DEV void edge_back(St& s, Ad& a, int method, const double* __restrict__ blk, double irho, bool is_exit,
                      EdgeMemo* mm = nullptr) {
  const double fx = blk[0];
  s.px = s.px - s.x * fx;
  double fy, dfy_px, dfy_d;
  edge_fy<MEMO>(s, method, blk, irho, is_exit, fy, dfy_px, dfy_d, mm);
  s.py = s.py + s.y * fy;
  // px' = px + x fx; py' = py - y fy(px, delta)
  a.x = a.x + a.px * fx;
  a.y = a.y - a.py * fy;
  a.px = a.px - a.py * s.y * dfy_px;
  a.d = a.d - a.py * s.y * dfy_d;
}

// RF kick (src/tracking.jl:508-513): delta' = delta - nv sin(phi)/beta^2, phi = w (z-lag)/c - philag
TP_DEV void rf_kick_back(St& s, Ad& a, const double* __restrict__ p) {
  const double w = p[3], lag = p[4], philag = p[5], nv = p[2], bb = p[6];
  const double phase = w * ((s.z - lag) / 2.99792458e8) - philag;
  const double sn = sin_lib(phase);
  s.d = s.d + nv * sn / bb;
  const double cph = cos_lib(phase);
  a.z = a.z - a.d * nv * cph * (w / 2.99792458e8) / bb;
}

// ---------------------------------------------------------------------------------------
// Forest maps of ExactSBend (src/tracking.jl:727-851) backwards.
// Inverses: yrot(-phi) undoes yrot(phi) and bend_edge(-theta) undoes bend_edge(theta)
// exactly in exact arithmetic (rotations of the reference frame, in free space and in the
// dipole field, compose additively in the angle; checked to 1e-40 at 40 digits).
// bend_fringe changes y, x, py and z with coefficients phi, dpx, dpy, dd that depend on the
// input (px, py, delta); px and delta pass through.  Its inverse solves
// py_out = py - phi(px, py, delta) yf for the input py by fixed-point steps (a contraction
// with factor |d phi/d py| |yf| ~ b0 |px py| |y|, below 1e-6 in any lattice that keeps the
// particle), then y = yf - dpy yf^2 / 2 (yf = 2y / (1 + sqrt(1 - 2 dpy y)) is the root of
// dpy yf^2 / 2 - yf + y = 0), x = x_out - dpx yf^2 / 2 and z = z_out + dd yf^2 / 2.
// VJPs: the Jacobian of each map at the reconstructed input, one input direction at a time,
// by forward-mode dual numbers (D1) through the same formulas; the maps run twice per bend
// per turn at most, so four evaluations cost little and keep the register need small.
// ---------------------------------------------------------------------------------------
// pxyz (src/tracking.jl:717-720): sqrt(B^2 - px^2 - py^2), 0 where that is not positive
template <class T>
TP_DEV T pxyz_g(T B, T px, T py) {
  const T v = B * B - px * px - py * py;
  return val(v) > 0.0 ? gsqrt(v) : T(0.0);
}

// yrot (src/tracking.jl:727-745) as functions of (x, px, py, B = beti + delta): the new x
// and px and the increments of y and z
template <class T>
TP_DEV void yrot_g(T x, T px, T py, T B, double co, double si, T& xo, T& pxo, T& dy, T& dz) {
  const T pz = pxyz_g(B, px, py);
  const T p = co * pz - si * px;
  pxo = si * pz + co * px;
  xo = x * pz / p;
  dy = x * py * si / p;
  dz = B * x * si / p;
}

// bend_edge (src/tracking.jl:822-851), same outputs
template <class T>
TP_DEV void edge_g(T x, T px, T py, T B, double rhoinv, double co, double si, double s2t, double theta, T& xo,
                   T& pxo, T& dy, T& dz) {
  const T pz = pxyz_g(B, px, py);
  const T d2 = pxyz_g(B, T(0.0), py);
  pxo = px * co + (pz - rhoinv * x) * si;
  const T dasin = gasin_c(px / d2) - gasin_c(pxo / d2);
  const T num = x * (px * s2t + si * si * (2.0 * pz - rhoinv * x));
  const T den = pxyz_g(B, pxo, py) + pz * co - px * si;
  xo = x * co + num / den;
  dy = py * (theta / rhoinv + dasin / rhoinv);
  dz = B / rhoinv * (theta + dasin);
}

// The coefficients of bend_fringe (src/tracking.jl:752-815) as functions of (px, py, B)
template <class T>
TP_DEV void fringe_coef_g(T px, T py, T B, double b0, double gK, T& phi, T& dpx, T& dpy, T& dd) {
  const T pz = pxyz_g(B, px, py);
  const T xp = px / pz;
  const T yp = py / pz;
  const T term1 = 1.0 + xp * xp * (2.0 + yp * yp);
  const T term2 = xp / (1.0 + yp * yp);
  phi = -b0 * gtan(b0 * gK * term1 * pz - gatan(term2));
  const T px2 = px * px, py2 = py * py, pz2 = pz * pz;
  const T pz4 = pz2 * pz2;
  const T py2z2 = py2 + pz2;
  const T arg_sec = (b0 * gK * (pz4 + px2 * (py2 + 2.0 * pz2))) / (pz2 * pz) - gatan(px * pz / (py2 + pz2));
  const T secv = 1.0 / gcos(arg_sec);
  const T powsec = secv * secv;
  const T py4 = py2 * py2, py6 = py4 * py2, pz5 = pz4 * pz, pz6 = pz4 * pz2;
  const T py2z2_sq = py2z2 * py2z2;
  const T denom = pz5 * (py4 + px2 * pz2 + 2.0 * py2 * pz2 + pz4);
  dpx = -(b0 * (px2 * pz4 * (py2 - pz2) - pz6 * (py2 + pz2) +
                b0 * gK * px *
                    (pz2 * py2z2_sq * (2.0 * py2 + 3.0 * pz2) + px2 * px2 * (3.0 * py2 * pz2 + 2.0 * pz4) +
                     px2 * (3.0 * py6 + 8.0 * py4 * pz2 + 9.0 * py2 * pz4 + 5.0 * pz6)))) *
        powsec / denom;
  dpy = -(b0 * py *
          (px * pz4 * (py2 + pz2) +
           b0 * gK * (-(pz4 * py2z2_sq) + px2 * px2 * (3.0 * py2 * pz2 + 4.0 * pz4) +
                      px2 * (3.0 * py6 + 10.0 * py4 * pz2 + 11.0 * py2 * pz4 + 3.0 * pz6)))) *
        powsec / denom;
  dd = (b0 * B *
        (px * pz4 * (py2 - pz2) + b0 * gK *
                                      (-(pz4 * py2z2_sq) + px2 * px2 * (3.0 * py2 * pz2 + 2.0 * pz4) +
                                       px2 * (3.0 * py6 + 8.0 * py4 * pz2 + 7.0 * py2 * pz4 + pz6)))) *
       powsec / denom;
}

// bend_fringe as functions of (y, px, py, B): the increment of x, the new y, the decrement
// of py and the decrement of z
template <class T>
TP_DEV void fringe_g(T y, T px, T py, T B, double b0, double gK, T& dxf, T& yf, T& dpyf, T& dct) {
  T phi, dpx, dpy, dd;
  fringe_coef_g(px, py, B, b0, gK, phi, dpx, dpy, dd);
  yf = (2.0 * y) / (1.0 + gsqrt(1.0 - 2.0 * dpy * y));
  dxf = 0.5 * dpx * (yf * yf);
  dct = 0.5 * dd * (yf * yf);
  dpyf = phi * yf;
}

// State and adjoint of one point, returned together by the out-of-line maps below (they are
// called at most a few times per bend and turn; out of line, their temporaries do not raise
// the register allocation of the element loop).
struct SA {
  St s;
  Ad a;
};

// yrot backwards: s enters as the output and leaves as the input (yrot with -phi).  x and px
// are replaced, y and z incremented, py and delta pass through.
__device__ __noinline__ SA yrot_back(St s, Ad a, CxE c, double co, double si) {
  s = yrot(s, c, co, -si);
  double g0 = 0.0, g1 = 0.0, g2 = 0.0, g3 = 0.0;
#pragma nounroll
  for (int k = 0; k < 4; ++k) {
    D1 xo, pxo, dy, dz;
    yrot_g(D1(s.x, k == 0), D1(s.px, k == 1), D1(s.py, k == 2), D1(c.B, k == 3), co, si, xo, pxo, dy, dz);
    const double gk = a.x * xo.d + a.px * pxo.d + a.y * dy.d + a.z * dz.d;
    g0 = k == 0 ? gk : g0;
    g1 = k == 1 ? gk : g1;
    g2 = k == 2 ? gk : g2;
    g3 = k == 3 ? gk : g3;
  }
  a.x = g0;
  a.px = g1;
  a.py = a.py + g2;
  a.d = a.d + g3;
  return SA{s, a};
}

// bend_edge backwards (bend_edge with -theta), same structure as yrot
__device__ __noinline__ SA wedge_back(St s, Ad a, CxE c, double rhoinv, double co, double si, double s2t, double theta) {
  s = bend_edge(s, c, rhoinv, co, -si, -s2t, -theta);
  double g0 = 0.0, g1 = 0.0, g2 = 0.0, g3 = 0.0;
#pragma nounroll
  for (int k = 0; k < 4; ++k) {
    D1 xo, pxo, dy, dz;
    edge_g(D1(s.x, k == 0), D1(s.px, k == 1), D1(s.py, k == 2), D1(c.B, k == 3), rhoinv, co, si, s2t, theta, xo,
           pxo, dy, dz);
    const double gk = a.x * xo.d + a.px * pxo.d + a.y * dy.d + a.z * dz.d;
    g0 = k == 0 ? gk : g0;
    g1 = k == 1 ? gk : g1;
    g2 = k == 2 ? gk : g2;
    g3 = k == 3 ? gk : g3;
  }
  a.x = g0;
  a.px = g1;
  a.py = a.py + g2;
  a.d = a.d + g3;
  return SA{s, a};
}

// bend_fringe backwards: y is replaced, x incremented, py and z decremented; px and delta
// pass through.  Five fixed-point steps for the input py (each gains ~6 digits or more).
__device__ __noinline__ SA fringe_back(St s, Ad a, CxE c, double b0, double gK) {
  const double yf = s.y;
  double py = s.py, phi, dpx, dpy, dd;
#pragma nounroll
  for (int it = 0; it < 5; ++it) {
    fringe_coef_g(s.px, py, c.B, b0, gK, phi, dpx, dpy, dd);
    py = s.py + phi * yf;
  }
  fringe_coef_g(s.px, py, c.B, b0, gK, phi, dpx, dpy, dd);
  s.x = s.x - 0.5 * dpx * (yf * yf);
  s.y = yf - 0.5 * dpy * (yf * yf);
  s.py = py;
  s.z = s.z + 0.5 * dd * (yf * yf);
  double g0 = 0.0, g1 = 0.0, g2 = 0.0, g3 = 0.0;
#pragma nounroll
  for (int k = 0; k < 4; ++k) {
    D1 dxf, yo, dpyf, dct;
    fringe_g(D1(s.y, k == 0), D1(s.px, k == 1), D1(s.py, k == 2), D1(c.B, k == 3), b0, gK, dxf, yo, dpyf, dct);
    const double gk = a.x * dxf.d + a.y * yo.d - a.py * dpyf.d - a.z * dct.d;
    g0 = k == 0 ? gk : g0;
    g1 = k == 1 ? gk : g1;
    g2 = k == 2 ? gk : g2;
    g3 = k == 3 ? gk : g3;
  }
  a.y = g0;
  a.px = a.px + g1;
  a.py = a.py + g2;
  a.d = a.d + g3;
  return SA{s, a};
}

// ---------------------------------------------------------------------------------------
// r1 / r2 rotations (the R * r products of every magnet pass!, e.g. src/tracking.jl:387-391):
// r' = M r with M column-major (M[i + 6 j]); undone with the inverse Minv that the host
// computes once per table, and the VJP is a = M^T a'.
// ---------------------------------------------------------------------------------------
TP_DEV void mat6_back(St& s, Ad& a, const double* __restrict__ M, const double* __restrict__ Minv) {
  const double r[6] = {s.x, s.px, s.y, s.py, s.z, s.d};
  const double l[6] = {a.x, a.px, a.y, a.py, a.z, a.d};
  double o[6], g[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    double acc = Minv[i] * r[0], gacc = M[6 * i] * l[0];
#pragma unroll
    for (int j = 1; j < 6; ++j) {
      acc = acc + Minv[i + 6 * j] * r[j];
      gacc = gacc + M[j + 6 * i] * l[j];
    }
    o[i] = acc;
    g[i] = gacc;
  }
  s = St{o[0], o[1], o[2], o[3], o[4], o[5]};
  a = Ad{g[0], g[1], g[2], g[3], g[4], g[5]};
}

// ---------------------------------------------------------------------------------------
// The other element families (tp_extra.h; src/tracking.jl:529-605, 1076-1218, 1531-1663)
// backwards.  Inverses:
//   * Solenoid and the LBend body are flows of quadratic Hamiltonians (plus, for LBend, a
//     constant dispersion term): the map of -L undoes them in x, px, y, py (checked on the
//     oracle to the last bit / 1e-19), and z, which they change by a function of the other
//     coordinates, is restored by subtracting that function at the recovered input;
//   * Corrector, the Lorentz boosts and LongitudinalRFMap in closed form;
//   * YRotation by the rotation of -angle (the same group as yrot);
//   * Translation, the Gaussian beam-beam kicks, AccelCavity and the wakes are shears: they
//     add to some coordinates a function of coordinates they leave alone, so the input is
//     out - (M(out) - out);
//   * CrabCavity: its kick phase is taken at the entrance z, before the first half drift;
//     the entrance z is found by fixed-point steps through the undone kick and drift
//     (a contraction with factor ~ L |k V x| / 2).
// VJPs by dual numbers through the same map text (vjp6), except the crab's (by hand) and
// LBend's edges (edge_back); the wiggler's 2nd-order maps and radiation kicks one by one
// (wiggler_back).
// ---------------------------------------------------------------------------------------
template <class F>
TP_DEV Ad vjp6(const St& s, const Ad& a, F&& map) {
  double g0 = 0.0, g1 = 0.0, g2 = 0.0, g3 = 0.0, g4 = 0.0, g5 = 0.0;
#pragma nounroll
  for (int k = 0; k < 6; ++k) {
    const SD q{D1(s.x, k == 0), D1(s.px, k == 1), D1(s.y, k == 2), D1(s.py, k == 3), D1(s.z, k == 4), D1(s.d, k == 5)};
    const SD o = map(q);
    const double gk = a.x * o.x.d + a.px * o.px.d + a.y * o.y.d + a.py * o.py.d + a.z * o.z.d + a.d * o.d.d;
    g0 = k == 0 ? gk : g0;
    g1 = k == 1 ? gk : g1;
    g2 = k == 2 ? gk : g2;
    g3 = k == 3 ? gk : g3;
    g4 = k == 4 ? gk : g4;
    g5 = k == 5 ? gk : g5;
  }
  return Ad{g0, g1, g2, g3, g4, g5};
}

// Input of a shear map M from its output: out - (M(out) - out)
TP_DEV St unshear(const St& o, const St& m) {
  return St{2.0 * o.x - m.x, 2.0 * o.px - m.px, 2.0 * o.y - m.y, 2.0 * o.py - m.py, 2.0 * o.z - m.z, 2.0 * o.d - m.d};
}

// drift6 inverse only (no VJP), exact or linearised
TP_DEV St drift_inv(St s, double L, double Lb, double beti, bool lin) {
  if (lin) {
    const double N = L / (1.0 + s.d);
    s.x = s.x - N * s.px;
    s.y = s.y - N * s.py;
    s.z = s.z - N * (s.px * s.px + s.py * s.py) / (2.0 * (1.0 + s.d));
    return s;
  }
  Cx c;
  mkctx(s, c, beti);
  const double NormL = L / sqrt(c.A - s.px * s.px - s.py * s.py);
  s.x = s.x - NormL * s.px;
  s.y = s.y - NormL * s.py;
  s.z = s.z - NormL * c.B + Lb;
  return s;
}

// CrabCavity backwards (src/tracking.jl:1531-1547): [half drift] kick [half drift], the
// kick px += V sin(a), delta -= V' x cos(a) with a = (k z_entry + phi) beti.
TP_DEV SA crab_back(St s, Ad a, uint32_t fl, const double* __restrict__ p, double beti, bool lin) {
  const bool dr = (fl & TP_F_RF_DRIFT) != 0;
  Cx c;
  if (dr) {
    mkctx(s, c, beti);
    drift_back(s, a, c, p[0], p[1], lin);  // s: the kicked state
  }
  const St k = s;
  St s1 = k, s0 = k;
  double z0 = k.z, ph = 0.0;
#pragma nounroll
  for (int it = 0; it < (dr ? 6 : 1); ++it) {
    ph = (p[3] * z0 + p[4]) * beti;
    s1.px = k.px - p[2] * sin(ph);
    s1.d = k.d + p[5] * k.x * cos(ph);
    s0 = dr ? drift_inv(s1, p[0], p[1], beti, lin) : s1;
    z0 = s0.z;
  }
  // kick VJP at s1: d px'/d a = V cos(a) beti, d delta'/d a = V' x sin(a) beti, d delta'/dx
  const double a_ph = a.px * p[2] * cos(ph) * beti + a.d * p[5] * k.x * sin(ph) * beti;
  a.x = a.x - a.d * p[5] * cos(ph);
  if (dr) {
    mkctx(s1, c, beti);
    St t = s1;
    drift_back(t, a, c, p[0], p[1], lin);
  }
  a.z = a.z + a_ph * p[3];
  return SA{s0, a};
}

// Inverse of wig_map_2nd(dl) started at Zw_in: its shears (functions of x, y only) and drifts
// (functions of px or py and delta only) undone in reverse order.
TP_DEV St wig_map_2nd_inv(St s, const double* __restrict__ h, int nh, int nv, double dl, double Zw_in) {
  const double delta = s.d;
  const double inv1pd = 1.0 / (1.0 + delta);
  const double dld = dl * inv1pd;
  const double dl2 = dl / 2.0;
  const double dl2d = dl2 * inv1pd;
  const double Zw = Zw_in + dl2;
  double x = s.x, px = s.px, y = s.y, py = s.py, z = s.z, a, b;
  wig_ay(h, nh, nv, x, y, Zw, a, b);
  px = px - b;
  py = py - a;
  z = z - (dl2d / 2.0) * (py * py) * inv1pd;
  y = y - dl2d * py;
  wig_ay(h, nh, nv, x, y, Zw, a, b);
  px = px + b;
  py = py + a;
  wig_ax(h, nh, nv, x, y, Zw, a, b);
  px = px - a;
  py = py - b;
  z = z - (dld / 2.0) * (px * px) * inv1pd;
  x = x - dld * px;
  wig_ax(h, nh, nv, x, y, Zw, a, b);
  px = px + a;
  py = py + b;
  wig_ay(h, nh, nv, x, y, Zw, a, b);
  px = px - b;
  py = py - a;
  z = z - (dl2d / 2.0) * (py * py) * inv1pd;
  y = y - dl2d * py;
  wig_ay(h, nh, nv, x, y, Zw, a, b);
  px = px + b;
  py = py + a;
  return St{x, px, y, py, z, delta};
}

// Inverse of wig_rad_kick: 1 + delta' = u - K u^2 (u = 1 + delta, K = Crad B^2/H^2 dl) is
// solved for the root near u', u = 2u' / (1 + sqrt(1 - 4 K u')); px, py are rescaled by
// 1 - K u^2 between the ax / ay shears.
TP_DEV St wig_rad_kick_inv(St s, const double* __restrict__ p, const double* __restrict__ h, int nh, int nv,
                           double Zw) {
  double ax, axpy, ay, aypx;
  wig_ax(h, nh, nv, s.x, s.y, Zw, ax, axpy);
  wig_ay(h, nh, nv, s.x, s.y, Zw, ay, aypx);
  s.px = s.px - ax;
  s.py = s.py - ay;
  const double B2 = wig_b2(h, nh, nv, s.x, s.y, Zw);
  if (B2 != 0.0) {
    const double K = p[7] * (B2 / (p[6] * p[6])) * p[8];
    const double u1 = 1.0 + s.d;
    const double u = 2.0 * u1 / (1.0 + sqrt(1.0 - 4.0 * K * u1));
    const double scale = 1.0 - K * u * u;
    s.px = s.px / scale;
    s.py = s.py / scale;
    s.d = u - 1.0;
  }
  s.px = s.px + ax;
  s.py = s.py + ay;
  return s;
}

// Zw at the start of 2nd-order map j (0, 1, 2) of step k, accumulated as the forward pass does
TP_DEV double wig_zw(int k, int j, double dl1, double dl0) {
  double Zw = 0.0;
  for (int kk = 0; kk <= k; ++kk)
    for (int jj = 0; jj < 3; ++jj) {
      if (kk == k && jj == j) return Zw;
      const double dl2 = (jj == 1 ? dl0 : dl1) / 2.0;
      Zw = Zw + dl2;
      Zw = Zw + dl2;
    }
  return Zw;
}

// Wiggler (src/tracking.jl:1380-1525) backwards: every 2nd-order map and radiation kick undone
// in reverse order, each with the VJP of the same map text at its reconstructed input.
TP_DEV SA wiggler_back(St s, Ad a, int nstep, const double* __restrict__ p) {
  const double dl1 = p[1], dl0 = p[2];
  const int nh = (int)p[3], nv = (int)p[4];
  const bool rad = p[5] != 0.0;
  const double* h = p + 9;
  for (int k = nstep - 1; k >= 0; --k) {
    if (rad) {
      const double Zr = wig_zw(k + 1, 0, dl1, dl0);
      const St in = wig_rad_kick_inv(s, p, h, nh, nv, Zr);
      a = vjp6(in, a, [&](SD r) {
        wig_rad_kick(r, p, h, nh, nv, Zr);
        return r;
      });
      s = in;
    }
    for (int j = 2; j >= 0; --j) {
      const double dl = j == 1 ? dl0 : dl1;
      const double Zin = wig_zw(k, j, dl1, dl0);
      const St in = wig_map_2nd_inv(s, h, nh, nv, dl, Zin);
      a = vjp6(in, a, [&](SD r) {
        double z = Zin;
        wig_map_2nd(r, h, nh, nv, dl, z);
        return r;
      });
      s = in;
    }
  }
  if (rad) {
    const St in = wig_rad_kick_inv(s, p, h, nh, nv, 0.0);
    a = vjp6(in, a, [&](SD r) {
      wig_rad_kick(r, p, h, nh, nv, 0.0);
      return r;
    });
    s = in;
  }
  return SA{s, a};
}

__device__ __noinline__ SA extra_back(St s, Ad a, int tag, uint32_t fl, int n, const double* __restrict__ p,
                                      double beti, bool lin) {
  St in = s;
  switch (tag) {
    case TP_SOLENOID: {
      const double q[2] = {-p[0], p[1]};
      in = solenoid(s, q);
      St z0 = in;
      z0.z = 0.0;
      in.z = s.z - solenoid(z0, p).z;
      in.d = s.d;
      a = vjp6(in, a, [&](const SD& r) { return solenoid(r, p); });
      break;
    }
    case TP_CORRECTOR: {
      const double p_norm = 1.0 / (1.0 + s.d);
      const double NormL = p[0] * p_norm;
      in.px = s.px - p[1];
      in.py = s.py - p[2];
      in.x = s.x - NormL * (in.px + p[4]);
      in.y = s.y - NormL * (in.py + p[5]);
      St z0 = in;
      z0.z = 0.0;
      in.z = s.z - corrector(z0, p).z;
      a = vjp6(in, a, [&](const SD& r) { return corrector(r, p); });
      break;
    }
    case TP_TRANSLATION:
      in = unshear(s, translation(s, p, beti));
      a = vjp6(in, a, [&](const SD& r) { return translation(r, p, beti); });
      break;
    case TP_YROTATION: {
      const double q[3] = {p[0], -p[1], -p[2]};
      in = y_rotation(s, q, beti);
      a = vjp6(in, a, [&](const SD& r) { return y_rotation(r, p, beti); });
      break;
    }
    case TP_CRAB:
      return crab_back(s, a, fl, p, beti, lin);
    case TP_ACCEL:
      in = unshear(s, accel_kick(s, p));
      a = vjp6(in, a, [&](const SD& r) { return accel_kick(r, p); });
      break;
    case TP_LONGRF:
      in.z = s.z + p[0] * s.d;
      a = vjp6(in, a, [&](const SD& r) { return longrf_slip(r, p); });
      break;
    case TP_BOOST:
    case TP_INVBOOST: {
      const bool inv = tag == TP_INVBOOST;
      in.z = s.z / p[1];
      in.px = s.px / p[1];
      in.py = s.py / p[1];
      in.x = inv ? s.x + p[0] * in.z : s.x - p[0] * in.z;
      in.d = inv ? s.d - p[0] * in.px : s.d + p[0] * in.px;
      a = vjp6(in, a, [&](const SD& r) { return boost(r, p, inv); });
      break;
    }
    case TP_GAUSSBB:
      in = unshear(s, gauss_bb(s, n, p));
      a = vjp6(in, a, [&](const SD& r) { return gauss_bb(r, n, p); });
      break;
    case TP_RLCWAKE:
      in = unshear(s, rlc_wake(s, p));
      a = vjp6(in, a, [&](const SD& r) { return rlc_wake(r, p); });
      break;
    case TP_TABWAKE:
      in = unshear(s, tab_wake(s, n, p));
      a = vjp6(in, a, [&](const SD& r) { return tab_wake(r, n, p); });
      break;
    case TP_LBEND: {
      const double irho = p[4];
      edge_back(s, a, 1, p + 9, irho, true);
      const double q[5] = {-p[0], p[1], p[2], p[3], p[4]};
      in = lbend_body(s, q);
      St z0 = in;
      z0.z = 0.0;
      in.z = s.z - lbend_body(z0, p).z;
      in.d = s.d;
      a = vjp6(in, a, [&](const SD& r) { return lbend_body(r, p); });
      edge_back(in, a, 1, p + 5, irho, false);
      break;
    }
    case TP_WIGGLER:
      return wiggler_back(s, a, n, p);
    default:  // TP_NAN: a particle that passes it is lost, so none reaches it backwards
      break;
  }
  return SA{in, a};
}

}  // namespace tpk

#endif  // TP_ADJOINT_H
