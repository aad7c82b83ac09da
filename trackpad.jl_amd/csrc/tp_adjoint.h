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

#include "tp_maps.h"

namespace tpk {

struct Ad {
  double x, px, y, py, z, d;
};

// ---------------------------------------------------------------------------------------
// drift6 (exact Hamiltonian): inverse + VJP.  s is the drift's output on entry and its
// input on return.  NormL depends only on px, py, delta, which the drift leaves alone.
// ---------------------------------------------------------------------------------------
TP_DEV void drift_back(St& s, Ad& a, const Cx& c, double L, double Lb) {
  const double pz2 = c.A - s.px * s.px - s.py * s.py;
#if TP_FAST
  const double rz = rsqrt_nr(pz2);
  const double NormL = L * rz;
  const double ipz2 = rz * rz;
#else
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
  if (KK == TP_KICK_NONE) return;
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
// Exact sector-bend body (Forest 12.18, src/tracking.jl:858-894) backwards.  s enters as
// the body's output and leaves as its input.  The inverse is the body of length -L (sn ->
// -sn); its intermediates are the forward map's: the inverse's pz is the forward output's
// pzn, its px_new is the input px and its pzn the input pz, and its path term is minus the
// forward dasin.  The VJP at the input reuses them instead of re-evaluating the map.
// rho = 1/irho and id2 = 1/d2 are formed by the caller once per element (d2 = pxyz(dp1, 0,
// py) is constant through the body).  pz_io enters as pxyz(dp1, px, py) of the output and
// leaves as that of the input, which is the output pz of the body before it when the kick
// between them is the identity.
// ---------------------------------------------------------------------------------------
TP_DEV void ebody_back(St& s, Ad& a, const Cx& c, double irho, double rho, double L, double Lb, double cs, double sn,
                       double d2, double id2, double& pz_io) {
  const double B = c.B;
  const double py = s.py;
  const double pxn = s.px;                       // output px
  const double pzn = pz_io;                      // output pz
  const double pzmx_o = pzn - (1.0 + s.x * irho);
  const double px = pxn * cs - pzmx_o * sn;      // input px
  const double pz = pxyz(c, px, py);             // input pz
#if TP_FAST
  // asin(px/d2) - asin(pxn/d2) as one series asin, as in the forward body
  const double dasin = L + asin_diff(px, pz, pxn, pzn, d2, id2) * rho;
  const double x_in = (pz - pzmx_o * cs - pxn * sn - 1.0) * rho;
  const double ipz = rcp_nr(pz), ipzn = rcp_nr(pzn);
#else
  const double dasin = L + (asin_s(clamp1(px / d2)) - asin_s(clamp1(pxn / d2))) / irho;
  const double x_in = (pz - pzmx_o * cs - pxn * sn - 1.0) / irho;
  const double ipz = 1.0 / pz, ipzn = 1.0 / pzn;
  (void)id2;
#endif
  const double iirho = rho;
  // partials of the forward map at the input with respect to (x, px, py, delta); y and z
  // enter the outputs with coefficient 1
  const double dpz_px = -px * ipz, dpz_py = -py * ipz, dpz_d = B * ipz;
  const double dpxn_x = -irho * sn, dpxn_px = cs + sn * dpz_px, dpxn_py = sn * dpz_py, dpxn_d = sn * dpz_d;
  const double dpzn_x = -pxn * dpxn_x * ipzn, dpzn_px = -pxn * dpxn_px * ipzn;
  const double dpzn_py = (-pxn * dpxn_py - py) * ipzn, dpzn_d = (B - pxn * dpxn_d) * ipzn;
  // x' irho = pzn - pzmx cs + px sn - 1, with d pzmx = (-irho, dpz_px, dpz_py, dpz_d)
  const double dx_x = (dpzn_x + cs * irho) * iirho;
  const double dx_px = (dpzn_px - cs * dpz_px + sn) * iirho;
  const double dx_py = (dpzn_py - cs * dpz_py) * iirho;
  const double dx_d = (dpzn_d - cs * dpz_d) * iirho;
  // dasin = L + (asin(px/d2) - asin(pxn/d2)) / irho, d asin(u) = du / sqrt(1-u^2),
  // sqrt(1 - (px/d2)^2) = pz/d2 and sqrt(1 - (pxn/d2)^2) = pzn/d2; d2 = sqrt(B^2 - py^2)
  const double dd2_py = -py * id2, dd2_d = B * id2;
  const double da_x = (-(dpxn_x * ipzn)) * iirho;
  const double da_px = (ipz - dpxn_px * ipzn) * iirho;
  const double da_py = ((-px * dd2_py * id2) * ipz - (dpxn_py - pxn * dd2_py * id2) * ipzn) * iirho;
  const double da_d = ((-px * dd2_d * id2) * ipz - (dpxn_d - pxn * dd2_d * id2) * ipzn) * iirho;
  // outputs: x', px', y' = y + py dasin, z' = z + B dasin - L beti
  const double ayz = a.y * py + a.z * B;
  const double ax = a.x * dx_x + a.px * dpxn_x + ayz * da_x;
  const double apx = a.x * dx_px + a.px * dpxn_px + ayz * da_px;
  const double apy = a.py + a.x * dx_py + a.px * dpxn_py + ayz * da_py + a.y * dasin;
  const double ad = a.d + a.x * dx_d + a.px * dpxn_d + ayz * da_d + a.z * dasin;
  a.x = ax;
  a.px = apx;
  a.py = apy;
  a.d = ad;
  s.y = s.y - py * dasin;
#if TP_FAST
  s.z = s.z - __builtin_fma(B, dasin, -Lb);
#else
  s.z = s.z - B * dasin + Lb;
#endif
  s.x = x_in;
  s.px = px;
  pz_io = pz;
}

// ---------------------------------------------------------------------------------------
// Linear dipole edge (src/tracking.jl:160-221): inverse + VJP.  blk = [fx, fyc, e, fc].
// ---------------------------------------------------------------------------------------
TP_DEV void edge_fy(const St& s, int method, const double* __restrict__ blk, double irho, bool is_exit, double& fy,
                    double& dfy_px, double& dfy_d) {
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
    const double tt = tan_lib(e - fc / opd);
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

TP_DEV void edge_back(St& s, Ad& a, int method, const double* __restrict__ blk, double irho, bool is_exit) {
  const double fx = blk[0];
  s.px = s.px - s.x * fx;
  double fy, dfy_px, dfy_d;
  edge_fy(s, method, blk, irho, is_exit, fy, dfy_px, dfy_d);
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

}  // namespace tpk

#endif  // TP_ADJOINT_H
