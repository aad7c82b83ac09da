// tp_maps.h -- per-particle element maps for the fused gfx950 tracking kernels.
//
// Included by tp_track_kernels.hip, which is compiled twice:
//   TP_FAST=0  with -ffp-contract=off: the parity build.  Every map performs the
//              reference's floating-point operations in the reference's order, so for
//              + - * / sqrt it reproduces the reference bit for bit; per-particle
//              transcendentals (asin, sin, tan, atan, cos) come from the device math
//              library and agree to a few ulp.
//   TP_FAST=1  with -ffp-contract=fast: FMA contraction, reciprocal square roots by
//              a refined hardware estimate or (drifts at small angles) a series about
//              a hoisted 1/sqrt(A), multiplications by hoisted reciprocals and a series
//              asin for small arguments.  Validated against the parity build and the
//              oracle to tolerance (1e-12 relative after one turn).
//
// Per-particle state lives in registers; every element parameter is read from the pool
// at a wave-uniform address (scalar loads).  Quantities that depend only on delta
// (which changes only at RF cavities and misalignments) are hoisted into Cx; each hoist
// is the exact leading subexpression of the reference formula, so hoisting does not
// change a single bit.
#ifndef TP_MAPS_H
#define TP_MAPS_H

#include <math.h>
#include "tp_device.h"

#define TP_DEV __device__ __forceinline__

namespace tpk {

// Yoshida-4 constants, src/tracking.jl:29-32 (only used by the adjoint; the forward
// kernels read the writer's precomputed L1, L2, K1, K2).
struct St {
  double x, px, y, py, z, d;
};

// delta-dependent hoists.  A and B are the leading subexpressions of drift6's
// pz2 = ((1 + (2 delta) beti) + delta^2) - px^2 - py^2 and of (beti + delta)
// (src/tracking.jl:90, 98); BB = dp1^2 of pxyz (src/tracking.jl:718) with dp1 = beti+delta;
// C = 1 + delta of the linearised drift (src/tracking.jl:102).
// The fast build adds R = 1/sqrt(A), H = 1/(2A), H3 = 3/(4A) and WB = 2^-18 A for the
// drift's series form (inv_pz); the out-of-line maps take only the base (CxE), so the
// additions are not passed to them.
struct CxE {
  double A, B, BB, C;
};
struct Cx : CxE {
#if TP_FAST
  double R, H, H3, WB;
#endif
};

TP_DEV void mkctx(const St& s, Cx& c, double beti) {
  c.A = 1.0 + 2.0 * s.d * beti + s.d * s.d;
  c.B = beti + s.d;
  c.BB = c.B * c.B;
  c.C = 1.0 + s.d;
#if TP_FAST
  // correctly rounded: these are formed once per change of delta, not per drift
  c.R = 1.0 / sqrt(c.A);
  c.H = 0.5 / c.A;
  c.H3 = 1.5 * c.H;
  c.WB = c.A * 0x1p-18;  // exact
#endif
}

#if TP_FAST
// 1/sqrt(v): the hardware estimate y (relative error below 2^-24 on gfx950, measured over
// [0.5, 2]) refined by the binomial series of (1 - e)^(-1/2) = 1 + e/2 + 3e^2/8 + ...,
// e = 1 - v y^2 < 2^-23: the dropped terms are below 2^-67, so one step gives the double
// result within about one ulp (two Newton steps cost three more instructions).  v <= 0
// gives NaN without a test: rsq(v < 0) = NaN, rsq(+-0) = +-inf and then v y = 0 * inf = NaN.
TP_DEV double rsqrt_nr(double v) {
  const double y = __builtin_amdgcn_rsq(v);
  const double t = v * y;
  const double e = __builtin_fma(-t, y, 1.0);
  return __builtin_fma(y * e, __builtin_fma(e, 0.375, 0.5), y);
}
// L / pz of the exact-Hamiltonian drift, pz = sqrt(A - w), w = px^2 + py^2, A = the
// (1 + delta)^2 form hoisted in Cx.  With R = 1/sqrt(A),
//   L / pz = L R (1 - w/A)^(-1/2) = L R (1 + (1/2)(w/A) + (3/8)(w/A)^2 + (5/16)(w/A)^3 + ...),
// and for w < 2^-18 A the terms after (w/A)^2 sum to below (5/16) 2^-54 / (1 - 2^-18) <
// 2^-55.6 relative, a quarter of an ulp, so the quadratic
//   L / pz = Lr + w (c1 + w c2),  Lr = L R,  c1 = Lr H = L R / (2A),  c2 = c1 H3 = 3 L R / (8 A^2)
// gives L / pz within about two ulp from two FMAs; Lr, c1, c2 are per element and
// particle (hoisted out of a magnet's steps), and their roundings weigh in with a factor
// w/A < 2^-18.  |p| < 2^-9 rad covers the beams of the ring workloads (|p| ~ 1e-4); larger
// angles (and NaN, pz2 <= 0, where w >= A) take rsqrt_nr lane by lane, so each lane's
// result depends only on its own coordinates.
TP_DEV double inv_pz(const St& s, const Cx& c, double L) {
  const double w = __builtin_fma(s.px, s.px, s.py * s.py);
  const double Lr = L * c.R;
  const double c1 = Lr * c.H;
  const double c2 = c1 * c.H3;
  double NormL = __builtin_fma(w, __builtin_fma(w, c2, c1), Lr);
  if (__builtin_expect(!(w < c.WB), 0)) NormL = L * rsqrt_nr(c.A - w);
  return NormL;
}
// sqrt(v) = v / sqrt(v) = t (1 - e)^(-1/2) with t = v y, from the same estimate
TP_DEV double sqrt_nr(double v) {
  const double y = __builtin_amdgcn_rsq(v);
  const double t = v * y;
  const double e = __builtin_fma(-t, y, 1.0);
  return __builtin_fma(t * e, __builtin_fma(e, 0.375, 0.5), t);
}
// 1/v by two Newton-Raphson steps from the hardware estimate.
TP_DEV double rcp_nr(double v) {
  double y = __builtin_amdgcn_rcp(v);
  double e = 1.0 - v * y;
  y = y + y * e;
  e = 1.0 - v * y;
  y = y + y * e;
  return y;
}
#endif

// Outlined library calls: executed rarely (large arguments, once-per-element edge maps,
// one RF kick per turn), kept out of line so their temporaries do not raise the register
// allocation of the hot loop.
__device__ __noinline__ double asin_lib(double u) { return asin(u); }
__device__ __noinline__ double tan_lib(double u) { return tan(u); }
__device__ __noinline__ double sin_lib(double u) { return sin(u); }
__device__ __noinline__ double cos_lib(double u) { return cos(u); }

// asin(u) for |u| <= 1/8 by its Maclaurin series
//   asin(u) = sum_k C(2k,k) / (4^k (2k+1)) u^(2k+1),
// truncated after k = 8 (remainder below 2^-60 |u|, so the result is within rounding
// of the true value); larger |u| calls the library asin.
TP_DEV double asin_s(double u) {
  const double w = u * u;
  double p = 6435.0 / 557056.0;
  p = p * w + 143.0 / 10240.0;
  p = p * w + 231.0 / 13312.0;
  p = p * w + 63.0 / 2816.0;
  p = p * w + 35.0 / 1152.0;
  p = p * w + 5.0 / 112.0;
  p = p * w + 3.0 / 40.0;
  p = p * w + 1.0 / 6.0;
  const double r = u + (u * w) * p;
  if (__builtin_expect(!(fabs(u) <= 0.125), 0)) return asin_lib(u);
  return r;
}

TP_DEV double clamp1(double v) { return v > 1.0 ? 1.0 : (v < -1.0 ? -1.0 : v); }

#if TP_FAST
// asin(a) - asin(b) of the exact bend body, a = px/d2, b = pxn/d2, whose cosines are
// sqrt(1 - a^2) = pz/d2 and sqrt(1 - b^2) = pzn/d2 (pz, pzn are the body's pxyz values):
// the difference has the sine u = (px pzn - pxn pz)/d2^2.  When |u| <= 1/64 and
// pz + pzn > d2/16 the difference is asin(u) itself -- the other root, near +-pi, needs
// both cosines below sin(asin(1/64)) = 1/64 -- and the series through u^9 gives it (first
// dropped term 63/2816 u^11 < 2^-65 |u|).  Otherwise (large angles, clamped arguments) the
// reference's two clamped arcsines.
TP_DEV double asin_diff(double px, double pz, double pxn, double pzn, double d2, double id2) {
  const double u = (px * pzn - pxn * pz) * (id2 * id2);
  if (__builtin_expect(fabs(u) <= 0.015625 && pz + pzn > 0.0625 * d2, 1)) {
    const double w = u * u;
    double p = __builtin_fma(w, 35.0 / 1152.0, 5.0 / 112.0);
    p = __builtin_fma(p, w, 3.0 / 40.0);
    p = __builtin_fma(p, w, 1.0 / 6.0);
    return __builtin_fma(u * w, p, u);
  }
  return asin_s(clamp1(px * id2)) - asin_s(clamp1(pxn * id2));
}
#endif

TP_DEV double nan_d() { return __builtin_nan(""); }

// ---------------------------------------------------------------------------------------
// drift6, src/tracking.jl:87-108
// ---------------------------------------------------------------------------------------
template <bool EXH, bool SER = true>
TP_DEV void drift(St& s, const Cx& c, double L, double Lb) {
  if (EXH) {
#if TP_FAST
    // NaN when pz2 <= 0 (inv_pz), so x, y, z become NaN and the particle is lost at
    // this element's exit with its entry coordinates, as after the reference's NaN map.
    // (Yoshida steps use drift_xy and sum the path length per step.)
    // The path-length change NormL B - L beti is formed with one rounding.
    const double NormL = SER ? inv_pz(s, c, L) : L * rsqrt_nr(c.A - s.px * s.px - s.py * s.py);
    s.x = s.x + NormL * s.px;
    s.y = s.y + NormL * s.py;
    s.z = s.z + __builtin_fma(NormL, c.B, -Lb);
#else
    double pz2 = c.A - s.px * s.px - s.py * s.py;
    double NormL = L / sqrt(pz2);
    NormL = (pz2 <= 0.0) ? nan_d() : NormL;  // reference returns NaN x 6 (:91-94)
    s.x = s.x + NormL * s.px;
    s.y = s.y + NormL * s.py;
    s.z = s.z + NormL * c.B - Lb;
#endif
  } else {
    double NormL = L / c.C;
    s.x = s.x + NormL * s.px;
    s.y = s.y + NormL * s.py;
    s.z = s.z + NormL * (s.px * s.px + s.py * s.py) / (2.0 * c.C);
  }
}

#if TP_FAST
// drift6 (exact Hamiltonian) without its path-length update: returns NormL, NaN for
// pz2 <= 0 (x and y become NaN).  A Yoshida step adds the four drifts' path lengths at once,
// z += (beti + delta) sum(NormL) - sum(L beti), one FMA per step instead of two operations
// per drift; the one rounding of the step's sum is below the reference's own rounding of
// z + NormL (beti + delta) at every drift.
template <bool SER>
TP_DEV double drift_xy(St& s, const Cx& c, double L) {
  const double NormL = SER ? inv_pz(s, c, L) : L * rsqrt_nr(c.A - s.px * s.px - s.py * s.py);
  s.x = s.x + NormL * s.px;
  s.y = s.y + NormL * s.py;
  return NormL;
}
// Two drifts of length L with nothing between them (the last half-drift of a Yoshida step
// and the first of the next): the momenta, and so NormL, are the same for both, and
// x + 2 NormL px replaces (x + NormL px) + NormL px with one rounding less.  Returns the NormL
// of one drift.
template <bool SER>
TP_DEV double drift_xy_pair(St& s, const Cx& c, double L) {
  const double NormL = SER ? inv_pz(s, c, L) : L * rsqrt_nr(c.A - s.px * s.px - s.py * s.py);
  const double N2 = NormL + NormL;
  s.x = __builtin_fma(N2, s.px, s.x);
  s.y = __builtin_fma(N2, s.py, s.y);
  return NormL;
}

// Rotation by an angle as its (cos, sin) pair; rot_mul composes two (the angles add).
struct Rot {
  double c, s;
};
TP_DEV Rot rot_mul(const Rot& u, const Rot& v) {
  return Rot{__builtin_fma(u.c, v.c, -(u.s * v.s)), __builtin_fma(u.s, v.c, u.c * v.s)};
}
// The rotation of a pure-dipole exact bend's whole body: each Yoshida step rotates by
// 2 phi1 + 2 phi2 (sub-bodies L1 L2 L2 L1, cos/sin in the record), ns steps in all; the
// power by repeated squaring.  Wave-uniform values.
TP_DEV Rot rot_body(double c1, double s1, double c2, double s2, int ns) {
  const Rot h = rot_mul(Rot{c1, s1}, Rot{c2, s2});
  Rot st = rot_mul(h, h);
  Rot r{1.0, 0.0};
  for (int n = ns; n > 0; n >>= 1) {
    if (n & 1) r = rot_mul(r, st);
    if (n > 1) st = rot_mul(st, st);
  }
  return r;
}
#endif

// ---------------------------------------------------------------------------------------
// Horner sums of strthinkick / bndthinkick, src/tracking.jl:251-259, 290-298.
// poly is interleaved (b0, a0, b1, a1, ...).
// ---------------------------------------------------------------------------------------
template <int KK>
TP_DEV void horner(const St& s, const double* __restrict__ poly, int m, double& re, double& im) {
  if (KK == TP_KICK_NONE) {
    re = 0.0;
    im = 0.0;
  } else if (KK == TP_KICK_QUAD) {
    const double b1 = poly[2];
    re = b1 * s.x;
    im = b1 * s.y;
  } else if (KK == TP_KICK_SEXT) {
    const double b2 = poly[4];
    const double rx = b2 * s.x, iy = b2 * s.y;
    re = rx * s.x - iy * s.y;
    im = iy * s.x + rx * s.y;
  } else {
    double ReSum = poly[2 * m];
    double ImSum = poly[2 * m + 1];
    for (int i = m - 1; i >= 0; --i) {
      double t = ReSum * s.x - ImSum * s.y + poly[2 * i];
      ImSum = ImSum * s.x + ReSum * s.y + poly[2 * i + 1];
      ReSum = t;
    }
    re = ReSum;
    im = ImSum;
  }
}

// Horner sum of a fixed order M on coefficients already in registers, cf = (b0, a0, ...,
// bM, aM): the loop of horner() unrolled, so the 2M+2 coefficients are loaded once per
// element instead of once per kick and per order.  The parity build keeps the reference's
// expression; the fast build forms each line with two FMAs.
template <int M>
TP_DEV void horner_cf(const St& s, const double* cf, double& re, double& im) {
  double ReSum = cf[2 * M], ImSum = cf[2 * M + 1];
#pragma unroll
  for (int i = M - 1; i >= 0; --i) {
#if TP_FAST
    const double t = __builtin_fma(ReSum, s.x, __builtin_fma(-ImSum, s.y, cf[2 * i]));
    ImSum = __builtin_fma(ImSum, s.x, __builtin_fma(ReSum, s.y, cf[2 * i + 1]));
#else
    const double t = ReSum * s.x - ImSum * s.y + cf[2 * i];
    ImSum = ImSum * s.x + ReSum * s.y + cf[2 * i + 1];
#endif
    ReSum = t;
  }
  re = ReSum;
  im = ImSum;
}

template <int M>
TP_DEV void kick_cf(St& s, const double* cf, double K) {
  double re, im;
  horner_cf<M>(s, cf, re, im);
  s.px = s.px - K * re;
  s.py = s.py + K * im;
}

// strthinkick, src/tracking.jl:284-305
template <int KK>
TP_DEV void kick(St& s, const double* __restrict__ poly, int m, double K) {
  if (KK == TP_KICK_NONE) return;
  double re, im;
  horner<KK>(s, poly, m, re, im);
  s.px = s.px - K * re;
  s.py = s.py + K * im;
}

#if TP_FAST
// strthinkick of the sextupole specialisation with the step's kick factor folded into its
// one coefficient, Kb = K b2 (formed once per element): px -= Kb (x^2 - y^2) and
// py += 2 Kb x y as two products and four FMAs.  The rounding of Kb is a fixed relative
// error of at most 2^-53 in the sextupole strength, far below the rounding of px itself
// since the kick is second order in x.  (The same folding for quadrupoles makes a fixed
// error in the linear focusing, whose phase error grows coherently over turns: measured
// on the FODO ring it triples the 100-turn error, so quadrupoles keep K (b1 x).)
TP_DEV void kick_sext_kb(St& s, double Kb) {
  const double rx = Kb * s.x, iy = Kb * s.y;
  s.px = __builtin_fma(-rx, s.x, s.px);
  s.px = __builtin_fma(iy, s.y, s.px);
  s.py = __builtin_fma(iy, s.x, s.py);
  s.py = __builtin_fma(rx, s.y, s.py);
}
#endif

// bndthinkick, src/tracking.jl:243-270 (Ki = K*irho, the reference's leading L*irho)
template <int KK>
TP_DEV void bkick(St& s, const double* __restrict__ poly, int m, double K, double Ki, double irho, double beti) {
  double re, im;
  horner<KK>(s, poly, m, re, im);
  s.px = s.px - K * (re - (s.d - s.x * irho) * irho);
  if (KK != TP_KICK_NONE) s.py = s.py + K * im;
  s.z = s.z + Ki * s.x * beti;
}

// ---------------------------------------------------------------------------------------
// Linear dipole edge, src/tracking.jl:160-221.  blk = [fx, fy_const, e, fc].
// ---------------------------------------------------------------------------------------
// Memo of the fringe term tan(e - fc / (1 + delta)) of edge methods 1 and 2: the bends of a
// ring usually share e and fc, and delta changes at few elements per turn, so each particle
// keeps its last value with the delta it was formed from (in LDS, one slot per lane, the
// lane's own), and the wave keeps the bit patterns of the (e, fc) the slots belong to.  A
// hit returns the bits the call would return: the same function of the same arguments.
// Every lane of the wave passes every edge (dead lanes too), so the slots of all lanes
// always belong to the wave's key.
struct EdgeMemo {
  uint64_t ke, kf;         // bits of e and fc (wave-uniform)
  double* __restrict__ w;  // the wave's slots for this particle (wave-uniform address): w[lane]
                           // delta of the lane's value (NaN: empty), w[64 + lane] the value
};

TP_DEV int lane_id() { return (int)__builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }
// The same, formed where it is used: the compiler cannot hoist it out of the element loop,
// where a lane index held across the loop costs a register the hot paths need.
TP_DEV int lane_id_here() {
  int l;
  asm volatile("v_mbcnt_lo_u32_b32 %0, -1, 0\n\tv_mbcnt_hi_u32_b32 %0, -1, %0" : "=v"(l));
  return l;
}

TP_DEV double edge_tan(EdgeMemo& mm, const St& s, double e, double fc) {
  const uint64_t be = __builtin_bit_cast(uint64_t, e), bf = __builtin_bit_cast(uint64_t, fc);
  const bool same = be == mm.ke && bf == mm.kf;
  mm.ke = be;
  mm.kf = bf;
  double* __restrict__ slot = mm.w + lane_id_here();
  double t;
  if (same && slot[0] == s.d) {
    t = slot[64];
  } else {
    t = tan_lib(e - fc / (1.0 + s.d));
    slot[0] = s.d;
    slot[64] = t;
  }
  return t;
}

template <bool MEMO = false>
TP_DEV void edge_lin(St& s, int method, const double* __restrict__ blk, double irho, bool is_exit,
                     EdgeMemo* mm = nullptr) {
  const double fx = blk[0], fyc = blk[1], e = blk[2], fc = blk[3];
  double fy;
  if (method == 2) {
    double t = (fc == 0.0) ? fyc : irho * (MEMO ? edge_tan(*mm, s, e, fc) : tan_lib(e - fc / (1.0 + s.d)));
    fy = t / (1.0 + s.d);
  } else if (method == 3) {
    fy = is_exit ? irho * tan_lib(e - fc - s.px / (1.0 + s.d)) : irho * tan_lib(e - fc + s.px / (1.0 + s.d));
  } else {  // Brown (1) and the fallback
    fy = (fc == 0.0) ? fyc : irho * (MEMO ? edge_tan(*mm, s, e, fc) : tan_lib(e - fc / (1.0 + s.d)));
  }
  const double px_new = s.px + s.x * fx;
  s.py = s.py - s.y * fy;
  s.px = px_new;
}

// ---------------------------------------------------------------------------------------
// Exact-bend helpers (Forest), src/tracking.jl:717-894
// ---------------------------------------------------------------------------------------
// pxyz(dp1, px, py) with dp1^2 = c.BB, src/tracking.jl:717-720
TP_DEV double pxyz(const CxE& c, double px, double py) {
  double v = c.BB - px * px - py * py;
#if TP_FAST
  // sqrt_nr is NaN for v <= 0 (and for NaN v), and fmax returns its other operand for a
  // NaN: 0, as the reference's test v > 0 gives (one v_max instead of a compare and selects)
  return __builtin_fmax(sqrt_nr(v), 0.0);
#else
  return v > 0.0 ? sqrt(v) : 0.0;
#endif
}


// yrot, src/tracking.jl:727-745 (caller skips phi == 0)
__device__ __noinline__ St yrot(St s, CxE c, double co, double si) {
  const double pz = pxyz(c, s.px, s.py);
  const double p = co * pz - si * s.px;
  const double px_new = si * pz + co * s.px;
  const double x_new = s.x * pz / p;
  s.y = s.y + s.x * s.py * si / p;
  s.z = s.z + c.B * s.x * si / p;
  s.x = x_new;
  s.px = px_new;
  return s;
}

// bend_fringe (Forest 13.13), src/tracking.jl:752-815
__device__ __noinline__ St bend_fringe(St s, CxE c, double irho, double gK) {
  const double b0 = irho;
  const double dp1 = c.B;
  const double pz = pxyz(c, s.px, s.py);
  const double px = s.px;
  const double py = s.py;
  const double xp = px / pz;
  const double yp = py / pz;
  const double term1 = 1.0 + xp * xp * (2.0 + yp * yp);
  const double term2 = xp / (1.0 + yp * yp);
  const double phi = -b0 * tan(b0 * gK * term1 * pz - atan(term2));
  const double px2 = px * px;
  const double py2 = py * py;
  const double pz2 = pz * pz;
  const double pz4 = pz2 * pz2;
  const double py2z2 = (py2 + pz2);
  const double arg_sec = (b0 * gK * (pz4 + px2 * (py2 + 2.0 * pz2))) / (pz2 * pz) - atan(px * pz / (py2 + pz2));
  const double secv = 1.0 / cos(arg_sec);
  const double powsec = secv * secv;
  const double py4 = py2 * py2;
  const double py6 = py4 * py2;
  const double pz5 = pz4 * pz;
  const double pz6 = pz4 * pz2;
  const double py2z2_sq = py2z2 * py2z2;
  const double denom = pz5 * (py4 + px2 * pz2 + 2.0 * py2 * pz2 + pz4);
  const double dpx_num = -(b0 * (px2 * pz4 * (py2 - pz2) - pz6 * (py2 + pz2) +
                                 b0 * gK * px * (pz2 * py2z2_sq * (2.0 * py2 + 3.0 * pz2) +
                                                 px2 * px2 * (3.0 * py2 * pz2 + 2.0 * pz4) +
                                                 px2 * (3.0 * py6 + 8.0 * py4 * pz2 + 9.0 * py2 * pz4 + 5.0 * pz6)))) *
                         powsec;
  const double dpx = dpx_num / denom;
  const double dpy_num = -(b0 * py * (px * pz4 * (py2 + pz2) +
                                      b0 * gK * (-(pz4 * py2z2_sq) + px2 * px2 * (3.0 * py2 * pz2 + 4.0 * pz4) +
                                                 px2 * (3.0 * py6 + 10.0 * py4 * pz2 + 11.0 * py2 * pz4 + 3.0 * pz6)))) *
                         powsec;
  const double dpy = dpy_num / denom;
  const double dd_num = (b0 * dp1 * (px * pz4 * (py2 - pz2) + b0 * gK *
                                     (-(pz4 * py2z2_sq) + px2 * px2 * (3.0 * py2 * pz2 + 2.0 * pz4) +
                                      px2 * (3.0 * py6 + 8.0 * py4 * pz2 + 7.0 * py2 * pz4 + pz6)))) *
                        powsec;
  const double dd = dd_num / denom;
  const double yf = (2.0 * s.y) / (1.0 + sqrt(1.0 - 2.0 * dpy * s.y));
  const double dxf = 0.5 * dpx * (yf * yf);
  const double dct = 0.5 * dd * (yf * yf);
  const double dpyf = phi * yf;
  s.x = s.x + dxf;
  s.y = yf;
  s.py = s.py - dpyf;
  s.z = s.z - dct;
  return s;
}

// bend_edge (Forest 12.41), src/tracking.jl:822-851 (caller skips |rhoinv| < 1e-6).
__device__ __noinline__ St bend_edge(St s, CxE c, double rhoinv, double co, double si, double s2t, double theta) {
  const double pz = pxyz(c, s.px, s.py);
  const double d2 = pxyz(c, 0.0, s.py);
  const double px_new = s.px * co + (pz - rhoinv * s.x) * si;
  const double val1 = clamp1(s.px / d2);
  const double val2 = clamp1(px_new / d2);
  const double dasin = asin_s(val1) - asin_s(val2);
  const double num = s.x * (s.px * s2t + si * si * (2.0 * pz - rhoinv * s.x));
  const double den = pxyz(c, px_new, s.py) + pz * co - s.px * si;
  const double x_new = s.x * co + num / den;
  s.y = s.y + s.py * (theta / rhoinv + dasin / rhoinv);
  s.z = s.z + c.B / rhoinv * (theta + dasin);
  s.x = x_new;
  s.px = px_new;
  return s;
}

// exact_bend_body (Forest 12.18), src/tracking.jl:858-894, for |irho| >= 1e-6.
// d2 = pxyz(dp1, 0, py) is passed in: py is unchanged by the body, and by the kicks
// between bodies when the kick is the identity, so the caller may hoist it.
// pz enters as pxyz(dp1, px, py) of the input and leaves as that of the output (pzn): the
// next body of a Yoshida step with identity kicks starts from exactly these arguments, so
// the caller passes it on instead of evaluating pxyz again (the same value bit for bit).
// rho = 1/irho and id2 = 1/d2 are used by the fast build only.  Returns the path term
// dasin; with YZ false the body leaves y and z to the caller (the fast build sums dasin over
// a Yoshida step of identity kicks, where py is constant: y += py sum(dasin) and
// z += (beti + delta) sum(dasin) - sum(L beti)).
template <bool YZ = true>
TP_DEV double ebody(St& s, const Cx& c, double irho, double rho, double L, double Lb, double cs, double sn,
                    double d2, double id2, double& pz_io) {
  const double pz = pz_io;
  const double pzmx = pz - (1.0 + s.x * irho);
  const double px_new = s.px * cs + pzmx * sn;
  const double pzn = pxyz(c, px_new, s.py);
#if TP_FAST
  // one series arcsine of the difference instead of two (asin_diff)
  const double dif = asin_diff(s.px, pz, px_new, pzn, d2, id2);
  const double dasin = L + dif * rho;
  const double x_new = (pzn - pzmx * cs + s.px * sn - 1.0) * rho;
#else
  const double val1 = clamp1(s.px / d2);
  const double val2 = clamp1(px_new / d2);
  const double dasin = L + (asin_s(val1) - asin_s(val2)) / irho;
  const double x_new = (pzn - pzmx * cs + s.px * sn - 1.0) / irho;
  (void)rho;
  (void)id2;
#endif
  if constexpr (YZ) {
    s.y = s.y + s.py * dasin;
#if TP_FAST
    s.z = s.z + __builtin_fma(c.B, dasin, -Lb);
#else
    s.z = s.z + c.B * dasin - Lb;
#endif
  }
  s.x = x_new;
  s.px = px_new;
  pz_io = pzn;
  return dasin;
}

// Drift limit of exact_bend_body (|irho| < 1e-6), src/tracking.jl:862-872
TP_DEV void ebody_driftlimit(St& s, const Cx& c, double L, double Lb) {
  const double pz = pxyz(c, s.px, s.py);
  const double NormL = L / pz;
  s.x = s.x + s.px * NormL;
  s.y = s.y + s.py * NormL;
  s.z = s.z + NormL * c.B - Lb;
}

// Sum over the 64 lanes of a wave (butterfly; every lane gets the same total, formed in
// the same order whatever the wave's position in the grid).
// One DPP move of a double (two 32-bit halves): lanes whose source is outside the row, or
// whose row is masked off, read zero.
template <int CTRL, int ROWS>
TP_DEV double dpp_f64(double v) {
  const uint64_t u = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)u, CTRL, ROWS, 0xf, true);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(u >> 32), CTRL, ROWS, 0xf, true);
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}
// Sum over the 64 lanes of a wave (every lane active), returned to every lane: sums within
// quads (quad_perm xor 1, xor 2), within rows of 16 (row_shr 4, 8: lane 15 of each row holds
// the row's sum), across rows (row_bcast 15 into rows 1 and 3, row_bcast 31 into rows 2 and
// 3: lane 63 holds the total), then lane 63 read back as a wave-uniform value.  A fixed
// order of additions, so the result is bit-reproducible.
TP_DEV double wave_sum(double v) {
  v = v + dpp_f64<0xb1, 0xf>(v);   // quad_perm [1, 0, 3, 2]
  v = v + dpp_f64<0x4e, 0xf>(v);   // quad_perm [2, 3, 0, 1]
  v = v + dpp_f64<0x114, 0xf>(v);  // row_shr 4
  v = v + dpp_f64<0x118, 0xf>(v);  // row_shr 8
  v = v + dpp_f64<0x142, 0xa>(v);  // row_bcast 15
  v = v + dpp_f64<0x143, 0xc>(v);  // row_bcast 31
  const uint64_t u = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, 63);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), 63);
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// Whether an element changes delta, and with it the delta hoists in Cx: the RF kick, the
// other families (their maps end with mkctx) and misaligned elements (t1 / t2 / r1 / r2 may
// move delta).  Wave-uniform: a test of the record on the scalar unit.  The element loops run
// such elements outside their hot loop, where the hoists are invariant: renewing them in the
// loop's own dispatch makes every iteration move them out of and back into their registers.
template <unsigned FEAT>
TP_DEV bool changes_delta(const tp_rec& rec) {
  if ((FEAT & FT_RF) != 0 && rec.tag == TP_RFCAV) return (rec.flags & TP_F_RF_KICK) != 0;
  if ((FEAT & FT_EXTRA) != 0 && rec.tag >= TP_SOLENOID) return true;
  if ((FEAT & FT_MISALIGN) != 0 && (rec.flags & TP_F_MISALIGN) != 0) return true;
  return false;
}

}  // namespace tpk

#endif  // TP_MAPS_H
