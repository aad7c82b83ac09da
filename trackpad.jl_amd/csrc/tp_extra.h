// tp_extra.h -- device maps of the element families outside the hot lattices: solenoid,
// corrector, translation, y-rotation, crab / accelerating cavities, the longitudinal RF
// map, Lorentz boosts, strong Gaussian beam-beam kicks, longitudinal wakes, the linear bend
// and the NaN map (parameter blocks in include/trackpad_table.h).
//
// Every map restates the reference formula in its operation order (src/tracking.jl:529-605,
// 1076-1218, 1531-1663; Julia evaluates a*b*c as (a*b)*c and x^2 as x*x) so that the parity
// build reproduces it.  The maps are compiled only into the kernel variant that holds every
// feature, and they run in one out-of-line function, so their registers and transcendentals
// do not weigh on the hot element loop.  The maps of the families the reverse pass
// covers are written over the state type S (St, or SD of dual numbers: tp_dual.h) so that
// the same text gives the forward map and its Jacobian columns.
#ifndef TP_EXTRA_H
#define TP_EXTRA_H

#include "tp_dual.h"

namespace tpk {

TP_DEV St nan_state() {
  const double n = nan_d();
  return St{n, n, n, n, n, n};
}
TP_DEV SD nan_state(SD) {
  const double n = nan_d();
  return SD{n, n, n, n, n, n};
}
TP_DEV St nan_state(St) { return nan_state(); }

// pz of Translation / YRotation: sqrt(1 + 2 delta beti + delta^2 - px^2 - py^2), NaN map
// when not positive (src/tracking.jl:1183-1186, 1203-1206)
template <class S, class T>
TP_DEV bool pz_of(const S& s, double beti, T& pz) {
  const T pz2 = 1.0 + 2.0 * s.d * beti + s.d * s.d - s.px * s.px - s.py * s.py;
  if (!(val(pz2) > 0.0)) return false;
  pz = gsqrt(pz2);
  return true;
}

// Solenoid with ks != 0, src/tracking.jl:542-560
template <class S>
TP_DEV S solenoid(S s, const double* __restrict__ p) {
  using T = decltype(s.x);
  const double L = p[0], ks = p[1];
  const T p_norm = 1.0 / (1.0 + s.d);
  const T x = s.x, xpr = s.px * p_norm, y = s.y, ypr = s.py * p_norm;
  const T H = ks * p_norm / 2.0;
  const T S_ = gsin(L * H), C = gcos(L * H);
  S o;
  o.x = x * C * C + xpr * C * S_ / H + y * C * S_ + ypr * S_ * S_ / H;
  o.px = (-x * H * C * S_ + xpr * C * C - y * H * S_ * S_ + ypr * C * S_) / p_norm;
  o.y = -x * C * S_ - xpr * S_ * S_ / H + y * C * C + ypr * C * S_ / H;
  o.py = (x * H * S_ * S_ - xpr * C * S_ - y * C * S_ * H + ypr * C * C) / p_norm;
  o.z = s.z + L * (H * H * (x * x + y * y) + 2.0 * H * (xpr * y - ypr * x) + xpr * xpr + ypr * ypr) / 2.0;
  o.d = s.d;
  return o;
}

// Corrector, src/tracking.jl:586-596 (p[3] = xkick^2/3 + ykick^2/3, p[4..5] = kicks/2)
template <class S>
TP_DEV S corrector(S s, const double* __restrict__ p) {
  using T = decltype(s.x);
  const T p_norm = 1.0 / (1.0 + s.d);
  const T NormL = p[0] * p_norm;
  S o = s;
  o.z = s.z + NormL * p_norm * (p[3] + s.px * s.px + s.py * s.py + s.px * p[1] + s.py * p[2]) / 2.0;
  o.x = s.x + NormL * (s.px + p[4]);
  o.px = s.px + p[1];
  o.y = s.y + NormL * (s.py + p[5]);
  o.py = s.py + p[2];
  return o;
}

// Translation, src/tracking.jl:1182-1192
template <class S>
TP_DEV S translation(S s, const double* __restrict__ p, double beti) {
  decltype(s.x) pz;
  if (!pz_of(s, beti, pz)) return nan_state(s);
  S o = s;
  o.x = s.x - (p[0] + p[2] * s.px / pz);
  o.y = s.y - (p[1] + p[2] * s.py / pz);
  o.z = s.z + p[2] * (beti + s.d) / pz;
  return o;
}

// YRotation, src/tracking.jl:1194-1218 (p = cos, sin, tan of -angle; angle 0 is a marker)
template <class S>
TP_DEV S y_rotation(S s, const double* __restrict__ p, double beti) {
  using T = decltype(s.x);
  const double ca = p[0], sa = p[1], ta = p[2];
  T pz;
  if (!pz_of(s, beti, pz)) return nan_state(s);
  const T ptt = 1.0 - ta * s.px / pz;
  if (fabs(val(ptt)) < 1.4901161193847656e-08) return nan_state(s);  // sqrt(eps(Float64))
  S o = s;
  o.x = s.x / (ca * ptt);
  o.px = ca * s.px + sa * pz;
  o.y = s.y + ta * s.x * s.py / (pz * ptt);
  o.z = s.z + ta * s.x * (beti + s.d) / (pz * ptt);
  return o;
}

// CrabCavity, src/tracking.jl:1531-1547: the phase is taken at the entrance z, before the
// first half drift
template <bool EXH>
TP_DEV St crab(St s, uint32_t fl, const double* __restrict__ p, double beti) {
  const double ang = p[3] * s.z + p[4];
  Cx c;
  if (fl & TP_F_RF_DRIFT) {
    mkctx(s, c, beti);
    drift<EXH, false>(s, c, p[0], p[1]);
  }
  const double a = ang * beti;
  s.px = s.px + p[2] * sin(a);
  s.d = s.d - p[5] * s.x * cos(a);
  if (fl & TP_F_RF_DRIFT) {
    mkctx(s, c, beti);
    drift<EXH, false>(s, c, p[0], p[1]);
  }
  return s;
}

// StrongGaussianBeam / StrongThinGaussianBeam, src/tracking.jl:1610-1641
template <class S>
TP_DEV S gauss_bb(S s, int nslice, const double* __restrict__ p) {
  using T = decltype(s.x);
  const double sx2 = p[0], sy2 = p[1], factor = p[2];
  T px = s.px, py = s.py;
  for (int i = 0; i < nslice; ++i) {
    const double* q = p + 3 + 3 * i;
    const T dx = s.x - q[1];
    const T dy = s.y - q[2];
    const T g = q[0] * gexp(-0.5 * (dx * dx / sx2 + dy * dy / sy2));
    px = px - factor * g * dx / sx2;
    py = py - factor * g * dy / sy2;
  }
  s.px = px;
  s.py = py;
  return s;
}

// LongitudinalWake: _linear_interpolate of src/elements/wake.jl:67-84 (first interval
// with x < times[i], as the reference scans)
template <class T>
TP_DEV T lin_interp(T x, const double* __restrict__ xs, const double* __restrict__ ys, int n) {
  if (val(x) <= xs[0]) {
    const double slope = (ys[1] - ys[0]) / (xs[1] - xs[0]);
    return ys[0] + slope * (x - xs[0]);
  } else if (val(x) >= xs[n - 1]) {
    const double slope = (ys[n - 1] - ys[n - 2]) / (xs[n - 1] - xs[n - 2]);
    return ys[n - 1] + slope * (x - xs[n - 1]);
  }
  for (int i = 1; i < n; ++i) {
    if (val(x) < xs[i]) {
      const double slope = (ys[i] - ys[i - 1]) / (xs[i] - xs[i - 1]);
      return ys[i - 1] + slope * (x - xs[i - 1]);
    }
  }
  return T(ys[n - 1]);
}

// LBend body, src/tracking.jl:1076-1150 (p = L, K, angle, by_error, angle/L)
template <class S>
TP_DEV S lbend_body(S s, const double* __restrict__ p) {
  using T = decltype(s.x);
  const double L = p[0], grd = p[1], by_error = p[3], Kx = p[4];
  const T p_norm = 1.0 / (1.0 + s.d);
  const T G1 = (Kx * Kx + grd) * p_norm;
  const T G2 = -grd * p_norm;
  const double tol = 1.4901161193847656e-08;  // sqrt(eps(Float64))
  T MHD = 1.0, M12 = L, M21 = 0.0, MVD = 1.0, M34 = L, M43 = 0.0;
  T arg1 = 0.0, sqrtG1 = 0.0;
  if (fabs(val(G1)) >= tol) {
    if (val(G1) > 0.0) {
      sqrtG1 = gsqrt(G1);
      arg1 = L * sqrtG1;
      MHD = gcos(arg1);
      M12 = gsin(arg1) / sqrtG1;
      M21 = -gsin(arg1) * sqrtG1;
    } else {
      sqrtG1 = gsqrt(-G1);
      arg1 = L * sqrtG1;
      MHD = gcosh(arg1);
      M12 = gsinh(arg1) / sqrtG1;
      M21 = gsinh(arg1) * sqrtG1;
    }
  }
  if (fabs(val(G2)) >= tol) {
    if (val(G2) > 0.0) {
      const T sqrtG2 = gsqrt(G2);
      const T arg2 = L * sqrtG2;
      MVD = gcos(arg2);
      M34 = gsin(arg2) / sqrtG2;
      M43 = -gsin(arg2) * sqrtG2;
    } else {
      const T sqrtG2 = gsqrt(-G2);
      const T arg2 = L * sqrtG2;
      MVD = gcosh(arg2);
      M34 = gsinh(arg2) / sqrtG2;
      M43 = gsinh(arg2) * sqrtG2;
    }
  }
  const T x = s.x, xpr = s.px * p_norm, y = s.y, ypr = s.py * p_norm;
  const T dterm = s.d * p_norm - by_error;
  T x_new = MHD * x + M12 * xpr;
  T px_new = (M21 * x + MHD * xpr) / p_norm;
  if (fabs(val(G1)) < tol) {
    x_new = x_new + dterm * L * L * Kx / 2.0;
    px_new = px_new + dterm * L * Kx / p_norm;
  } else if (val(G1) > 0.0) {
    x_new = x_new + dterm * (1.0 - gcos(arg1)) * Kx / G1;
    px_new = px_new + dterm * gsin(arg1) * Kx / (sqrtG1 * p_norm);
  } else {
    x_new = x_new + dterm * (1.0 - gcosh(arg1)) * Kx / G1;
    px_new = px_new + dterm * gsinh(arg1) * Kx / (sqrtG1 * p_norm);
  }
  const T y_new = MVD * y + M34 * ypr;
  const T py_new = (M43 * y + MVD * ypr) / p_norm;
  T z_new = s.z + xpr * xpr * (L + MHD * M12) / 4.0;
  if (fabs(val(G1)) >= tol) {
    z_new = z_new + (L - MHD * M12) * (x * x * G1 + dterm * dterm * Kx * Kx / G1 - 2.0 * x * Kx * dterm) / 4.0;
    z_new = z_new + M12 * M21 * (x * xpr - xpr * dterm * Kx / G1) / 2.0;
    z_new = z_new + (Kx * x * M12 + xpr * (1.0 - MHD) * Kx / G1 + dterm * (L - M12) * Kx * Kx / G1);
  }
  z_new = z_new + ((L - MVD * M34) * y * y * G2 + ypr * ypr * (L + MVD * M34)) / 4.0;
  z_new = z_new + M34 * M43 * y * ypr / 2.0;
  return S{x_new, px_new, y_new, py_new, z_new, s.d};
}

// The z- and delta-kicks of the families without drifts: AccelCavity (src/tracking.jl:
// 1549-1559), LongitudinalRFMap (:1575-1585), LongitudinalRLCWake (:1647-1654,
// src/elements/wake.jl:25-36), LongitudinalWake (:1656-1663, src/elements/wake.jl:86-91),
// and the Lorentz boosts (:1587-1604)
template <class S>
TP_DEV S accel_kick(S s, const double* __restrict__ p) {
  s.d = s.d + p[0] * (gsin(p[1] * s.z + p[2]) - p[3]);
  return s;
}
template <class S>
TP_DEV S longrf_slip(S s, const double* __restrict__ p) {
  s.z = s.z - p[0] * s.d;
  return s;
}
template <class S>
TP_DEV S rlc_wake(S s, const double* __restrict__ p) {
  using T = decltype(s.x);
  const T zc = s.z / 2.99792458e8;
  const T t = val(zc) < 0.0 ? zc : T(0.0);  // min(z / c, 0)
  const T a = p[2] * t;
  const T w = p[1] * (gcos(a) + gsin(a) / p[3]) * gexp(p[4] * t / p[5]);
  s.d = s.d - p[0] * w;
  return s;
}
template <class S>
TP_DEV S tab_wake(S s, int n, const double* __restrict__ p) {
  using T = decltype(s.x);
  const T t = s.z / 2.99792458e8;
  const T w = (val(t) > p[2]) ? T(0.0) : lin_interp(t * p[1], p + 3, p + 3 + n, n);
  s.d = s.d - p[0] * w;
  return s;
}
template <class S>
TP_DEV S boost(S s, const double* __restrict__ p, bool inverse) {
  if (inverse) {
    const auto x_new = s.x - p[0] * s.z;
    const auto d_new = s.d + p[0] * s.px;
    return S{x_new, s.px * p[1], s.y, s.py * p[1], s.z * p[1], d_new};
  }
  const auto x_new = s.x + p[0] * s.z;
  const auto d_new = s.d - p[0] * s.px;
  return S{x_new, s.px * p[1], s.y, s.py * p[1], s.z * p[1], d_new};
}

// ---------------------------------------------------------------------------------------
// Wiggler, src/tracking.jl:1224-1525.  Harmonic h (11 doubles): kx ky kz tz, then the
// leading factors of the ax, axpy, ay, aypx terms, of Bx and By, and the sinc flag.
// ---------------------------------------------------------------------------------------
template <class T>
TP_DEV T sinc_taylor(T x) {  // _sinc_taylor, :1224-1227
  const T x2 = x * x;
  return 1.0 - x2 / 6.0 * (1.0 - x2 / 20.0 * (1.0 - x2 / 42.0 * (1.0 - x2 / 72.0)));
}

// _wig_ax_axpy (:1229-1279) and _wig_ay_aypx (:1281-1331) at (x, y, Zw)
template <class T>
TP_DEV void wig_ax(const double* __restrict__ h, int nh, int nv, T x, T y, double Zw, T& ax, T& axpy) {
  ax = 0.0;
  axpy = 0.0;
  for (int i = 0; i < nh; ++i, h += 11) {
    const double kx = h[0], ky = h[1], kz = h[2], tz = h[3];
    const T cx = gcos(kx * x), chy = gcosh(ky * y);
    const double sz = sin(kz * Zw + tz);
    ax = ax + h[4] * cx * chy * sz;
    const T shy = gsinh(ky * y);
    const T sxkx = h[10] != 0.0 ? gsin(kx * x) / kx : x * sinc_taylor(kx * x);
    axpy = axpy + h[5] * sxkx * shy * sz;
  }
  for (int i = 0; i < nv; ++i, h += 11) {
    const double kx = h[0], ky = h[1], kz = h[2], tz = h[3];
    const T shx = gsinh(kx * x), sy = gsin(ky * y);
    const double sz = sin(kz * Zw + tz);
    ax = ax + h[4] * shx * sy * sz;
    const T chx = gcosh(kx * x), cy = gcos(ky * y);
    axpy = axpy + h[5] * chx * cy * sz;
  }
}

template <class T>
TP_DEV void wig_ay(const double* __restrict__ h, int nh, int nv, T x, T y, double Zw, T& ay, T& aypx) {
  ay = 0.0;
  aypx = 0.0;
  for (int i = 0; i < nh; ++i, h += 11) {
    const double kx = h[0], ky = h[1], kz = h[2], tz = h[3];
    const T sx = gsin(kx * x), shy = gsinh(ky * y);
    const double sz = sin(kz * Zw + tz);
    ay = ay + h[6] * sx * shy * sz;
    const T cx = gcos(kx * x), chy = gcosh(ky * y);
    aypx = aypx + h[7] * cx * chy * sz;
  }
  for (int i = 0; i < nv; ++i, h += 11) {
    const double kx = h[0], ky = h[1], kz = h[2], tz = h[3];
    const T chx = gcosh(kx * x), cy = gcos(ky * y);
    const double sz = sin(kz * Zw + tz);
    ay = ay + h[6] * chx * cy * sz;
    const T shx = gsinh(kx * x);
    const T syky = h[10] != 0.0 ? gsin(ky * y) / ky : y * sinc_taylor(ky * y);
    aypx = aypx + h[7] * shx * syky * sz;
  }
}

// _wig_map_2nd, :1333-1379 (Zw is advanced by the caller's copy)
template <class S>
TP_DEV void wig_map_2nd(S& s, const double* __restrict__ h, int nh, int nv, double dl, double& Zw) {
  using T = decltype(s.x);
  const T delta = s.d;
  const T inv1pd = 1.0 / (1.0 + delta);
  const T dld = dl * inv1pd;
  const double dl2 = dl / 2.0;
  const T dl2d = dl2 * inv1pd;
  Zw = Zw + dl2;
  T x = s.x, px = s.px, y = s.y, py = s.py, z = s.z;
  T a, b;
  wig_ay(h, nh, nv, x, y, Zw, a, b);
  px = px - b;
  py = py - a;
  y = y + dl2d * py;
  z = z + (dl2d / 2.0) * (py * py) * inv1pd;
  wig_ay(h, nh, nv, x, y, Zw, a, b);
  px = px + b;
  py = py + a;
  wig_ax(h, nh, nv, x, y, Zw, a, b);
  px = px - a;
  py = py - b;
  x = x + dld * px;
  z = z + (dld / 2.0) * (px * px) * inv1pd;
  wig_ax(h, nh, nv, x, y, Zw, a, b);
  px = px + a;
  py = py + b;
  wig_ay(h, nh, nv, x, y, Zw, a, b);
  px = px - b;
  py = py - a;
  y = y + dl2d * py;
  z = z + (dl2d / 2.0) * (py * py) * inv1pd;
  wig_ay(h, nh, nv, x, y, Zw, a, b);
  px = px + b;
  py = py + a;
  Zw = Zw + dl2;
  s = S{x, px, y, py, z, delta};
}

// The field's bending term B^2 of _wig_B (Bx, By; :1401-1457) at (x, y, Zw)
template <class T>
TP_DEV T wig_b2(const double* __restrict__ h, int nh, int nv, T x, T y, double Zw) {
  T Bx = 0.0, By = 0.0;
  const double* q = h;
  for (int i = 0; i < nh; ++i, q += 11) {
    const double kx = q[0], ky = q[1], kz = q[2], tz = q[3];
    const T sx = gsin(kx * x), cx = gcos(kx * x), chy = gcosh(ky * y), shy = gsinh(ky * y);
    const double cz = cos(kz * Zw + tz);
    Bx = Bx + q[8] * sx * shy * cz;
    By = By - q[9] * cx * chy * cz;
  }
  for (int i = 0; i < nv; ++i, q += 11) {
    const double kx = q[0], ky = q[1], kz = q[2], tz = q[3];
    const T shx = gsinh(kx * x), chx = gcosh(kx * x), sy = gsin(ky * y), cy = gcos(ky * y);
    const double cz = cos(kz * Zw + tz);
    Bx = Bx + q[8] * chx * cy * cz;
    By = By - q[9] * shx * sy * cz;
  }
  return Bx * Bx + By * By;
}

// _wig_radiation_kicks (:1459-1470), with the ax / ay kicks around them as
// _wig_pass_4th_rad applies them (:1472-1507)
template <class S>
TP_DEV void wig_rad_kick(S& s, const double* __restrict__ p, const double* __restrict__ h, int nh, int nv, double Zw) {
  using T = decltype(s.x);
  T ax, axpy, ay, aypx;
  wig_ax(h, nh, nv, s.x, s.y, Zw, ax, axpy);
  wig_ay(h, nh, nv, s.x, s.y, Zw, ay, aypx);
  s.px = s.px - ax;
  s.py = s.py - ay;
  const T B2 = wig_b2(h, nh, nv, s.x, s.y, Zw);
  if (val(B2) != 0.0) {
    const double H = p[6];
    const T irho2 = B2 / (H * H);
    const T dFactor = (1.0 + s.d) * (1.0 + s.d);
    const T dDelta = -p[7] * dFactor * irho2 * p[8];
    const T scale = 1.0 + dDelta;
    s.px = s.px * scale;
    s.py = s.py * scale;
    s.d = s.d + dDelta;
  }
  s.px = s.px + ax;
  s.py = s.py + ay;
}

TP_DEV St wiggler(St s, int nstep, const double* __restrict__ p) {
  const double dl1 = p[1], dl0 = p[2];
  const int nh = (int)p[3], nv = (int)p[4];
  const bool rad = p[5] != 0.0;
  const double* h = p + 9;
  double Zw = 0.0;
  if (rad) wig_rad_kick(s, p, h, nh, nv, Zw);
  for (int k = 0; k < nstep; ++k) {
    wig_map_2nd(s, h, nh, nv, dl1, Zw);
    wig_map_2nd(s, h, nh, nv, dl0, Zw);
    wig_map_2nd(s, h, nh, nv, dl1, Zw);
    if (rad) wig_rad_kick(s, p, h, nh, nv, Zw);
  }
  return s;
}

// One element of a family above; the caller rebuilds the delta-dependent hoists after it.
template <bool EXH>
__device__ __noinline__ St extra_map(St s, int tag, uint32_t fl, int n, const double* __restrict__ p,
                                     double beti) {
  switch (tag) {
    case TP_SOLENOID:
      return solenoid(s, p);
    case TP_CORRECTOR:
      return corrector(s, p);
    case TP_TRANSLATION:
      return translation(s, p, beti);
    case TP_YROTATION:
      return y_rotation(s, p, beti);
    case TP_CRAB:
      return crab<EXH>(s, fl, p, beti);
    case TP_ACCEL:
      return accel_kick(s, p);
    case TP_LONGRF:
      return longrf_slip(s, p);
    case TP_BOOST:
      return boost(s, p, false);
    case TP_INVBOOST:
      return boost(s, p, true);
    case TP_GAUSSBB:
      return gauss_bb(s, n, p);
    case TP_RLCWAKE:
      return rlc_wake(s, p);
    case TP_TABWAKE:
      return tab_wake(s, n, p);
    case TP_LBEND: {  // src/tracking.jl:1152-1170: entrance edge, body, exit edge (method 1)
      const double irho = p[4];
      edge_lin(s, 1, p + 5, irho, false);
      s = lbend_body(s, p);
      edge_lin(s, 1, p + 9, irho, true);
      return s;
    }
    case TP_WIGGLER:
      return wiggler(s, n, p);
    default:  // TP_NAN
      return nan_state();
  }
}

}  // namespace tpk

#endif  // TP_EXTRA_H
