// tp_extra.h -- device maps of the element families outside the hot lattices: solenoid,
// corrector, translation, y-rotation, crab / accelerating cavities, the longitudinal RF
// map, Lorentz boosts, strong Gaussian beam-beam kicks, longitudinal wakes, the linear bend
// and the NaN map (parameter blocks in include/trackpad_table.h).
//
// Every map restates the reference formula in its operation order (src/tracking.jl:529-605,
// 1076-1218, 1531-1663; Julia evaluates a*b*c as (a*b)*c and x^2 as x*x) so that the parity
// build reproduces it.  The maps are compiled only into the kernel variant that holds every
// feature, and they run in one out-of-line function, so their registers and transcendentals
// do not weigh on the hot element loop.
#ifndef TP_EXTRA_H
#define TP_EXTRA_H

#include "tp_maps.h"

namespace tpk {

TP_DEV St nan_state() {
  const double n = nan_d();
  return St{n, n, n, n, n, n};
}

// pz of Translation / YRotation: sqrt(1 + 2 delta beti + delta^2 - px^2 - py^2), NaN map
// when not positive (src/tracking.jl:1183-1186, 1203-1206)
TP_DEV bool pz_of(const St& s, double beti, double& pz) {
  const double pz2 = 1.0 + 2.0 * s.d * beti + s.d * s.d - s.px * s.px - s.py * s.py;
  if (!(pz2 > 0.0)) return false;
  pz = sqrt(pz2);
  return true;
}

// Solenoid with ks != 0, src/tracking.jl:542-560
TP_DEV St solenoid(St s, const double* __restrict__ p) {
  const double L = p[0], ks = p[1];
  const double p_norm = 1.0 / (1.0 + s.d);
  const double x = s.x, xpr = s.px * p_norm, y = s.y, ypr = s.py * p_norm;
  const double H = ks * p_norm / 2.0;
  const double S = sin(L * H), C = cos(L * H);
  St o;
  o.x = x * C * C + xpr * C * S / H + y * C * S + ypr * S * S / H;
  o.px = (-x * H * C * S + xpr * C * C - y * H * S * S + ypr * C * S) / p_norm;
  o.y = -x * C * S - xpr * S * S / H + y * C * C + ypr * C * S / H;
  o.py = (x * H * S * S - xpr * C * S - y * C * S * H + ypr * C * C) / p_norm;
  o.z = s.z + L * (H * H * (x * x + y * y) + 2.0 * H * (xpr * y - ypr * x) + xpr * xpr + ypr * ypr) / 2.0;
  o.d = s.d;
  return o;
}

// Corrector, src/tracking.jl:586-596 (p[3] = xkick^2/3 + ykick^2/3, p[4..5] = kicks/2)
TP_DEV St corrector(St s, const double* __restrict__ p) {
  const double p_norm = 1.0 / (1.0 + s.d);
  const double NormL = p[0] * p_norm;
  St o = s;
  o.z = s.z + NormL * p_norm * (p[3] + s.px * s.px + s.py * s.py + s.px * p[1] + s.py * p[2]) / 2.0;
  o.x = s.x + NormL * (s.px + p[4]);
  o.px = s.px + p[1];
  o.y = s.y + NormL * (s.py + p[5]);
  o.py = s.py + p[2];
  return o;
}

// Translation, src/tracking.jl:1182-1192
TP_DEV St translation(St s, const double* __restrict__ p, double beti) {
  double pz;
  if (!pz_of(s, beti, pz)) return nan_state();
  St o = s;
  o.x = s.x - (p[0] + p[2] * s.px / pz);
  o.y = s.y - (p[1] + p[2] * s.py / pz);
  o.z = s.z + p[2] * (beti + s.d) / pz;
  return o;
}

// YRotation, src/tracking.jl:1194-1218 (p = cos, sin, tan of -angle; angle 0 is a marker)
TP_DEV St y_rotation(St s, const double* __restrict__ p, double beti) {
  const double ca = p[0], sa = p[1], ta = p[2];
  double pz;
  if (!pz_of(s, beti, pz)) return nan_state();
  const double ptt = 1.0 - ta * s.px / pz;
  if (fabs(ptt) < 1.4901161193847656e-08) return nan_state();  // sqrt(eps(Float64))
  St o = s;
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
    drift<EXH>(s, c, p[0], p[1]);
  }
  const double a = ang * beti;
  s.px = s.px + p[2] * sin(a);
  s.d = s.d - p[5] * s.x * cos(a);
  if (fl & TP_F_RF_DRIFT) {
    mkctx(s, c, beti);
    drift<EXH>(s, c, p[0], p[1]);
  }
  return s;
}

// StrongGaussianBeam / StrongThinGaussianBeam, src/tracking.jl:1610-1641
TP_DEV St gauss_bb(St s, int nslice, const double* __restrict__ p) {
  const double sx2 = p[0], sy2 = p[1], factor = p[2];
  double px = s.px, py = s.py;
  for (int i = 0; i < nslice; ++i) {
    const double* q = p + 3 + 3 * i;
    const double dx = s.x - q[1];
    const double dy = s.y - q[2];
    const double g = q[0] * exp(-0.5 * (dx * dx / sx2 + dy * dy / sy2));
    px = px - factor * g * dx / sx2;
    py = py - factor * g * dy / sy2;
  }
  s.px = px;
  s.py = py;
  return s;
}

// LongitudinalWake: _linear_interpolate of src/elements/wake.jl:67-84 (first interval
// with x < times[i], as the reference scans)
TP_DEV double lin_interp(double x, const double* __restrict__ xs, const double* __restrict__ ys, int n) {
  if (x <= xs[0]) {
    const double slope = (ys[1] - ys[0]) / (xs[1] - xs[0]);
    return ys[0] + slope * (x - xs[0]);
  } else if (x >= xs[n - 1]) {
    const double slope = (ys[n - 1] - ys[n - 2]) / (xs[n - 1] - xs[n - 2]);
    return ys[n - 1] + slope * (x - xs[n - 1]);
  }
  for (int i = 1; i < n; ++i) {
    if (x < xs[i]) {
      const double slope = (ys[i] - ys[i - 1]) / (xs[i] - xs[i - 1]);
      return ys[i - 1] + slope * (x - xs[i - 1]);
    }
  }
  return ys[n - 1];
}

// LBend body, src/tracking.jl:1076-1150 (p = L, K, angle, by_error, angle/L)
TP_DEV St lbend_body(St s, const double* __restrict__ p) {
  const double L = p[0], grd = p[1], by_error = p[3], Kx = p[4];
  const double p_norm = 1.0 / (1.0 + s.d);
  const double G1 = (Kx * Kx + grd) * p_norm;
  const double G2 = -grd * p_norm;
  const double tol = 1.4901161193847656e-08;  // sqrt(eps(Float64))
  double MHD = 1.0, M12 = L, M21 = 0.0, MVD = 1.0, M34 = L, M43 = 0.0;
  double arg1 = 0.0, sqrtG1 = 0.0;
  if (fabs(G1) >= tol) {
    if (G1 > 0.0) {
      sqrtG1 = sqrt(G1);
      arg1 = L * sqrtG1;
      MHD = cos(arg1);
      M12 = sin(arg1) / sqrtG1;
      M21 = -sin(arg1) * sqrtG1;
    } else {
      sqrtG1 = sqrt(-G1);
      arg1 = L * sqrtG1;
      MHD = cosh(arg1);
      M12 = sinh(arg1) / sqrtG1;
      M21 = sinh(arg1) * sqrtG1;
    }
  }
  if (fabs(G2) >= tol) {
    if (G2 > 0.0) {
      const double sqrtG2 = sqrt(G2);
      const double arg2 = L * sqrtG2;
      MVD = cos(arg2);
      M34 = sin(arg2) / sqrtG2;
      M43 = -sin(arg2) * sqrtG2;
    } else {
      const double sqrtG2 = sqrt(-G2);
      const double arg2 = L * sqrtG2;
      MVD = cosh(arg2);
      M34 = sinh(arg2) / sqrtG2;
      M43 = sinh(arg2) * sqrtG2;
    }
  }
  const double x = s.x, xpr = s.px * p_norm, y = s.y, ypr = s.py * p_norm;
  const double dterm = s.d * p_norm - by_error;
  double x_new = MHD * x + M12 * xpr;
  double px_new = (M21 * x + MHD * xpr) / p_norm;
  if (fabs(G1) < tol) {
    x_new = x_new + dterm * L * L * Kx / 2.0;
    px_new = px_new + dterm * L * Kx / p_norm;
  } else if (G1 > 0.0) {
    x_new = x_new + dterm * (1.0 - cos(arg1)) * Kx / G1;
    px_new = px_new + dterm * sin(arg1) * Kx / (sqrtG1 * p_norm);
  } else {
    x_new = x_new + dterm * (1.0 - cosh(arg1)) * Kx / G1;
    px_new = px_new + dterm * sinh(arg1) * Kx / (sqrtG1 * p_norm);
  }
  const double y_new = MVD * y + M34 * ypr;
  const double py_new = (M43 * y + MVD * ypr) / p_norm;
  double z_new = s.z + xpr * xpr * (L + MHD * M12) / 4.0;
  if (fabs(G1) >= tol) {
    z_new = z_new + (L - MHD * M12) * (x * x * G1 + dterm * dterm * Kx * Kx / G1 - 2.0 * x * Kx * dterm) / 4.0;
    z_new = z_new + M12 * M21 * (x * xpr - xpr * dterm * Kx / G1) / 2.0;
    z_new = z_new + (Kx * x * M12 + xpr * (1.0 - MHD) * Kx / G1 + dterm * (L - M12) * Kx * Kx / G1);
  }
  z_new = z_new + ((L - MVD * M34) * y * y * G2 + ypr * ypr * (L + MVD * M34)) / 4.0;
  z_new = z_new + M34 * M43 * y * ypr / 2.0;
  return St{x_new, px_new, y_new, py_new, z_new, s.d};
}

// ---------------------------------------------------------------------------------------
// Wiggler, src/tracking.jl:1224-1525.  Harmonic h (11 doubles): kx ky kz tz, then the
// leading factors of the ax, axpy, ay, aypx terms, of Bx and By, and the sinc flag.
// ---------------------------------------------------------------------------------------
TP_DEV double sinc_taylor(double x) {  // _sinc_taylor, :1224-1227
  const double x2 = x * x;
  return 1.0 - x2 / 6.0 * (1.0 - x2 / 20.0 * (1.0 - x2 / 42.0 * (1.0 - x2 / 72.0)));
}

// _wig_ax_axpy (:1229-1279) and _wig_ay_aypx (:1281-1331) at (x, y, Zw)
TP_DEV void wig_ax(const double* __restrict__ h, int nh, int nv, double x, double y, double Zw, double& ax,
                   double& axpy) {
  ax = 0.0;
  axpy = 0.0;
  for (int i = 0; i < nh; ++i, h += 11) {
    const double kx = h[0], ky = h[1], kz = h[2], tz = h[3];
    const double cx = cos(kx * x), chy = cosh(ky * y), sz = sin(kz * Zw + tz);
    ax = ax + h[4] * cx * chy * sz;
    const double shy = sinh(ky * y);
    const double sxkx = h[10] != 0.0 ? sin(kx * x) / kx : x * sinc_taylor(kx * x);
    axpy = axpy + h[5] * sxkx * shy * sz;
  }
  for (int i = 0; i < nv; ++i, h += 11) {
    const double kx = h[0], ky = h[1], kz = h[2], tz = h[3];
    const double shx = sinh(kx * x), sy = sin(ky * y), sz = sin(kz * Zw + tz);
    ax = ax + h[4] * shx * sy * sz;
    const double chx = cosh(kx * x), cy = cos(ky * y);
    axpy = axpy + h[5] * chx * cy * sz;
  }
}

TP_DEV void wig_ay(const double* __restrict__ h, int nh, int nv, double x, double y, double Zw, double& ay,
                   double& aypx) {
  ay = 0.0;
  aypx = 0.0;
  for (int i = 0; i < nh; ++i, h += 11) {
    const double kx = h[0], ky = h[1], kz = h[2], tz = h[3];
    const double sx = sin(kx * x), shy = sinh(ky * y), sz = sin(kz * Zw + tz);
    ay = ay + h[6] * sx * shy * sz;
    const double cx = cos(kx * x), chy = cosh(ky * y);
    aypx = aypx + h[7] * cx * chy * sz;
  }
  for (int i = 0; i < nv; ++i, h += 11) {
    const double kx = h[0], ky = h[1], kz = h[2], tz = h[3];
    const double chx = cosh(kx * x), cy = cos(ky * y), sz = sin(kz * Zw + tz);
    ay = ay + h[6] * chx * cy * sz;
    const double shx = sinh(kx * x);
    const double syky = h[10] != 0.0 ? sin(ky * y) / ky : y * sinc_taylor(ky * y);
    aypx = aypx + h[7] * shx * syky * sz;
  }
}

// _wig_map_2nd, :1333-1379 (Zw is advanced by the caller's copy)
TP_DEV void wig_map_2nd(St& s, const double* __restrict__ h, int nh, int nv, double dl, double& Zw) {
  const double delta = s.d;
  const double inv1pd = 1.0 / (1.0 + delta);
  const double dld = dl * inv1pd;
  const double dl2 = dl / 2.0;
  const double dl2d = dl2 * inv1pd;
  Zw = Zw + dl2;
  double x = s.x, px = s.px, y = s.y, py = s.py, z = s.z;
  double a, b;
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
  s = St{x, px, y, py, z, delta};
}

// _wig_B (Bx, By; :1401-1457) and _wig_radiation_kicks (:1459-1470), with the ax / ay
// kicks around them as _wig_pass_4th_rad applies them (:1472-1507)
TP_DEV void wig_rad_kick(St& s, const double* __restrict__ p, const double* __restrict__ h, int nh, int nv,
                         double Zw) {
  double ax, axpy, ay, aypx;
  wig_ax(h, nh, nv, s.x, s.y, Zw, ax, axpy);
  wig_ay(h, nh, nv, s.x, s.y, Zw, ay, aypx);
  s.px = s.px - ax;
  s.py = s.py - ay;
  double Bx = 0.0, By = 0.0;
  const double* q = h;
  for (int i = 0; i < nh; ++i, q += 11) {
    const double kx = q[0], ky = q[1], kz = q[2], tz = q[3];
    const double sx = sin(kx * s.x), cx = cos(kx * s.x), chy = cosh(ky * s.y), shy = sinh(ky * s.y);
    const double cz = cos(kz * Zw + tz);
    Bx = Bx + q[8] * sx * shy * cz;
    By = By - q[9] * cx * chy * cz;
  }
  for (int i = 0; i < nv; ++i, q += 11) {
    const double kx = q[0], ky = q[1], kz = q[2], tz = q[3];
    const double shx = sinh(kx * s.x), chx = cosh(kx * s.x), sy = sin(ky * s.y), cy = cos(ky * s.y);
    const double cz = cos(kz * Zw + tz);
    Bx = Bx + q[8] * chx * cy * cz;
    By = By - q[9] * shx * sy * cz;
  }
  const double B2 = Bx * Bx + By * By;
  if (B2 != 0.0) {
    const double H = p[6];
    const double irho2 = B2 / (H * H);
    const double dFactor = (1.0 + s.d) * (1.0 + s.d);
    const double dDelta = -p[7] * dFactor * irho2 * p[8];
    const double scale = 1.0 + dDelta;
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
    case TP_ACCEL:  // src/tracking.jl:1549-1559
      s.d = s.d + p[0] * (sin(p[1] * s.z + p[2]) - p[3]);
      return s;
    case TP_LONGRF:  // src/tracking.jl:1575-1585
      s.z = s.z - p[0] * s.d;
      return s;
    case TP_BOOST: {  // src/tracking.jl:1587-1595
      const double x_new = s.x + p[0] * s.z;
      const double d_new = s.d - p[0] * s.px;
      return St{x_new, s.px * p[1], s.y, s.py * p[1], s.z * p[1], d_new};
    }
    case TP_INVBOOST: {  // src/tracking.jl:1597-1604
      const double x_new = s.x - p[0] * s.z;
      const double d_new = s.d + p[0] * s.px;
      return St{x_new, s.px * p[1], s.y, s.py * p[1], s.z * p[1], d_new};
    }
    case TP_GAUSSBB:
      return gauss_bb(s, n, p);
    case TP_RLCWAKE: {  // src/tracking.jl:1647-1654, src/elements/wake.jl:25-36
      const double zc = s.z / 2.99792458e8;
      const double t = zc < 0.0 ? zc : 0.0;  // min(z / c, 0)
      const double a = p[2] * t;
      const double w = p[1] * (cos(a) + sin(a) / p[3]) * exp(p[4] * t / p[5]);
      s.d = s.d - p[0] * w;
      return s;
    }
    case TP_TABWAKE: {  // src/tracking.jl:1656-1663, src/elements/wake.jl:86-91
      const double t = s.z / 2.99792458e8;
      const double w = (t > p[2]) ? 0.0 : lin_interp(t * p[1], p + 3, p + 3 + n, n);
      s.d = s.d - p[0] * w;
      return s;
    }
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
