// tp_dual.h -- forward-mode dual numbers for the reverse pass.
//
// D1 carries a value and one directional derivative through + - * / and the elementary
// functions the element maps use.  Maps written once over a scalar type T (double or D1)
// give the reference's arithmetic for T = double and, for T = D1 with a unit seed on one
// input, one column of the map's Jacobian; the reverse pass forms the vector-Jacobian
// products of its rarer maps (Forest's bend maps, the other element families) that way, at
// the input it has reconstructed.  The g* functions are the elementary functions for both
// types (for double, the same calls the forward maps make).
#ifndef TP_DUAL_H
#define TP_DUAL_H

#include "tp_maps.h"

namespace tpk {

struct D1 {
  double v, d;
  D1() = default;
  TP_DEV D1(double x) : v(x), d(0.0) {}
  TP_DEV D1(double x, double dx) : v(x), d(dx) {}
};
TP_DEV D1 operator+(D1 a, D1 b) { return D1(a.v + b.v, a.d + b.d); }
TP_DEV D1 operator-(D1 a, D1 b) { return D1(a.v - b.v, a.d - b.d); }
TP_DEV D1 operator-(D1 a) { return D1(-a.v, -a.d); }
TP_DEV D1 operator*(D1 a, D1 b) { return D1(a.v * b.v, a.d * b.v + a.v * b.d); }
TP_DEV D1 operator/(D1 a, D1 b) {
  const double q = a.v / b.v;
  return D1(q, (a.d - q * b.d) / b.v);
}
TP_DEV D1 operator+(double a, D1 b) { return D1(a + b.v, b.d); }
TP_DEV D1 operator+(D1 a, double b) { return D1(a.v + b, a.d); }
TP_DEV D1 operator-(double a, D1 b) { return D1(a - b.v, -b.d); }
TP_DEV D1 operator-(D1 a, double b) { return D1(a.v - b, a.d); }
TP_DEV D1 operator*(double a, D1 b) { return D1(a * b.v, a * b.d); }
TP_DEV D1 operator*(D1 a, double b) { return D1(a.v * b, a.d * b); }
TP_DEV D1 operator/(D1 a, double b) { return D1(a.v / b, a.d / b); }
TP_DEV D1 operator/(double a, D1 b) {
  const double q = a / b.v;
  return D1(q, -q * b.d / b.v);
}
TP_DEV double val(double a) { return a; }
TP_DEV double val(D1 a) { return a.v; }
TP_DEV double gsqrt(double a) { return sqrt(a); }
TP_DEV D1 gsqrt(D1 a) {
  const double r = sqrt(a.v);
  return D1(r, 0.5 * a.d / r);
}
TP_DEV double gtan(double a) { return tan_lib(a); }
TP_DEV D1 gtan(D1 a) {
  const double t = tan_lib(a.v);
  return D1(t, a.d * (1.0 + t * t));
}
TP_DEV double gatan(double a) { return atan(a); }
TP_DEV D1 gatan(D1 a) { return D1(atan(a.v), a.d / (1.0 + a.v * a.v)); }
TP_DEV double gcos(double a) { return cos(a); }
TP_DEV D1 gcos(D1 a) { return D1(cos(a.v), -sin(a.v) * a.d); }
TP_DEV double gsin(double a) { return sin(a); }
TP_DEV D1 gsin(D1 a) { return D1(sin(a.v), cos(a.v) * a.d); }
TP_DEV double gexp(double a) { return exp(a); }
TP_DEV D1 gexp(D1 a) {
  const double e = exp(a.v);
  return D1(e, e * a.d);
}
TP_DEV double gsinh(double a) { return sinh(a); }
TP_DEV D1 gsinh(D1 a) { return D1(sinh(a.v), cosh(a.v) * a.d); }
TP_DEV double gcosh(double a) { return cosh(a); }
TP_DEV D1 gcosh(D1 a) { return D1(cosh(a.v), sinh(a.v) * a.d); }
// asin of the argument clamped to [-1, 1] (src/tracking.jl:839-840); zero slope where clamped
TP_DEV double gasin_c(double a) { return asin_s(clamp1(a)); }
TP_DEV D1 gasin_c(D1 a) {
  if (!(fabs(a.v) < 1.0)) return D1(asin_s(clamp1(a.v)), 0.0);
  return D1(asin_s(a.v), a.d / sqrt(1.0 - a.v * a.v));
}
// The particle state over D1 (St over double)
struct SD {
  D1 x, px, y, py, z, d;
};

}  // namespace tpk

#endif  // TP_DUAL_H
