// trackpad_oracle.cpp -- CPU restatement of TrackPad.jl's multi-particle tracking path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the correctness oracle for the HIP library and
// the timed CPU baseline; it is never linked into, loaded by or called from the product.
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it.
//
// Every function restates the reference Julia code operation by operation, in the same
// evaluation order (Julia never contracts a*b+c and x^2 is x*x), so that it is compiled
// with -ffp-contract=off.  Citations are file:line in MSU-Beam-Dynamics/TrackPad.jl.
//
// The scalar type is a template parameter so that the same restatement also runs on
// forward-mode dual numbers (exact Jacobians and parameter derivatives, used to check
// the adjoint and the symplecticity/tune invariants).
//
// Pinning (what checks this restatement; see tests/test_oracle_*.py):
//   drift6, strthinkick, bndthinkick, Yoshida composition, exact bend body, linear edge,
//   RF kick, Beam: 50-digit mpmath restatements built from independent formulations
//   (straight-line drift geometry, complex power sums, circle-arc geometry) in
//   tests/golden/, 4th-order convergence of the Yoshida composition, det(J)=1 and
//   reversibility invariants.
//   bend_fringe (Forest 13.13), yrot and bend_edge: the 50-digit restatement of the same
//   formulas, J^T S J = S of the whole exact bend (dual numbers), and the fringe's x/y/z
//   updates as the derivatives of its py kick at 50 digits; the JuTrack values the
//   reference tests compare against are not available (sign conventions relative to
//   JuTrack: parity unpinned).
//   r1/r2 rotation summation order (StaticArrays SMatrix*SVector): parity unpinned.
//   Julia Base.Math sin/cos/tan/asin/atan last-ulp values: parity unpinned (glibc here).
//
// Deviations from the reference, all outside its tested domain:
//   * sqrt of a negative argument in bend_fringe raises DomainError in Julia; here it
//     yields NaN and the particle is flagged lost at element exit.
//   * order-10 multipoles: the reference polynomials are SVector{4} (max_order <= 3,
//     src/elements/quad.jl:10-11); here arrays hold ORC_MAXORD+1 = 11 coefficients and
//     the same Horner recurrence (src/tracking.jl:290-298) runs over max_order.
//   * an optional circular aperture x^2+y^2 > R^2, checked together with check_lost at
//     element exit (the reference stores r_apertures but never reads them).

#include <quadmath.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_MAXORD 10
#define ORC_NP (ORC_MAXORD + 1)

// ---------------------------------------------------------------------------------------
// Forward-mode dual numbers (K tangent directions).
// ---------------------------------------------------------------------------------------
template <int K>
struct Dual {
  double v;
  double d[K];
  Dual() : v(0.0) { for (int k = 0; k < K; ++k) d[k] = 0.0; }
  Dual(double x) : v(x) { for (int k = 0; k < K; ++k) d[k] = 0.0; }
};
template <int K> inline Dual<K> mk(double v, const double* d) { Dual<K> r(v); for (int k = 0; k < K; ++k) r.d[k] = d[k]; return r; }
template <int K> inline Dual<K> operator+(const Dual<K>& a, const Dual<K>& b) { Dual<K> r(a.v + b.v); for (int k = 0; k < K; ++k) r.d[k] = a.d[k] + b.d[k]; return r; }
template <int K> inline Dual<K> operator-(const Dual<K>& a, const Dual<K>& b) { Dual<K> r(a.v - b.v); for (int k = 0; k < K; ++k) r.d[k] = a.d[k] - b.d[k]; return r; }
template <int K> inline Dual<K> operator-(const Dual<K>& a) { Dual<K> r(-a.v); for (int k = 0; k < K; ++k) r.d[k] = -a.d[k]; return r; }
template <int K> inline Dual<K> operator*(const Dual<K>& a, const Dual<K>& b) { Dual<K> r(a.v * b.v); for (int k = 0; k < K; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k]; return r; }
template <int K> inline Dual<K> operator/(const Dual<K>& a, const Dual<K>& b) { Dual<K> r(a.v / b.v); for (int k = 0; k < K; ++k) r.d[k] = (a.d[k] - r.v * b.d[k]) / b.v; return r; }
template <int K> inline Dual<K> operator+(const Dual<K>& a, double b) { return a + Dual<K>(b); }
template <int K> inline Dual<K> operator+(double a, const Dual<K>& b) { return Dual<K>(a) + b; }
template <int K> inline Dual<K> operator-(const Dual<K>& a, double b) { return a - Dual<K>(b); }
template <int K> inline Dual<K> operator-(double a, const Dual<K>& b) { return Dual<K>(a) - b; }
template <int K> inline Dual<K> operator*(const Dual<K>& a, double b) { return a * Dual<K>(b); }
template <int K> inline Dual<K> operator*(double a, const Dual<K>& b) { return Dual<K>(a) * b; }
template <int K> inline Dual<K> operator/(const Dual<K>& a, double b) { return a / Dual<K>(b); }
template <int K> inline Dual<K> operator/(double a, const Dual<K>& b) { return Dual<K>(a) / b; }
template <int K> inline bool operator<(const Dual<K>& a, const Dual<K>& b) { return a.v < b.v; }
template <int K> inline bool operator>(const Dual<K>& a, const Dual<K>& b) { return a.v > b.v; }
template <int K> inline bool operator<=(const Dual<K>& a, const Dual<K>& b) { return a.v <= b.v; }
template <int K> inline bool operator>=(const Dual<K>& a, const Dual<K>& b) { return a.v >= b.v; }
template <int K> inline bool operator==(const Dual<K>& a, const Dual<K>& b) { return a.v == b.v; }
template <int K> inline bool operator!=(const Dual<K>& a, const Dual<K>& b) { return a.v != b.v; }
template <int K> inline Dual<K> chain(const Dual<K>& a, double fv, double dfdx) { Dual<K> r(fv); for (int k = 0; k < K; ++k) r.d[k] = dfdx * a.d[k]; return r; }
template <int K> inline Dual<K> sqrt(const Dual<K>& a) { double s = std::sqrt(a.v); return chain(a, s, 0.5 / s); }
template <int K> inline Dual<K> sin(const Dual<K>& a) { return chain(a, std::sin(a.v), std::cos(a.v)); }
template <int K> inline Dual<K> cos(const Dual<K>& a) { return chain(a, std::cos(a.v), -std::sin(a.v)); }
template <int K> inline Dual<K> tan(const Dual<K>& a) { double t = std::tan(a.v); return chain(a, t, 1.0 + t * t); }
template <int K> inline Dual<K> asin(const Dual<K>& a) { return chain(a, std::asin(a.v), 1.0 / std::sqrt(1.0 - a.v * a.v)); }
template <int K> inline Dual<K> atan(const Dual<K>& a) { return chain(a, std::atan(a.v), 1.0 / (1.0 + a.v * a.v)); }
template <int K> inline Dual<K> fabs(const Dual<K>& a) { return a.v < 0 ? -a : a; }
template <int K> inline Dual<K> exp(const Dual<K>& a) { double e = std::exp(a.v); return chain(a, e, e); }
template <int K> inline Dual<K> sinh(const Dual<K>& a) { return chain(a, std::sinh(a.v), std::cosh(a.v)); }
template <int K> inline Dual<K> cosh(const Dual<K>& a) { return chain(a, std::cosh(a.v), std::sinh(a.v)); }
template <int K> inline bool isnan_(const Dual<K>& a) { return std::isnan(a.v); }
template <int K> inline double val(const Dual<K>& a) { return a.v; }

inline bool isnan_(double a) { return std::isnan(a); }
inline double val(double a) { return a; }
inline bool isnan_(long double a) { return std::isnan(a); }
inline double val(long double a) { return (double)a; }
using std::sqrt; using std::sin; using std::cos; using std::tan; using std::asin; using std::atan; using std::fabs;
using std::exp; using std::sinh; using std::cosh;

// Quad precision (__float128, libquadmath): the yardstick for finite differences, whose
// rounding noise must stay far below the differences being resolved.
typedef __float128 f128;
inline f128 sqrt(f128 a) { return sqrtq(a); }
inline f128 sin(f128 a) { return sinq(a); }
inline f128 cos(f128 a) { return cosq(a); }
inline f128 tan(f128 a) { return tanq(a); }
inline f128 asin(f128 a) { return asinq(a); }
inline f128 atan(f128 a) { return atanq(a); }
inline f128 fabs(f128 a) { return fabsq(a); }
inline f128 exp(f128 a) { return expq(a); }
inline f128 sinh(f128 a) { return sinhq(a); }
inline f128 cosh(f128 a) { return coshq(a); }
inline bool isnan_(f128 a) { return isnanq(a) != 0; }
inline double val(f128 a) { return (double)a; }

// ---------------------------------------------------------------------------------------
// Raw element description: the fields of the reference's element structs
// (src/elements/{drift,quad,multipole,bend,rf}.jl), one union-like record.
// ---------------------------------------------------------------------------------------
enum OrcKind {
  ORC_MARKER = 0,         // src/elements/marker.jl
  ORC_DRIFT = 1,          // src/elements/drift.jl:6-30
  ORC_QUAD = 2,           // src/elements/quad.jl:6-47
  ORC_SEXT = 3,           // src/elements/multipole.jl:6-54
  ORC_OCT = 4,            // src/elements/multipole.jl:61-109
  ORC_THINMP = 5,         // src/elements/multipole.jl:116-162
  ORC_SBEND = 6,          // src/elements/bend.jl:6-74
  ORC_EXACTSBEND = 7,     // src/elements/bend.jl:88-149
  ORC_RFCAVITY = 8,       // src/elements/rf.jl:6-21
  ORC_EXACTSBEND_LE = 9,  // exact body (src/tracking.jl:858-894) with the linear edge maps
                          // of src/tracking.jl:160-221 (no reference element; composition)
  // the other families; their extra fields travel in orc_elem.x (listed per kind below)
  ORC_SOLENOID = 10,      // src/elements/solenoid.jl:6-26       L, k = ks
  ORC_CORRECTOR = 11,     // src/elements/corrector.jl:6-27      L, x = xkick, ykick
  ORC_TRANSLATION = 12,   // src/elements/SCelements.jl:420-433  x = dx, dy, ds
  ORC_YROTATION = 13,     // src/elements/SCelements.jl:444-455  angle
  ORC_CRAB = 14,          // src/elements/rf.jl:35-54            L, volt, energy, x = k, phi, err1, err2
  ORC_ACCEL = 15,         // src/elements/rf.jl:69-87            volt, energy, x = k, phis
  ORC_LONGRF = 16,        // src/elements/rf.jl:102-105          x = alphac, rf kind, rf freq, rf k, rf h
  ORC_BOOST = 17,         // src/elements/beambeam.jl:6-19       x = cosang, tanang, mode
  ORC_INVBOOST = 18,      // src/elements/beambeam.jl:30-43      x = sinang, cosang, mode
  ORC_THINBB = 19,        // src/elements/beambeam.jl:54-69      x = amplitude, sx, sy, xoff, yoff
  ORC_SGB = 20,           // src/elements/beambeam.jl:83-119     x = charge, total_energy, sx, sy, n,
                          //                                     [8..) npar[n], xoff[n], yoff[n]
  ORC_RLCWAKE = 21,       // src/elements/wake.jl:6-36           x = freq, Rshunt, Q0, scale
  ORC_TABWAKE = 22,       // src/elements/wake.jl:44-91          x = scale, fliphalf, n, times[n], wakes[n]
  ORC_DRIFTSC = 23,       // src/elements/SCelements.jl:6-38     L
  ORC_QUADSC = 24,        // src/elements/SCelements.jl:51-101   L, x = k0..k3, polynom_a, ...
  ORC_SEXTSC = 25,        // src/elements/SCelements.jl:116-166
  ORC_OCTSC = 26,         // src/elements/SCelements.jl:181-231
  ORC_SBENDSC = 27,       // src/elements/SCelements.jl:246-311  as SBend
  ORC_LBEND = 28,         // src/elements/SCelements.jl:346-379  L, angle, e1, e2, fint1, fint2, gap =
                          //                                     full_gap, x = K, by_error
  ORC_SPACECHARGE = 29,   // src/elements/SCelements.jl:394-408
  ORC_WIGGLER = 30,       // src/elements/wiggler.jl:6-44         L, energy, x = lw, Bmax, Nsteps, rad,
                          //   NHharm, NVharm, By[6 NHharm], Bx[6 NVharm]
};
#define ORC_NX 128

extern "C" {
typedef struct orc_elem {
  int32_t kind, max_order, num_int_steps, fringe_bend_entrance;
  int32_t fringe_bend_exit, has_r1, has_r2, pad;
  double L, k, angle, e1, e2, fint1, fint2, gap, gk;
  double kick_angle[2];
  double polynom_a[ORC_NP], polynom_b[ORC_NP];
  double t1[6], t2[6];
  double r1[36], r2[36];  // column-major, as SMatrix{6,6}
  double volt, freq, h, lag, philag, energy;
  double x[ORC_NX];  // per-kind extra fields (OrcKind)
} orc_elem;
}

template <class T>
struct Elem {
  int kind, max_order, num_int_steps, fbe, fbx, has_r1, has_r2;
  T L, k, angle, e1, e2, fint1, fint2, gap, gk;
  T kick_angle[2];
  T pa[ORC_NP], pb[ORC_NP];
  T t1[6], t2[6];
  T r1[36], r2[36];
  T volt, freq, h, lag, philag, energy;
  T x[ORC_NX];
};

template <class T>
static void load_elem(const orc_elem& e, Elem<T>& o) {
  o.kind = e.kind; o.max_order = e.max_order; o.num_int_steps = e.num_int_steps;
  o.fbe = e.fringe_bend_entrance; o.fbx = e.fringe_bend_exit; o.has_r1 = e.has_r1; o.has_r2 = e.has_r2;
  o.L = T(e.L); o.k = T(e.k); o.angle = T(e.angle); o.e1 = T(e.e1); o.e2 = T(e.e2);
  o.fint1 = T(e.fint1); o.fint2 = T(e.fint2); o.gap = T(e.gap); o.gk = T(e.gk);
  o.kick_angle[0] = T(e.kick_angle[0]); o.kick_angle[1] = T(e.kick_angle[1]);
  for (int i = 0; i < ORC_NP; ++i) { o.pa[i] = T(e.polynom_a[i]); o.pb[i] = T(e.polynom_b[i]); }
  for (int i = 0; i < 6; ++i) { o.t1[i] = T(e.t1[i]); o.t2[i] = T(e.t2[i]); }
  for (int i = 0; i < 36; ++i) { o.r1[i] = T(e.r1[i]); o.r2[i] = T(e.r2[i]); }
  o.volt = T(e.volt); o.freq = T(e.freq); o.h = T(e.h); o.lag = T(e.lag); o.philag = T(e.philag); o.energy = T(e.energy);
  for (int i = 0; i < ORC_NX; ++i) o.x[i] = T(e.x[i]);
}

template <class T> struct R6 { T v[6]; };

// Yoshida-4 coefficients: src/tracking.jl:29-32 (decimal literals parsed to Float64).
static const double DRIFT1 = 0.6756035959798286638;
static const double DRIFT2 = -0.1756035959798286639;
static const double KICK1 = 1.351207191959657328;
static const double KICK2 = -1.702414383919314656;
static const double C_LIGHT = 2.99792458e8;     // src/elements/rf.jl:349
static const double TWO_PI = 6.283185307179586;  // Julia Float64(2pi)

// check_lost: src/tracking.jl:49-52 (COORD_LIMIT = ANGLE_LIMIT = 1.0, :25-26)
template <class T>
static inline bool check_lost(const R6<T>& r) {
  return isnan_(r.v[0]) || fabs(r.v[0]) > T(1.0) || fabs(r.v[2]) > T(1.0) ||
         fabs(r.v[1]) > T(1.0) || fabs(r.v[3]) > T(1.0);
}

template <class T>
static inline R6<T> nan6() { R6<T> o; for (int i = 0; i < 6; ++i) o.v[i] = T(NAN); return o; }

// apply_misalignment: src/tracking.jl:60-67 (r + t), and R*r for a nonzero R
// (StaticArrays SMatrix*SVector; summed left to right, order parity unpinned).
template <class T>
static inline R6<T> add6(const R6<T>& r, const T* t) { R6<T> o; for (int i = 0; i < 6; ++i) o.v[i] = r.v[i] + t[i]; return o; }
template <class T>
static inline R6<T> mat6(const T* M, const R6<T>& r) {
  R6<T> o;
  for (int i = 0; i < 6; ++i) {
    T s = M[i] * r.v[0];
    for (int j = 1; j < 6; ++j) s = s + M[i + 6 * j] * r.v[j];
    o.v[i] = s;
  }
  return o;
}
template <class T>
static inline R6<T> entrance(const Elem<T>& e, R6<T> r) {
  r = add6(r, e.t1);
  if (e.has_r1) r = mat6(e.r1, r);
  return r;
}
template <class T>
static inline R6<T> exit_(const Elem<T>& e, R6<T> r) {
  if (e.has_r2) r = mat6(e.r2, r);
  return add6(r, e.t2);
}

// drift6: src/tracking.jl:87-108
template <class T>
static inline R6<T> drift6(const R6<T>& r, T L, T beti, bool exact) {
  R6<T> o = r;
  if (exact) {
    T pz2 = T(1.0) + T(2.0) * r.v[5] * beti + r.v[5] * r.v[5] - r.v[1] * r.v[1] - r.v[3] * r.v[3];
    if (pz2 <= T(0.0)) return nan6<T>();
    T NormL = L / sqrt(pz2);
    o.v[0] = r.v[0] + NormL * r.v[1];
    o.v[2] = r.v[2] + NormL * r.v[3];
    o.v[4] = r.v[4] + NormL * (beti + r.v[5]) - L * beti;
  } else {
    T NormL = L / (T(1.0) + r.v[5]);
    o.v[0] = r.v[0] + NormL * r.v[1];
    o.v[2] = r.v[2] + NormL * r.v[3];
    o.v[4] = r.v[4] + NormL * (r.v[1] * r.v[1] + r.v[3] * r.v[3]) / (T(2.0) * (T(1.0) + r.v[5]));
  }
  return o;
}

// edge_fringe_entrance / edge_fringe_exit: src/tracking.jl:160-221
template <class T>
static inline R6<T> edge_fringe(const R6<T>& r, T inv_rho, T edge_angle, T fint, T gap, int method, bool is_exit) {
  T fringecorr;
  if (fint == T(0.0) || gap == T(0.0) || method == 0) {
    fringecorr = T(0.0);
  } else {
    T sedge = sin(edge_angle);
    T cedge = cos(edge_angle);
    fringecorr = inv_rho * gap * fint * (T(1.0) + sedge * sedge) / cedge;
  }
  T fx = inv_rho * tan(edge_angle);
  T fy;
  if (method == 1) {
    fy = inv_rho * tan(edge_angle - fringecorr / (T(1.0) + r.v[5]));
  } else if (method == 2) {
    fy = inv_rho * tan(edge_angle - fringecorr / (T(1.0) + r.v[5])) / (T(1.0) + r.v[5]);
  } else if (method == 3) {
    if (!is_exit) fy = inv_rho * tan(edge_angle - fringecorr + r.v[1] / (T(1.0) + r.v[5]));
    else          fy = inv_rho * tan(edge_angle - fringecorr - r.v[1] / (T(1.0) + r.v[5]));
  } else {
    fy = inv_rho * tan(edge_angle - fringecorr / (T(1.0) + r.v[5]));
  }
  R6<T> o = r;
  o.v[1] = r.v[1] + r.v[0] * fx;
  o.v[3] = r.v[3] - r.v[2] * fy;
  return o;
}

// bndthinkick: src/tracking.jl:243-270
template <class T>
static inline R6<T> bndthinkick(const R6<T>& r, const T* pa, const T* pb, T L, T irho, int max_order, T beti) {
  T ReSum = pb[max_order];
  T ImSum = pa[max_order];
  for (int i = max_order - 1; i >= 0; --i) {
    T ReSumTemp = ReSum * r.v[0] - ImSum * r.v[2] + pb[i];
    ImSum = ImSum * r.v[0] + ReSum * r.v[2] + pa[i];
    ReSum = ReSumTemp;
  }
  R6<T> o = r;
  o.v[1] = r.v[1] - L * (ReSum - (r.v[5] - r.v[0] * irho) * irho);
  o.v[3] = r.v[3] + L * ImSum;
  o.v[4] = r.v[4] + L * irho * r.v[0] * beti;
  return o;
}

// strthinkick: src/tracking.jl:284-305
template <class T>
static inline R6<T> strthinkick(const R6<T>& r, const T* pa, const T* pb, T L, int max_order) {
  T ReSum = pb[max_order];
  T ImSum = pa[max_order];
  for (int i = max_order - 1; i >= 0; --i) {
    T ReSumTemp = ReSum * r.v[0] - ImSum * r.v[2] + pb[i];
    ImSum = ImSum * r.v[0] + ReSum * r.v[2] + pa[i];
    ReSum = ReSumTemp;
  }
  R6<T> o = r;
  o.v[1] = r.v[1] - L * ReSum;
  o.v[3] = r.v[3] + L * ImSum;
  return o;
}

// symplectic4_pass: src/tracking.jl:317-341
template <class T>
static inline R6<T> symplectic4_pass(R6<T> r, T L, const T* pa, const T* pb, int max_order, int num_steps, T beti, bool exact) {
  T SL = L / T(double(num_steps));
  T L1 = SL * T(DRIFT1);
  T L2 = SL * T(DRIFT2);
  T K1 = SL * T(KICK1);
  T K2 = SL * T(KICK2);
  for (int s = 0; s < num_steps; ++s) {
    r = drift6(r, L1, beti, exact);
    r = strthinkick(r, pa, pb, K1, max_order);
    r = drift6(r, L2, beti, exact);
    r = strthinkick(r, pa, pb, K2, max_order);
    r = drift6(r, L2, beti, exact);
    r = strthinkick(r, pa, pb, K1, max_order);
    r = drift6(r, L1, beti, exact);
  }
  return r;
}

// symplectic4_bend_pass: src/tracking.jl:350-375
template <class T>
static inline R6<T> symplectic4_bend_pass(R6<T> r, T L, const T* pa, const T* pb, T irho, int max_order, int num_steps, T beti, bool exact) {
  T SL = L / T(double(num_steps));
  T L1 = SL * T(DRIFT1);
  T L2 = SL * T(DRIFT2);
  T K1 = SL * T(KICK1);
  T K2 = SL * T(KICK2);
  for (int s = 0; s < num_steps; ++s) {
    r = drift6(r, L1, beti, exact);
    r = bndthinkick(r, pa, pb, K1, irho, max_order, beti);
    r = drift6(r, L2, beti, exact);
    r = bndthinkick(r, pa, pb, K2, irho, max_order, beti);
    r = drift6(r, L2, beti, exact);
    r = bndthinkick(r, pa, pb, K1, irho, max_order, beti);
    r = drift6(r, L1, beti, exact);
  }
  return r;
}

// pxyz: src/tracking.jl:717-720
template <class T>
static inline T pxyz(T dp1, T px, T py) {
  T v = dp1 * dp1 - px * px - py * py;
  return v > T(0.0) ? sqrt(v) : T(0.0);
}

// yrot: src/tracking.jl:727-745
template <class T>
static inline R6<T> yrot(const R6<T>& r, T phi, T beti) {
  if (phi == T(0.0)) return r;
  T dp1 = beti + r.v[5];
  T c = cos(phi);
  T s = sin(phi);
  T pz = pxyz(dp1, r.v[1], r.v[3]);
  T p = c * pz - s * r.v[1];
  R6<T> o = r;
  o.v[1] = s * pz + c * r.v[1];
  o.v[0] = r.v[0] * pz / p;
  o.v[2] = r.v[2] + r.v[0] * r.v[3] * s / p;
  o.v[4] = r.v[4] + dp1 * r.v[0] * s / p;
  return o;
}

// bend_fringe: src/tracking.jl:752-815 (Forest 13.13)
template <class T>
static inline R6<T> bend_fringe(const R6<T>& r, T irho, T gK, T beti) {
  T b0 = irho;
  T dp1 = beti + r.v[5];
  T pz = pxyz(dp1, r.v[1], r.v[3]);
  T px = r.v[1];
  T py = r.v[3];
  T xp = px / pz;
  T yp = py / pz;
  T term1 = T(1.0) + xp * xp * (T(2.0) + yp * yp);
  T term2 = xp / (T(1.0) + yp * yp);
  T phi = -b0 * tan(b0 * gK * term1 * pz - atan(term2));
  T px2 = px * px;
  T py2 = py * py;
  T pz2 = pz * pz;
  T pz4 = pz2 * pz2;
  T py2z2 = (py2 + pz2);
  T arg_sec = (b0 * gK * (pz4 + px2 * (py2 + T(2.0) * pz2))) / (pz2 * pz) - atan(px * pz / (py2 + pz2));
  T secv = T(1.0) / cos(arg_sec);
  T powsec = secv * secv;
  T py4 = py2 * py2;
  T py6 = py4 * py2;
  T pz5 = pz4 * pz;
  T pz6 = pz4 * pz2;
  T py2z2_sq = py2z2 * py2z2;
  T denom = pz5 * (py4 + px2 * pz2 + T(2.0) * py2 * pz2 + pz4);
  T dpx_num = -(b0 * (px2 * pz4 * (py2 - pz2) - pz6 * (py2 + pz2) +
                      b0 * gK * px * (pz2 * py2z2_sq * (T(2.0) * py2 + T(3.0) * pz2) +
                                      px2 * px2 * (T(3.0) * py2 * pz2 + T(2.0) * pz4) +
                                      px2 * (T(3.0) * py6 + T(8.0) * py4 * pz2 + T(9.0) * py2 * pz4 + T(5.0) * pz6)))) * powsec;
  T dpx = dpx_num / denom;
  T dpy_num = -(b0 * py * (px * pz4 * (py2 + pz2) +
                           b0 * gK * (-(pz4 * py2z2_sq) + px2 * px2 * (T(3.0) * py2 * pz2 + T(4.0) * pz4) +
                                      px2 * (T(3.0) * py6 + T(10.0) * py4 * pz2 + T(11.0) * py2 * pz4 + T(3.0) * pz6)))) * powsec;
  T dpy = dpy_num / denom;
  T dd_num = (b0 * dp1 * (px * pz4 * (py2 - pz2) + b0 * gK *
                          (-(pz4 * py2z2_sq) + px2 * px2 * (T(3.0) * py2 * pz2 + T(2.0) * pz4) +
                           px2 * (T(3.0) * py6 + T(8.0) * py4 * pz2 + T(7.0) * py2 * pz4 + pz6)))) * powsec;
  T dd = dd_num / denom;
  T yf = (T(2.0) * r.v[2]) / (T(1.0) + sqrt(T(1.0) - T(2.0) * dpy * r.v[2]));
  T dxf = T(0.5) * dpx * (yf * yf);
  T dct = T(0.5) * dd * (yf * yf);
  T dpyf = phi * yf;
  R6<T> o;
  o.v[0] = r.v[0] + dxf;
  o.v[1] = r.v[1];
  o.v[2] = yf;
  o.v[3] = r.v[3] - dpyf;
  o.v[4] = r.v[4] - dct;
  o.v[5] = r.v[5];
  return o;
}

template <class T>
static inline T clamp1(T v) { return v > T(1.0) ? T(1.0) : (v < T(-1.0) ? T(-1.0) : v); }

// bend_edge: src/tracking.jl:822-851 (Forest 12.41)
template <class T>
static inline R6<T> bend_edge(const R6<T>& r, T rhoinv, T theta, T beti) {
  if (fabs(rhoinv) < T(1e-6)) return r;
  T dp1 = beti + r.v[5];
  T c = cos(theta);
  T s = sin(theta);
  T pz = pxyz(dp1, r.v[1], r.v[3]);
  T d2 = pxyz(dp1, T(0.0), r.v[3]);
  T px_new = r.v[1] * c + (pz - rhoinv * r.v[0]) * s;
  T val1 = clamp1(r.v[1] / d2);
  T val2 = clamp1(px_new / d2);
  T dasin = asin(val1) - asin(val2);
  T num = r.v[0] * (r.v[1] * sin(T(2.0) * theta) + s * s * (T(2.0) * pz - rhoinv * r.v[0]));
  T den = pxyz(dp1, px_new, r.v[3]) + pz * c - r.v[1] * s;
  R6<T> o = r;
  o.v[0] = r.v[0] * c + num / den;
  o.v[1] = px_new;
  o.v[2] = r.v[2] + r.v[3] * (theta / rhoinv + dasin / rhoinv);
  o.v[4] = r.v[4] + dp1 / rhoinv * (theta + dasin);
  return o;
}

// exact_bend_body: src/tracking.jl:858-894 (Forest 12.18)
template <class T>
static inline R6<T> exact_bend_body(const R6<T>& r, T irho, T L, T beti) {
  T dp1 = beti + r.v[5];
  T pz = pxyz(dp1, r.v[1], r.v[3]);
  R6<T> o = r;
  if (fabs(irho) < T(1e-6)) {
    T NormL = L / pz;
    o.v[0] = r.v[0] + r.v[1] * NormL;
    o.v[2] = r.v[2] + r.v[3] * NormL;
    o.v[4] = r.v[4] + NormL * dp1 - L * beti;
    return o;
  }
  T pzmx = pz - (T(1.0) + r.v[0] * irho);
  T cs = cos(irho * L);
  T sn = sin(irho * L);
  T px_new = r.v[1] * cs + pzmx * sn;
  T d2 = pxyz(dp1, T(0.0), r.v[3]);
  T val1 = clamp1(r.v[1] / d2);
  T val2 = clamp1(px_new / d2);
  T dasin = L + (asin(val1) - asin(val2)) / irho;
  o.v[0] = (pxyz(dp1, px_new, r.v[3]) - pzmx * cs + r.v[1] * sn - T(1.0)) / irho;
  o.v[1] = px_new;
  o.v[2] = r.v[2] + r.v[3] * dasin;
  o.v[4] = r.v[4] + dp1 * dasin - L * beti;
  return o;
}

// Polynomials of a magnet after the reference's per-call substitutions:
// Quadrupole b[1] <- k1 (src/tracking.jl:395), Sextupole b[2] <- k2/2 (:446),
// Octupole b[3] <- k3/6 (:478); kick-angle correction for Quad/SBend/ExactSBend when
// L > 0 (:398-414, :670-683, :935-942).
template <class T>
static inline void magnet_polys(const Elem<T>& e, T* pa, T* pb, bool kick_corr) {
  for (int i = 0; i < ORC_NP; ++i) { pa[i] = e.pa[i]; pb[i] = e.pb[i]; }
  if (e.kind == ORC_QUAD) pb[1] = e.k;
  if (e.kind == ORC_SEXT) pb[2] = e.k / T(2.0);
  if (e.kind == ORC_OCT) pb[3] = e.k / T(6.0);
  if (kick_corr && e.L > T(0.0)) {
    pb[0] = pb[0] - sin(e.kick_angle[0]) / e.L;
    pa[0] = pa[0] + sin(e.kick_angle[1]) / e.L;
  }
}

static const double EPS64 = 2.220446049250313e-16;      // eps(Float64)
static const double SQRT_EPS64 = 1.4901161193847656e-08;  // sqrt(eps(Float64))

template <class T> static inline T tmax(T a, T b) { return a > b ? a : b; }

// src/elements/wake.jl:67-84
template <class T>
static T linear_interpolate(T x, const T* xs, const T* ys, int n) {
  if (x <= xs[0]) {
    T slope = (ys[1] - ys[0]) / (xs[1] - xs[0]);
    return ys[0] + slope * (x - xs[0]);
  } else if (x >= xs[n - 1]) {
    T slope = (ys[n - 1] - ys[n - 2]) / (xs[n - 1] - xs[n - 2]);
    return ys[n - 1] + slope * (x - xs[n - 1]);
  }
  for (int i = 1; i < n; ++i) {
    if (x < xs[i]) {
      T slope = (ys[i] - ys[i - 1]) / (xs[i] - xs[i - 1]);
      return ys[i - 1] + slope * (x - xs[i - 1]);
    }
  }
  return ys[n - 1];
}

// _lbend_body: src/tracking.jl:1076-1150
template <class T>
static R6<T> lbend_body(R6<T> r, T L, T grd, T b_angle, T by_error) {
  if (L == T(0.0)) return r;
  T p_norm = T(1.0) / (T(1.0) + r.v[5]);
  T Kx = b_angle / L;
  T G1 = (Kx * Kx + grd) * p_norm;
  T G2 = -grd * p_norm;
  T tol = T(SQRT_EPS64);
  T MHD = T(1.0), M12 = L, M21 = T(0.0);
  T MVD = T(1.0), M34 = L, M43 = T(0.0);
  T arg1 = T(0.0), arg2 = T(0.0), sqrtG1 = T(0.0), sqrtG2 = T(0.0);
  if (fabs(G1) >= tol) {
    if (G1 > T(0.0)) {
      sqrtG1 = sqrt(G1); arg1 = L * sqrtG1;
      MHD = cos(arg1); M12 = sin(arg1) / sqrtG1; M21 = -sin(arg1) * sqrtG1;
    } else {
      sqrtG1 = sqrt(-G1); arg1 = L * sqrtG1;
      MHD = cosh(arg1); M12 = sinh(arg1) / sqrtG1; M21 = sinh(arg1) * sqrtG1;
    }
  }
  if (fabs(G2) >= tol) {
    if (G2 > T(0.0)) {
      sqrtG2 = sqrt(G2); arg2 = L * sqrtG2;
      MVD = cos(arg2); M34 = sin(arg2) / sqrtG2; M43 = -sin(arg2) * sqrtG2;
    } else {
      sqrtG2 = sqrt(-G2); arg2 = L * sqrtG2;
      MVD = cosh(arg2); M34 = sinh(arg2) / sqrtG2; M43 = sinh(arg2) * sqrtG2;
    }
  }
  T x = r.v[0], xpr = r.v[1] * p_norm, y = r.v[2], ypr = r.v[3] * p_norm, delta = r.v[5];
  T dterm = delta * p_norm - by_error;
  T x_new = MHD * x + M12 * xpr;
  T px_new = (M21 * x + MHD * xpr) / p_norm;
  if (fabs(G1) < tol) {
    x_new = x_new + dterm * L * L * Kx / T(2.0);
    px_new = px_new + dterm * L * Kx / p_norm;
  } else if (G1 > T(0.0)) {
    x_new = x_new + dterm * (T(1.0) - cos(arg1)) * Kx / G1;
    px_new = px_new + dterm * sin(arg1) * Kx / (sqrtG1 * p_norm);
  } else {
    x_new = x_new + dterm * (T(1.0) - cosh(arg1)) * Kx / G1;
    px_new = px_new + dterm * sinh(arg1) * Kx / (sqrtG1 * p_norm);
  }
  T y_new = MVD * y + M34 * ypr;
  T py_new = (M43 * y + MVD * ypr) / p_norm;
  T z_new = r.v[4] + xpr * xpr * (L + MHD * M12) / T(4.0);
  if (fabs(G1) >= tol) {
    z_new = z_new + (L - MHD * M12) * (x * x * G1 + dterm * dterm * Kx * Kx / G1 - T(2.0) * x * Kx * dterm) / T(4.0);
    z_new = z_new + M12 * M21 * (x * xpr - xpr * dterm * Kx / G1) / T(2.0);
    z_new = z_new + (Kx * x * M12 + xpr * (T(1.0) - MHD) * Kx / G1 + dterm * (L - M12) * Kx * Kx / G1);
  }
  z_new = z_new + ((L - MVD * M34) * y * y * G2 + ypr * ypr * (L + MVD * M34)) / T(4.0);
  z_new = z_new + M34 * M43 * y * ypr / T(2.0);
  R6<T> o = {{x_new, px_new, y_new, py_new, z_new, r.v[5]}};
  return o;
}

// Wiggler: src/tracking.jl:1224-1525.  The harmonic tables are read from the raw By / Bx
// integers every call, as the reference's _wig_* functions read them.
static const double M_ELECTRON_EV = 0.51099895069e6;  // src/lattice.jl:42

template <class T>
struct Wig {
  const Elem<T>* e;
  int nh, nv;
  T lw, Bmax, Aw, Po, kw;
  const T* by() const { return e->x + 6; }
  const T* bx() const { return e->x + 6 + 6 * nh; }
};

template <class T>
static T sinc_taylor(T x) {  // :1224-1227
  T x2 = x * x;
  return T(1.0) - x2 / T(6.0) * (T(1.0) - x2 / T(20.0) * (T(1.0) - x2 / T(42.0) * (T(1.0) - x2 / T(72.0))));
}

template <class T>
static void wig_ax_axpy(const Wig<T>& w, T x, T y, T Zw, T& ax, T& axpy) {  // :1229-1279
  T kw = w.kw;
  ax = T(0.0);
  axpy = T(0.0);
  for (int i = 0; i < w.nh; ++i) {
    const T* b = w.by() + 6 * i;
    T HCw = b[1] * w.Aw / w.Po;
    T kx = b[2] * kw, ky = b[3] * kw, kz = b[4] * kw, tz = b[5];
    T cx = cos(kx * x), chy = cosh(ky * y), sz = sin(kz * Zw + tz);
    ax = ax + HCw * (kw / kz) * cx * chy * sz;
    T shy = sinh(ky * y);
    T sxkx = fabs(kx / kw) > T(1e-6) ? sin(kx * x) / kx : x * sinc_taylor(kx * x);
    axpy = axpy + HCw * (kw / kz) * ky * sxkx * shy * sz;
  }
  for (int i = 0; i < w.nv; ++i) {
    const T* b = w.bx() + 6 * i;
    T VCw = b[1] * w.Aw / w.Po;
    T kx = b[2] * kw, ky = b[3] * kw, kz = b[4] * kw, tz = b[5];
    T shx = sinh(kx * x), sy = sin(ky * y), sz = sin(kz * Zw + tz);
    ax = ax + VCw * (kw / kz) * (ky / kx) * shx * sy * sz;
    T chx = cosh(kx * x), cy = cos(ky * y);
    axpy = axpy + VCw * (kw / kz) * ((ky / kx) * (ky / kx)) * chx * cy * sz;
  }
}

template <class T>
static void wig_ay_aypx(const Wig<T>& w, T x, T y, T Zw, T& ay, T& aypx) {  // :1281-1331
  T kw = w.kw;
  ay = T(0.0);
  aypx = T(0.0);
  for (int i = 0; i < w.nh; ++i) {
    const T* b = w.by() + 6 * i;
    T HCw = b[1] * w.Aw / w.Po;
    T kx = b[2] * kw, ky = b[3] * kw, kz = b[4] * kw, tz = b[5];
    T sx = sin(kx * x), shy = sinh(ky * y), sz = sin(kz * Zw + tz);
    ay = ay + HCw * (kw / kz) * (kx / ky) * sx * shy * sz;
    T cx = cos(kx * x), chy = cosh(ky * y);
    aypx = aypx + HCw * (kw / kz) * ((kx / ky) * (kx / ky)) * cx * chy * sz;
  }
  for (int i = 0; i < w.nv; ++i) {
    const T* b = w.bx() + 6 * i;
    T VCw = b[1] * w.Aw / w.Po;
    T kx = b[2] * kw, ky = b[3] * kw, kz = b[4] * kw, tz = b[5];
    T chx = cosh(kx * x), cy = cos(ky * y), sz = sin(kz * Zw + tz);
    ay = ay + VCw * (kw / kz) * chx * cy * sz;
    T shx = sinh(kx * x);
    T syky = fabs(ky / kw) > T(1e-6) ? sin(ky * y) / ky : y * sinc_taylor(ky * y);
    aypx = aypx + VCw * (kw / kz) * kx * shx * syky * sz;
  }
}

template <class T>
static R6<T> wig_map_2nd(const Wig<T>& w, R6<T> r, T dl, T& Zw) {  // :1333-1379
  T delta = r.v[5];
  T inv1pd = T(1.0) / (T(1.0) + delta);
  T dld = dl * inv1pd, dl2 = dl / T(2.0), dl2d = dl2 * inv1pd;
  Zw = Zw + dl2;
  T x = r.v[0], px = r.v[1], y = r.v[2], py = r.v[3], z = r.v[4];
  T a, b;
  wig_ay_aypx(w, x, y, Zw, a, b);
  px = px - b; py = py - a;
  y = y + dl2d * py;
  z = z + (dl2d / T(2.0)) * (py * py) * inv1pd;
  wig_ay_aypx(w, x, y, Zw, a, b);
  px = px + b; py = py + a;
  wig_ax_axpy(w, x, y, Zw, a, b);
  px = px - a; py = py - b;
  x = x + dld * px;
  z = z + (dld / T(2.0)) * (px * px) * inv1pd;
  wig_ax_axpy(w, x, y, Zw, a, b);
  px = px + a; py = py + b;
  wig_ay_aypx(w, x, y, Zw, a, b);
  px = px - b; py = py - a;
  y = y + dl2d * py;
  z = z + (dl2d / T(2.0)) * (py * py) * inv1pd;
  wig_ay_aypx(w, x, y, Zw, a, b);
  px = px + b; py = py + a;
  Zw = Zw + dl2;
  R6<T> o = {{x, px, y, py, z, delta}};
  return o;
}

template <class T>
static void wig_B(const Wig<T>& w, T x, T y, T Zw, T& Bx, T& By) {  // :1401-1457 (Bz unused)
  T kw = w.kw, PB0 = w.Bmax;
  Bx = T(0.0);
  By = T(0.0);
  for (int i = 0; i < w.nh; ++i) {
    const T* b = w.by() + 6 * i;
    T HCw = b[1];
    T kx = b[2] * kw, ky = b[3] * kw, kz = b[4] * kw, tz = b[5];
    T sx = sin(kx * x), cx = cos(kx * x), chy = cosh(ky * y), shy = sinh(ky * y);
    T cz = cos(kz * Zw + tz);
    Bx = Bx + PB0 * HCw * (kx / ky) * sx * shy * cz;
    By = By - PB0 * HCw * cx * chy * cz;
  }
  for (int i = 0; i < w.nv; ++i) {
    const T* b = w.bx() + 6 * i;
    T VCw = b[1];
    T kx = b[2] * kw, ky = b[3] * kw, kz = b[4] * kw, tz = b[5];
    T shx = sinh(kx * x), chx = cosh(kx * x), sy = sin(ky * y), cy = cos(ky * y);
    T cz = cos(kz * Zw + tz);
    Bx = Bx + PB0 * VCw * chx * cy * cz;
    By = By - PB0 * VCw * (ky / kx) * shx * sy * cz;
  }
}

// the ax / ay transform around _wig_B and _wig_radiation_kicks, :1459-1470, 1491-1496
template <class T>
static R6<T> wig_rad(const Wig<T>& w, R6<T> r, T Zw, T srCoef, T SL) {
  T ax, ay, unused;
  wig_ax_axpy(w, r.v[0], r.v[2], Zw, ax, unused);
  wig_ay_aypx(w, r.v[0], r.v[2], Zw, ay, unused);
  r.v[1] = r.v[1] - ax;
  r.v[3] = r.v[3] - ay;
  T Bx, By;
  wig_B(w, r.v[0], r.v[2], Zw, Bx, By);
  T B2 = Bx * Bx + By * By;
  if (B2 != T(0.0)) {
    T H = w.Po * T(M_ELECTRON_EV) / T(C_LIGHT);
    T irho2 = B2 / (H * H);
    T dFactor = (T(1.0) + r.v[5]) * (T(1.0) + r.v[5]);
    T dDelta = -srCoef * dFactor * irho2 * SL;
    T scale = T(1.0) + dDelta;
    r.v[1] = r.v[1] * scale;
    r.v[3] = r.v[3] * scale;
    r.v[5] = r.v[5] + dDelta;
  }
  r.v[1] = r.v[1] + ax;
  r.v[3] = r.v[3] + ay;
  return r;
}

// _wig_pass_4th / _wig_pass_4th_rad, :1381-1399, 1472-1507
template <class T>
static R6<T> wiggler_pass(const Elem<T>& e, R6<T> r) {
  Wig<T> w;
  w.e = &e;
  w.lw = e.x[0];
  w.Bmax = e.x[1];
  int PN = (int)val(e.x[2]);
  int rad = (int)val(e.x[3]);
  w.nh = (int)val(e.x[4]);
  w.nv = (int)val(e.x[5]);
  long Nw = (long)std::nearbyint(val(e.L) / val(w.lw));  // round(Int, L / lw): ties to even
  long Nstep = (long)PN * std::max(1L, Nw);
  T SL = w.lw / T((double)PN);
  T dl1 = SL * T(1.3512071919596573), dl0 = SL * T(-1.7024143839193146);
  T gamma = e.energy / T(M_ELECTRON_EV);
  w.Po = sqrt(gamma * gamma - T(1.0));
  w.Aw = T(1e-9) * T(C_LIGHT) / (T(M_ELECTRON_EV) * T(1e-9)) / T(TWO_PI) * w.lw * w.Bmax;
  w.kw = T(TWO_PI) / w.lw;
  T srCoef = T(2.0 / 3.0) * T(2.8179403205e-15) * (gamma * gamma * gamma);
  T Zw = T(0.0);
  if (rad == 1) r = wig_rad(w, r, Zw, srCoef, SL);
  for (long k = 0; k < Nstep; ++k) {
    r = wig_map_2nd(w, r, dl1, Zw);
    r = wig_map_2nd(w, r, dl0, Zw);
    r = wig_map_2nd(w, r, dl1, Zw);
    if (rad == 1) r = wig_rad(w, r, Zw, srCoef, SL);
  }
  return r;
}

// pass! of the other element families: src/tracking.jl:529-605, 981-1218, 1531-1663
template <class T>
static R6<T> pass_extra(const Elem<T>& e, R6<T> r, T beti, bool exact) {
  switch (e.kind) {
    case ORC_SOLENOID: {  // :529-566
      r = entrance(e, r);
      T ks = e.k, L = e.L;
      if (ks == T(0.0)) {
        r = drift6(r, L, beti, exact);
      } else {
        T p_norm = T(1.0) / (T(1.0) + r.v[5]);
        T x = r.v[0], xpr = r.v[1] * p_norm, y = r.v[2], ypr = r.v[3] * p_norm;
        T H = ks * p_norm / T(2.0);
        T S = sin(L * H), C = cos(L * H);
        T x_new = x * C * C + xpr * C * S / H + y * C * S + ypr * S * S / H;
        T px_new = (-x * H * C * S + xpr * C * C - y * H * S * S + ypr * C * S) / p_norm;
        T y_new = -x * C * S - xpr * S * S / H + y * C * C + ypr * C * S / H;
        T py_new = (x * H * S * S - xpr * C * S - y * C * S * H + ypr * C * C) / p_norm;
        T z_new = r.v[4] + L * (H * H * (x * x + y * y) + T(2.0) * H * (xpr * y - ypr * x) + xpr * xpr + ypr * ypr) / T(2.0);
        R6<T> o = {{x_new, px_new, y_new, py_new, z_new, r.v[5]}};
        r = o;
      }
      return exit_(e, r);
    }
    case ORC_CORRECTOR: {  // :579-605
      r = entrance(e, r);
      T xk = e.x[0], yk = e.x[1];
      T p_norm = T(1.0) / (T(1.0) + r.v[5]);
      T NormL = e.L * p_norm;
      T z_new = r.v[4] + NormL * p_norm *
                (xk * xk / T(3.0) + yk * yk / T(3.0) + r.v[1] * r.v[1] + r.v[3] * r.v[3] + r.v[1] * xk + r.v[3] * yk) / T(2.0);
      T x_new = r.v[0] + NormL * (r.v[1] + xk / T(2.0));
      T px_new = r.v[1] + xk;
      T y_new = r.v[2] + NormL * (r.v[3] + yk / T(2.0));
      T py_new = r.v[3] + yk;
      R6<T> o = {{x_new, px_new, y_new, py_new, z_new, r.v[5]}};
      return exit_(e, o);
    }
    case ORC_TRANSLATION: {  // :1182-1192
      T pz2 = T(1.0) + T(2.0) * r.v[5] * beti + r.v[5] * r.v[5] - r.v[1] * r.v[1] - r.v[3] * r.v[3];
      if (pz2 <= T(0.0)) return nan6<T>();
      T pz = sqrt(pz2);
      T dx = e.x[0], dy = e.x[1], ds = e.x[2];
      R6<T> o = r;
      o.v[0] = r.v[0] - (dx + ds * r.v[1] / pz);
      o.v[2] = r.v[2] - (dy + ds * r.v[3] / pz);
      o.v[4] = r.v[4] + ds * (beti + r.v[5]) / pz;
      return o;
    }
    case ORC_YROTATION: {  // :1194-1218
      T angle = -e.angle;
      if (angle == T(0.0)) return r;
      T ca = cos(angle), sa = sin(angle), ta = tan(angle);
      T pz2 = T(1.0) + T(2.0) * r.v[5] * beti + r.v[5] * r.v[5] - r.v[1] * r.v[1] - r.v[3] * r.v[3];
      if (pz2 <= T(0.0)) return nan6<T>();
      T pz = sqrt(pz2);
      T ptt = T(1.0) - ta * r.v[1] / pz;
      if (fabs(ptt) < T(SQRT_EPS64)) return nan6<T>();
      R6<T> o = r;
      o.v[0] = r.v[0] / (ca * ptt);
      o.v[1] = ca * r.v[1] + sa * pz;
      o.v[2] = r.v[2] + ta * r.v[0] * r.v[3] / (pz * ptt);
      o.v[4] = r.v[4] + ta * r.v[0] * (beti + r.v[5]) / (pz * ptt);
      return o;
    }
    case ORC_CRAB: {  // :1531-1547
      T k = e.x[0];
      T volt = e.volt * (T(1.0) + e.x[2]);
      T phi = e.x[1] + e.x[3];
      T E = tmax(fabs(e.energy), T(EPS64));
      T ang = k * r.v[4] + phi;
      if (e.L > T(0.0)) r = drift6(r, e.L / T(2.0), beti, exact);
      T px_new = r.v[1] + (volt / E) * sin(ang * beti);
      T delta_new = r.v[5] - (k * volt / E * beti) * r.v[0] * cos(ang * beti);
      r.v[1] = px_new;
      r.v[5] = delta_new;
      if (e.L > T(0.0)) r = drift6(r, e.L / T(2.0), beti, exact);
      return r;
    }
    case ORC_ACCEL: {  // :1549-1559
      T beta = T(1.0) / beti;
      T beta2 = beta * beta;
      if (beta2 <= T(EPS64)) return r;
      T E = tmax(fabs(e.energy), T(EPS64));
      T sv = sin(e.x[0] * r.v[4] + e.x[1]) - sin(e.x[1]);
      r.v[5] = r.v[5] + (e.volt / (beta2 * E)) * sv;
      return r;
    }
    case ORC_LONGRF: {  // :1561-1585
      int rk = (int)val(e.x[1]);
      T k = rk == 1 ? T(TWO_PI) * e.x[2] / T(C_LIGHT) : (rk == 2 || rk == 3 ? e.x[3] : T(0.0));
      if (fabs(k) <= T(EPS64)) return r;
      T h = (rk == 1 || rk == 2) ? e.x[4] : T(1.0);
      T beta = T(1.0) / beti;
      T eta = e.x[0] - (T(1.0) - beta * beta);
      r.v[4] = r.v[4] - (T(TWO_PI) * h * eta / k) * r.v[5];
      return r;
    }
    case ORC_BOOST: {  // :1587-1595
      if (val(e.x[2]) != 0.0) return r;
      T invcos = T(1.0) / e.x[0];
      T tanang = e.x[1];
      T x_new = r.v[0] + tanang * r.v[4];
      T delta_new = r.v[5] - tanang * r.v[1];
      R6<T> o = {{x_new, r.v[1] * invcos, r.v[2], r.v[3] * invcos, r.v[4] * invcos, delta_new}};
      return o;
    }
    case ORC_INVBOOST: {  // :1597-1604
      if (val(e.x[2]) != 0.0) return r;
      T sinang = e.x[0], cosang = e.x[1];
      T x_new = r.v[0] - sinang * r.v[4];
      T delta_new = r.v[5] + sinang * r.v[1];
      R6<T> o = {{x_new, r.v[1] * cosang, r.v[2], r.v[3] * cosang, r.v[4] * cosang, delta_new}};
      return o;
    }
    case ORC_THINBB: {  // :1610-1619
      T sx2 = tmax(e.x[1] * e.x[1], T(EPS64));
      T sy2 = tmax(e.x[2] * e.x[2], T(EPS64));
      T dx = r.v[0] - e.x[3];
      T dy = r.v[2] - e.x[4];
      T g = e.x[0] * exp(T(-0.5) * (dx * dx / sx2 + dy * dy / sy2));
      r.v[1] = r.v[1] - g * dx / sx2;
      r.v[3] = r.v[3] - g * dy / sy2;
      return r;
    }
    case ORC_SGB: {  // :1621-1641
      int n = (int)val(e.x[4]);
      if (n <= 0) return r;
      T sx2 = tmax(e.x[2] * e.x[2], T(EPS64));
      T sy2 = tmax(e.x[3] * e.x[3], T(EPS64));
      T factor = e.x[0] / tmax(fabs(e.x[1]), T(EPS64));
      T px = r.v[1], py = r.v[3];
      for (int i = 0; i < n; ++i) {
        T w = e.x[8 + i], xoff = e.x[8 + n + i], yoff = e.x[8 + 2 * n + i];
        T dx = r.v[0] - xoff;
        T dy = r.v[2] - yoff;
        T g = w * exp(T(-0.5) * (dx * dx / sx2 + dy * dy / sy2));
        px = px - factor * g * dx / sx2;
        py = py - factor * g * dy / sy2;
      }
      r.v[1] = px;
      r.v[3] = py;
      return r;
    }
    case ORC_RLCWAKE: {  // :1647-1654, src/elements/wake.jl:25-36
      T freq = e.x[0], Rshunt = e.x[1], Q0 = e.x[2], scale = e.x[3];
      if (scale == T(0.0)) return r;
      T t = r.v[4] / T(C_LIGHT);
      if (!(t < T(0.0))) t = T(0.0);  // min(z / c, 0)
      T Q0p = sqrt(Q0 * Q0 - T(0.25));
      T w0 = T(TWO_PI) * freq;
      T w0p = w0 / Q0 * Q0p;
      T wake = (t > T(0.0)) ? T(0.0)
                            : Rshunt * w0 / Q0 * (cos(w0p * t) + sin(w0p * t) / (T(2.0) * Q0p)) * exp(w0 * t / (T(2.0) * Q0));
      r.v[5] = r.v[5] - scale * wake;
      return r;
    }
    case ORC_TABWAKE: {  // :1656-1663, src/elements/wake.jl:86-91
      T scale = e.x[0], flip = e.x[1];
      int n = (int)val(e.x[2]);
      if (scale == T(0.0)) return r;
      T t = r.v[4] / T(C_LIGHT);
      const T* ts = e.x + 3;
      const T* ws = e.x + 3 + n;
      T wake = (t > ts[0] * flip) ? T(0.0) : linear_interpolate(t * flip, ts, ws, n);
      r.v[5] = r.v[5] - scale * wake;
      return r;
    }
    case ORC_DRIFTSC:  // :981-991
      r = entrance(e, r);
      r = drift6(r, e.L, beti, exact);
      return exit_(e, r);
    case ORC_QUADSC:   // :997-1015
    case ORC_SEXTSC:   // :1017-1027
    case ORC_OCTSC: {  // :1029-1039
      r = entrance(e, r);
      T pa[ORC_NP], pb[ORC_NP];
      for (int i = 0; i < ORC_NP; ++i) { pa[i] = e.pa[i]; pb[i] = T(0.0); }
      pb[0] = e.x[0]; pb[1] = e.x[1]; pb[2] = e.x[2] / T(2.0); pb[3] = e.x[3] / T(6.0);
      if (e.kind == ORC_QUADSC && e.L > T(0.0)) {
        pb[0] = pb[0] - sin(e.kick_angle[0]) / e.L;
        pa[0] = pa[0] + sin(e.kick_angle[1]) / e.L;
      }
      int m = std::min(std::max(e.max_order, 0), 3);
      r = symplectic4_pass(r, e.L, pa, pb, m, e.num_int_steps, beti, exact);
      return exit_(e, r);
    }
    case ORC_SBENDSC: {  // :1041-1074
      r = entrance(e, r);
      if (e.L == T(0.0)) return nan6<T>();
      T irho = e.angle / e.L;
      T pa[ORC_NP], pb[ORC_NP];
      for (int i = 0; i < ORC_NP; ++i) { pa[i] = e.pa[i]; pb[i] = e.pb[i]; }
      if (e.L > T(0.0)) {
        pb[0] = pb[0] - sin(e.kick_angle[0]) / e.L;
        pa[0] = pa[0] + sin(e.kick_angle[1]) / e.L;
      }
      if (e.fbe != 0) r = edge_fringe(r, irho, e.e1, e.fint1, e.gap, e.fbe, false);
      int m = std::min(std::max(e.max_order, 0), 3);
      r = symplectic4_bend_pass(r, e.L, pa, pb, irho, m, e.num_int_steps, beti, exact);
      if (e.fbx != 0) r = edge_fringe(r, irho, e.e2, e.fint2, e.gap, e.fbx, true);
      return exit_(e, r);
    }
    case ORC_LBEND: {  // :1152-1170
      r = entrance(e, r);
      if (e.L == T(0.0)) return nan6<T>();
      T irho = e.angle / e.L;
      r = edge_fringe(r, irho, e.e1, e.fint1, e.gap, 1, false);
      r = lbend_body(r, e.L, e.x[0], e.angle, e.x[1]);
      r = edge_fringe(r, irho, e.e2, e.fint2, e.gap, 1, true);
      return exit_(e, r);
    }
    case ORC_SPACECHARGE:  // :1176-1180
      return r;
    case ORC_WIGGLER:  // :1509-1525
      r = entrance(e, r);
      r = wiggler_pass(e, r);
      return exit_(e, r);
    default:
      return nan6<T>();
  }
}

// pass!: one element, one particle.  src/tracking.jl:115-141, 386-518, 616-706, 901-973
template <class T>
static R6<T> pass_elem(const Elem<T>& e, R6<T> r, T beti, bool exact) {
  T pa[ORC_NP], pb[ORC_NP];
  switch (e.kind) {
    case ORC_MARKER:  // :139-141
      return r;
    case ORC_DRIFT:  // :115-132
      r = entrance(e, r);
      r = drift6(r, e.L, beti, exact);
      return exit_(e, r);
    case ORC_QUAD:  // :386-427
      r = entrance(e, r);
      magnet_polys(e, pa, pb, true);
      r = symplectic4_pass(r, e.L, pa, pb, e.max_order, e.num_int_steps, beti, exact);
      return exit_(e, r);
    case ORC_SEXT:  // :438-459
    case ORC_OCT:   // :470-491
      r = entrance(e, r);
      magnet_polys(e, pa, pb, false);
      r = symplectic4_pass(r, e.L, pa, pb, e.max_order, e.num_int_steps, beti, exact);
      return exit_(e, r);
    case ORC_THINMP:  // :616-633
      r = entrance(e, r);
      r = strthinkick(r, e.pa, e.pb, T(1.0), e.max_order);
      return exit_(e, r);
    case ORC_SBEND: {  // :657-706
      r = entrance(e, r);
      T irho = e.angle / e.L;
      magnet_polys(e, pa, pb, true);
      if (e.fbe != 0) r = edge_fringe(r, irho, e.e1, e.fint1, e.gap, e.fbe, false);
      r = symplectic4_bend_pass(r, e.L, pa, pb, irho, e.max_order, e.num_int_steps, beti, exact);
      if (e.fbx != 0) r = edge_fringe(r, irho, e.e2, e.fint2, e.gap, e.fbx, true);
      return exit_(e, r);
    }
    case ORC_EXACTSBEND:      // :901-973
    case ORC_EXACTSBEND_LE: {
      bool le = (e.kind == ORC_EXACTSBEND_LE);
      r = entrance(e, r);
      T irho = e.angle / e.L;
      if (!le) {
        r = yrot(r, e.e1, beti);
        if (e.fbe != 0) r = bend_fringe(r, irho, e.gk, beti);
        r = bend_edge(r, irho, -e.e1, beti);
      } else {
        if (e.fbe != 0) r = edge_fringe(r, irho, e.e1, e.fint1, e.gap, e.fbe, false);
      }
      if (e.num_int_steps == 0) {
        r = exact_bend_body(r, irho, e.L, beti);
      } else {
        T SL = e.L / T(double(e.num_int_steps));
        T L1 = SL * T(DRIFT1);
        T L2 = SL * T(DRIFT2);
        T K1 = SL * T(KICK1);
        T K2 = SL * T(KICK2);
        magnet_polys(e, pa, pb, true);
        for (int s = 0; s < e.num_int_steps; ++s) {
          r = exact_bend_body(r, irho, L1, beti);
          r = strthinkick(r, pa, pb, K1, e.max_order);
          r = exact_bend_body(r, irho, L2, beti);
          r = strthinkick(r, pa, pb, K2, e.max_order);
          r = exact_bend_body(r, irho, L2, beti);
          r = strthinkick(r, pa, pb, K1, e.max_order);
          r = exact_bend_body(r, irho, L1, beti);
        }
      }
      if (!le) {
        r = bend_edge(r, irho, -e.e2, beti);
        if (e.fbx != 0) r = bend_fringe(r, -irho, e.gk, beti);
        r = yrot(r, e.e2, beti);
      } else {
        if (e.fbx != 0) r = edge_fringe(r, irho, e.e2, e.fint2, e.gap, e.fbx, true);
      }
      return exit_(e, r);
    }
    case ORC_RFCAVITY: {  // :502-518
      T beta = T(1.0) / beti;
      if (e.L > T(0.0)) r = drift6(r, e.L / T(2.0), beti, exact);
      if (e.energy > T(0.0)) {
        T nv = e.volt / e.energy;
        T phase = T(TWO_PI) * e.freq * ((r.v[4] - e.lag) / T(C_LIGHT)) - e.philag;
        r.v[5] = r.v[5] - nv * sin(phase) / (beta * beta);
      }
      if (e.L > T(0.0)) r = drift6(r, e.L / T(2.0), beti, exact);
      return r;
    }
    default:
      return pass_extra(e, r, beti, exact);
  }
}

template <class T>
static inline bool lost_after(const R6<T>& r, double ap_r2) {
  if (check_lost(r)) return true;
  if (ap_r2 > 0.0) {
    double x = val(r.v[0]), y = val(r.v[2]);
    if (x * x + y * y > ap_r2) return true;
  }
  return false;
}

template <class T>
static std::vector<Elem<T>> load_all(const orc_elem* elems, int n_elem) {
  std::vector<Elem<T>> v(n_elem);
  for (int i = 0; i < n_elem; ++i) load_elem(elems[i], v[i]);
  return v;
}

// linepass! over particles [i0, i1): src/lattice.jl:260-297 (element-outer, particle-inner).
// T = double restates the reference; T = long double runs the same algorithm with a
// 64-bit significand (the accuracy yardstick for implementations that round differently).
template <class T>
static void linepass_range(const std::vector<Elem<T>>& el, double* coords, int64_t n, int64_t i0, int64_t i1,
                           int64_t* lost_flags, T beti, double ap_r2, bool exact, int32_t* lost_turn,
                           int32_t* lost_elem, int32_t turn, T* wide) {
  const int ne = (int)el.size();
  for (int ie = 0; ie < ne; ++ie) {
    const Elem<T>& e = el[ie];
    for (int64_t i = i0; i < i1; ++i) {
      if (lost_flags[i] == 1) continue;
      R6<T> r;
      for (int k = 0; k < 6; ++k) r.v[k] = wide ? wide[i + k * n] : T(coords[i + k * n]);
      R6<T> rn = pass_elem(e, r, beti, exact);
      if (lost_after(rn, ap_r2)) {
        lost_flags[i] = 1;
        if (lost_turn) lost_turn[i] = turn;
        if (lost_elem) lost_elem[i] = ie;
        continue;
      }
      for (int k = 0; k < 6; ++k) {
        if (wide) wide[i + k * n] = rn.v[k];
        coords[i + k * n] = (double)rn.v[k];
      }
    }
  }
}

template <class T>
static void ringpass_t(const orc_elem* elems, int n_elem, double* coords, int64_t n, int64_t* lost_flags,
                       T beti, double ap_r2, int exact, int32_t* lost_turn, int32_t* lost_elem, int32_t turn0,
                       int32_t nturns, int nthreads) {
  std::vector<Elem<T>> el = load_all<T>(elems, n_elem);
  std::vector<T> wide;
  T* wp = nullptr;
  if (sizeof(T) > sizeof(double)) {
    wide.resize((size_t)(6 * n));
    for (int64_t i = 0; i < 6 * n; ++i) wide[(size_t)i] = T(coords[i]);
    wp = wide.data();
  }
  if (nthreads < 1) nthreads = 1;
#ifdef _OPENMP
#pragma omp parallel num_threads(nthreads)
  {
    int t = omp_get_thread_num(), nt = omp_get_num_threads();
    int64_t i0 = n * t / nt, i1 = n * (t + 1) / nt;
    for (int32_t k = 0; k < nturns; ++k)
      linepass_range<T>(el, coords, n, i0, i1, lost_flags, beti, ap_r2, exact != 0, lost_turn, lost_elem, turn0 + k, wp);
  }
#else
  for (int32_t k = 0; k < nturns; ++k)
    linepass_range<T>(el, coords, n, 0, n, lost_flags, beti, ap_r2, exact != 0, lost_turn, lost_elem, turn0 + k, wp);
#endif
}

extern "C" {

int orc_version(void) { return 1; }

// Beam(energy; mass): src/lattice.jl:50-54 and beti :61
double orc_beti(double energy, double mass) {
  double gamma = (energy + mass) / mass;
  double beta = std::sqrt(1.0 - 1.0 / (gamma * gamma));
  return 1.0 / beta;
}

// One particle, one element (for fixtures and unit checks).
int orc_pass(const orc_elem* e, double* r6, double beti, int exact) {
  Elem<double> el;
  load_elem(*e, el);
  R6<double> r;
  for (int k = 0; k < 6; ++k) r.v[k] = r6[k];
  r = pass_elem(el, r, beti, exact != 0);
  for (int k = 0; k < 6; ++k) r6[k] = r.v[k];
  return 0;
}

// linepass!: src/lattice.jl:260-297.  coords is N x 6 column-major (6 planes of N).
int orc_linepass(const orc_elem* elems, int n_elem, double* coords, int64_t n, int64_t* lost_flags,
                 double beti, double ap_r2, int exact, int32_t* lost_turn, int32_t* lost_elem, int32_t turn) {
  ringpass_t<double>(elems, n_elem, coords, n, lost_flags, beti, ap_r2, exact, lost_turn, lost_elem, turn, 1, 1);
  return 0;
}

// ringpass!: src/lattice.jl:305-317.  nthreads > 1 splits the particles into contiguous
// chunks, each tracked with the reference loop order (results are identical: particles
// do not interact).
int orc_ringpass(const orc_elem* elems, int n_elem, double* coords, int64_t n, int64_t* lost_flags,
                 double beti, double ap_r2, int exact, int32_t* lost_turn, int32_t* lost_elem,
                 int32_t turn0, int32_t nturns, int nthreads) {
  ringpass_t<double>(elems, n_elem, coords, n, lost_flags, beti, ap_r2, exact, lost_turn, lost_elem, turn0, nturns,
                     nthreads);
  return 0;
}

// The same ringpass! in quad precision; results rounded to double.  loss_out (may be NULL)
// receives sum over survivors of (x^2 + y^2), accumulated in quad precision.
int orc_ringpass_q(const orc_elem* elems, int n_elem, double* coords, int64_t n, int64_t* lost_flags,
                   double beti, double ap_r2, int exact, int32_t* lost_turn, int32_t* lost_elem,
                   int32_t turn0, int32_t nturns, int nthreads, double* loss_hi, double* loss_lo) {
  std::vector<Elem<f128>> el = load_all<f128>(elems, n_elem);
  std::vector<f128> wide((size_t)(6 * n));
  for (int64_t i = 0; i < 6 * n; ++i) wide[(size_t)i] = f128(coords[i]);
  if (nthreads < 1) nthreads = 1;
#ifdef _OPENMP
#pragma omp parallel num_threads(nthreads)
  {
    int t = omp_get_thread_num(), nt = omp_get_num_threads();
    int64_t i0 = n * t / nt, i1 = n * (t + 1) / nt;
    for (int32_t k = 0; k < nturns; ++k)
      linepass_range<f128>(el, coords, n, i0, i1, lost_flags, (f128)beti, ap_r2, exact != 0, lost_turn, lost_elem,
                           turn0 + k, wide.data());
  }
#else
  for (int32_t k = 0; k < nturns; ++k)
    linepass_range<f128>(el, coords, n, 0, n, lost_flags, (f128)beti, ap_r2, exact != 0, lost_turn, lost_elem,
                         turn0 + k, wide.data());
#endif
  f128 acc = 0;
  for (int64_t i = 0; i < n; ++i)
    if (lost_flags[i] != 1) acc += wide[(size_t)i] * wide[(size_t)i] + wide[(size_t)(2 * n + i)] * wide[(size_t)(2 * n + i)];
  // the quad sum split into a double pair hi + lo
  const double hi = (double)acc;
  if (loss_hi) *loss_hi = hi;
  if (loss_lo) *loss_lo = (double)(acc - (f128)hi);
  return 0;
}

// The same ringpass! in long double (x87 extended precision), results rounded to double.
int orc_ringpass_ld(const orc_elem* elems, int n_elem, double* coords, int64_t n, int64_t* lost_flags,
                    double beti, double ap_r2, int exact, int32_t* lost_turn, int32_t* lost_elem,
                    int32_t turn0, int32_t nturns, int nthreads) {
  ringpass_t<long double>(elems, n_elem, coords, n, lost_flags, (long double)beti, ap_r2, exact, lost_turn,
                          lost_elem, turn0, nturns, nthreads);
  return 0;
}

int orc_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

// Forward-mode Jacobian of nturns ring passes about r0 (one particle): J is 6x6 row-major,
// J[i*6+j] = d r_out[i] / d r0[j].  Returns 1 if the particle was lost.
int orc_jacobian(const orc_elem* elems, int n_elem, const double* r0, int32_t nturns, double beti, int exact,
                 double* r_out, double* J) {
  typedef Dual<6> D;
  std::vector<Elem<D>> el(n_elem);
  for (int i = 0; i < n_elem; ++i) load_elem(elems[i], el[i]);
  R6<D> r;
  for (int k = 0; k < 6; ++k) { r.v[k] = D(r0[k]); r.v[k].d[k] = 1.0; }
  D bt(beti);
  for (int t = 0; t < nturns; ++t)
    for (int ie = 0; ie < n_elem; ++ie) {
      R6<D> rn = pass_elem(el[ie], r, bt, exact != 0);
      if (check_lost(rn)) return 1;
      r = rn;
    }
  for (int i = 0; i < 6; ++i) {
    r_out[i] = r.v[i].v;
    for (int j = 0; j < 6; ++j) J[i * 6 + j] = r.v[i].d[j];
  }
  return 0;
}

// Directional derivative, for checking reverse-mode gradients:
//   loss = sum over surviving particles of (x^2 + y^2), after nturns,
//   d loss / d eps along  k_e -> k_e + eps * dk[e]  (e over elements: Quad k1, Sext k2,
//   Oct k3) and  r0 -> r0 + eps * dr0.  Particles lost in the primal are excluded.
//   Returns the loss in *loss and the directional derivative in *dloss.
int orc_loss_dir(const orc_elem* elems, int n_elem, const double* coords0, int64_t n, const double* dk,
                 const double* dr0, int32_t nturns, double beti, double ap_r2, int exact, double* loss,
                 double* dloss, int64_t* n_alive) {
  typedef Dual<1> D;
  std::vector<Elem<D>> el(n_elem);
  for (int i = 0; i < n_elem; ++i) {
    load_elem(elems[i], el[i]);
    if (dk) el[i].k.d[0] = dk[i];
  }
  D bt(beti);
  // particles in parallel, each into its own slot; the sums run in particle order, so the
  // result does not depend on the thread count
  std::vector<double> ls((size_t)(n > 0 ? n : 1), 0.0), dls((size_t)(n > 0 ? n : 1), 0.0);
  std::vector<char> ok((size_t)(n > 0 ? n : 1), 0);
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t p = 0; p < n; ++p) {
    R6<D> r;
    for (int k = 0; k < 6; ++k) { r.v[k] = D(coords0[p + k * n]); if (dr0) r.v[k].d[0] = dr0[p + k * n]; }
    bool lost = false;
    for (int t = 0; t < nturns && !lost; ++t)
      for (int ie = 0; ie < n_elem; ++ie) {
        R6<D> rn = pass_elem(el[ie], r, bt, exact != 0);
        if (lost_after(rn, ap_r2)) { lost = true; break; }
        r = rn;
      }
    if (lost) continue;
    D s = r.v[0] * r.v[0] + r.v[2] * r.v[2];
    ls[(size_t)p] = s.v;
    dls[(size_t)p] = s.d[0];
    ok[(size_t)p] = 1;
  }
  double L = 0.0, dL = 0.0;
  int64_t alive = 0;
  for (int64_t p = 0; p < n; ++p) {
    if (!ok[(size_t)p]) continue;
    ++alive;
    L += ls[(size_t)p];
    dL += dls[(size_t)p];
  }
  *loss = L;
  *dloss = dL;
  if (n_alive) *n_alive = alive;
  return 0;
}

}  // extern "C"
