// tp_track_kernels.hip -- fused multi-turn tracking kernel for gfx950.
//
// One lane tracks PPL particles (6 FP64 coordinates each, in VGPRs) through every
// element of every turn of a launch: the turn loop of ringpass! (src/lattice.jl:311-315),
// the element loop of linepass! (:266) and the particle loop (:268) collapse into one
// kernel.  The element table is read at wave-uniform addresses (scalar loads), so
// element dispatch is a scalar branch and every lane of a wave runs the same map.
//
// Loss semantics (src/lattice.jl:269-284): after every element a particle is tested
// with check_lost (src/tracking.jl:49-52) and the table's circular aperture; a lost
// particle keeps its element-entry coordinates and is never tracked again.  Here the
// entry state is written to HBM at the moment of loss together with the turn and
// lattice position; the lane then keeps executing (masked results are discarded), so
// the hot loop has no divergent control flow.
//
// Compiled twice: TP_FAST=0 (-ffp-contract=off, parity build) and TP_FAST=1.
#include "tp_extra.h"

#if TP_FAST
#define TP_LAUNCH tp_launch_track_fast
#define TP_TRACK_OCC tp_track_occupancy_fast
#define TP_BATCH_LAUNCH tp_launch_batch_fast
#define TP_BATCH_OCC tp_batch_occupancy_fast
#define TPK_BUILD fast
#else
#define TP_LAUNCH tp_launch_track_exact
#define TP_TRACK_OCC tp_track_occupancy_exact
#define TP_BATCH_LAUNCH tp_launch_batch_exact
#define TP_BATCH_OCC tp_batch_occupancy_exact
#define TPK_BUILD exact
#endif

namespace tpk {
namespace TPK_BUILD {  // kernel names carry the build flavour (profiles tell them apart)

// Thick multipole: Yoshida-4 with strthinkick, src/tracking.jl:317-341 / :386-491.
// MM > 0: general kick of order MM with its coefficients held in registers (kick_cf).
template <bool EXH, int KK, int PPL, bool SER, int MM = 0>
TP_DEV void mpole_steps(St* s, const Cx* c, const double* __restrict__ p, int nsteps, int m) {
  const double L1 = p[0], L2 = p[1], K1 = p[2], K2 = p[3], L1b = p[4], L2b = p[5];
  const double* poly = p + 6;
#if TP_FAST
  if constexpr (EXH) {
    // Path length summed per Yoshida step (drift_xy).  Consecutive steps meet at two
    // half-drifts of length L1 with no kick between them (... K1 D1 | D1 K1 ...): they see the
    // same momenta, so they are applied as one (drift_xy_pair), and the pair's NormL is
    // split between the two steps' path sums.
    const double Lsb = 2.0 * L1b + 2.0 * L2b;
    if (nsteps <= 0) return;
    double S_[PPL];
#define TP_YBODY(k, KICK1, KICK2)                         \
  KICK1;                                                  \
  S_[k] = S_[k] + drift_xy<SER>(s[k], c[k], L2);          \
  KICK2;                                                  \
  S_[k] = S_[k] + drift_xy<SER>(s[k], c[k], L2);          \
  KICK1;
#define TP_YSTEPS(KICK1, KICK2)                                                    \
  {                                                                                \
    _Pragma("unroll") for (int k = 0; k < PPL; ++k) S_[k] = drift_xy<SER>(s[k], c[k], L1); \
    for (int st = 1; st < nsteps; ++st) {                                          \
      _Pragma("unroll") for (int k = 0; k < PPL; ++k) {                            \
        TP_YBODY(k, KICK1, KICK2)                                                  \
        const double N_ = drift_xy_pair<SER>(s[k], c[k], L1);                     \
        s[k].z = s[k].z + __builtin_fma(S_[k] + N_, c[k].B, -Lsb);                 \
        S_[k] = N_;                                                                \
      }                                                                            \
    }                                                                              \
    _Pragma("unroll") for (int k = 0; k < PPL; ++k) {                              \
      TP_YBODY(k, KICK1, KICK2)                                                    \
      S_[k] = S_[k] + drift_xy<SER>(s[k], c[k], L1);                               \
      s[k].z = s[k].z + __builtin_fma(S_[k], c[k].B, -Lsb);                        \
    }                                                                              \
  }
    if constexpr (MM > 0) {
      double cf[2 * MM + 2];
#pragma unroll
      for (int i = 0; i < 2 * MM + 2; ++i) cf[i] = poly[i];
      TP_YSTEPS(kick_cf<MM>(s[k], cf, K1), kick_cf<MM>(s[k], cf, K2))
    } else if constexpr (KK == TP_KICK_SEXT) {
      const double Kb1 = K1 * poly[4], Kb2 = K2 * poly[4];
      TP_YSTEPS(kick_sext_kb(s[k], Kb1), kick_sext_kb(s[k], Kb2))
    } else {
      TP_YSTEPS(kick<KK>(s[k], poly, m, K1), kick<KK>(s[k], poly, m, K2))
    }
#undef TP_YSTEPS
#undef TP_YBODY
    return;
  }
#endif
  if constexpr (MM > 0) {
    double cf[2 * MM + 2];
#pragma unroll
    for (int i = 0; i < 2 * MM + 2; ++i) cf[i] = poly[i];
    for (int st = 0; st < nsteps; ++st) {
#pragma unroll
      for (int k = 0; k < PPL; ++k) {
        drift<EXH, SER>(s[k], c[k], L1, L1b);
        kick_cf<MM>(s[k], cf, K1);
        drift<EXH, SER>(s[k], c[k], L2, L2b);
        kick_cf<MM>(s[k], cf, K2);
        drift<EXH, SER>(s[k], c[k], L2, L2b);
        kick_cf<MM>(s[k], cf, K1);
        drift<EXH, SER>(s[k], c[k], L1, L1b);
      }
    }
    return;
  }
  for (int st = 0; st < nsteps; ++st) {
#pragma unroll
    for (int k = 0; k < PPL; ++k) {
      drift<EXH, SER>(s[k], c[k], L1, L1b);
      kick<KK>(s[k], poly, m, K1);
      drift<EXH, SER>(s[k], c[k], L2, L2b);
      kick<KK>(s[k], poly, m, K2);
      drift<EXH, SER>(s[k], c[k], L2, L2b);
      kick<KK>(s[k], poly, m, K1);
      drift<EXH, SER>(s[k], c[k], L1, L1b);
    }
  }
}

// Expanded-Hamiltonian bend body: Yoshida-4 with bndthinkick, src/tracking.jl:350-375
template <bool EXH, int KK, int PPL, bool SER>
TP_DEV void sbend_steps(St* s, const Cx* c, const double* __restrict__ p, int nsteps, int m, double beti) {
  const double L1 = p[0], L2 = p[1], K1 = p[2], K2 = p[3], L1b = p[4], L2b = p[5];
  const double irho = p[6], K1i = p[7], K2i = p[8];
  const double* poly = p + 17;
  for (int st = 0; st < nsteps; ++st) {
#pragma unroll
    for (int k = 0; k < PPL; ++k) {
      drift<EXH, SER>(s[k], c[k], L1, L1b);
      bkick<KK>(s[k], poly, m, K1, K1i, irho, beti);
      drift<EXH, SER>(s[k], c[k], L2, L2b);
      bkick<KK>(s[k], poly, m, K2, K2i, irho, beti);
      drift<EXH, SER>(s[k], c[k], L2, L2b);
      bkick<KK>(s[k], poly, m, K1, K1i, irho, beti);
      drift<EXH, SER>(s[k], c[k], L1, L1b);
    }
  }
}

// Exact bend body with Yoshida-4 sub-kicks, src/tracking.jl:923-953
template <int KK, int PPL, int MM = 0>
TP_DEV void esbend_body(St* s, const Cx* c, const double* __restrict__ p, int nsteps, int m, bool driftlimit) {
  const double irho = p[0];
  if (driftlimit) {
    if (nsteps == 0) {
#pragma unroll
      for (int k = 0; k < PPL; ++k) ebody_driftlimit(s[k], c[k], p[1], p[5]);
      return;
    }
    const double L1 = p[1], L2 = p[2], K1 = p[3], K2 = p[4], L1b = p[5], L2b = p[6];
    const double* poly = p + 24;
    for (int st = 0; st < nsteps; ++st) {
#pragma unroll
      for (int k = 0; k < PPL; ++k) {
        ebody_driftlimit(s[k], c[k], L1, L1b);
        kick<KK>(s[k], poly, m, K1);
        ebody_driftlimit(s[k], c[k], L2, L2b);
        kick<KK>(s[k], poly, m, K2);
        ebody_driftlimit(s[k], c[k], L2, L2b);
        kick<KK>(s[k], poly, m, K1);
        ebody_driftlimit(s[k], c[k], L1, L1b);
      }
    }
    return;
  }
#if TP_FAST
  const double rho = 1.0 / irho;
#else
  const double rho = 0.0;
#endif
  // py is constant through the body and through identity kicks: hoist d2 = pxyz(dp1,0,py).
  // pz = pxyz(dp1, px, py) is carried from one body to the next (ebody) and evaluated again
  // only after a kick that moves px or py.
  double d2[PPL], id2[PPL], pz[PPL];
#pragma unroll
  for (int k = 0; k < PPL; ++k) {
    d2[k] = pxyz(c[k], 0.0, s[k].py);
    pz[k] = pxyz(c[k], s[k].px, s[k].py);
#if TP_FAST
    id2[k] = rcp_nr(d2[k]);
#else
    id2[k] = 0.0;
#endif
  }
  if (nsteps == 0) {
#pragma unroll
    for (int k = 0; k < PPL; ++k) ebody(s[k], c[k], irho, rho, p[1], p[5], p[7], p[8], d2[k], id2[k], pz[k]);
    return;
  }
  const double L1 = p[1], L2 = p[2], K1 = p[3], K2 = p[4], L1b = p[5], L2b = p[6];
  const double c1 = p[7], s1 = p[8], c2 = p[9], s2 = p[10];
  const double* poly = p + 24;
#if TP_FAST
#define TP_AFTER_KICK(k)                        \
  d2[k] = pxyz(c[k], 0.0, s[k].py);             \
  id2[k] = rcp_nr(d2[k]);                       \
  pz[k] = pxyz(c[k], s[k].px, s[k].py)
#else
#define TP_AFTER_KICK(k)                        \
  d2[k] = pxyz(c[k], 0.0, s[k].py);             \
  pz[k] = pxyz(c[k], s[k].px, s[k].py)
#endif
#if TP_FAST
  if constexpr (KK == TP_KICK_NONE) {
    // Pure dipole (identity kicks): py, delta and hence d2 stay fixed, and each sub-body of
    // exact_bend_body (src/tracking.jl:858-894) is a rotation of (px, pzmx), pzmx = pz -
    // (1 + x irho), by the sub-step's angle: px' = px cs + pzmx sn, and from the body's own
    // x' = (pzn - pzmx cs + px sn - 1)/irho, pzmx' = pzn - (1 + x' irho) = pzmx cs - px sn.
    // With identity kicks between them the 4 nsteps sub-body rotations compose to one, by
    // the body's whole angle (rot_body, formed once per element: 4 instructions per lane
    // instead of a square root, a series arcsine and the O(1) cancellation of x' per
    // sub-body); x is formed once, at the end, from pz - 1 = (dp1^2 - 1 - px^2 - py^2) /
    // (pz + 1) without cancellation, and the path terms asin(px_k/d2) - asin(px_k+1/d2) of
    // the sub-bodies telescope to one difference.
    const double Ltot = (double)nsteps * (2.0 * L1 + 2.0 * L2);
    const double Lbtot = (double)nsteps * (2.0 * L1b + 2.0 * L2b);
    const Rot R = rot_body(c1, s1, c2, s2, nsteps);
#pragma unroll
    for (int k = 0; k < PPL; ++k) {
      const double Bm1 = c[k].B - 1.0;                 // exact: dp1 - 1 of the reference's dp1
      const double BBm1 = Bm1 * (c[k].B + 1.0);        // dp1^2 - 1
      const double dd = __builtin_fma(-s[k].py, s[k].py, BBm1);
      const double px0 = s[k].px, pz0 = pz[k];
      double a = px0;
      double b = __builtin_fma(-s[k].x, irho, __builtin_fma(-px0, px0, dd) * rcp_nr(pz0 + 1.0));
      // The rotation keeps a^2 + b^2; below d2^2 every sub-body has |px| < d2, so none of the
      // reference's clamps (pz at 0, arcsine arguments at +-1) acts and the rotations are its
      // map.  A particle past that (a grazing angle) takes the reference's sub-bodies one by one.
      if (__builtin_expect(!(__builtin_fma(a, a, b * b) < 0.999999 * (d2[k] * d2[k])), 0)) {
        const double Lsb = 2.0 * L1b + 2.0 * L2b;
        for (int st = 0; st < nsteps; ++st) {
          double D = ebody<false>(s[k], c[k], irho, rho, L1, L1b, c1, s1, d2[k], id2[k], pz[k]);
          D = D + ebody<false>(s[k], c[k], irho, rho, L2, L2b, c2, s2, d2[k], id2[k], pz[k]);
          D = D + ebody<false>(s[k], c[k], irho, rho, L2, L2b, c2, s2, d2[k], id2[k], pz[k]);
          D = D + ebody<false>(s[k], c[k], irho, rho, L1, L1b, c1, s1, d2[k], id2[k], pz[k]);
          s[k].y = s[k].y + s[k].py * D;
          s[k].z = s[k].z + __builtin_fma(c[k].B, D, -Lsb);
        }
        continue;
      }
      {
        const double a_ = __builtin_fma(a, R.c, b * R.s);
        b = __builtin_fma(b, R.c, -(a * R.s));
        a = a_;
      }
      const double pzn = pxyz(c[k], a, s[k].py);
      const double pn1 = __builtin_fma(-a, a, dd) * rcp_nr(pzn + 1.0);
      const double S = __builtin_fma(asin_diff(px0, pz0, a, pzn, d2[k], id2[k]), rho, Ltot);
      s[k].x = (pn1 - b) * rho;
      s[k].px = a;
      s[k].y = __builtin_fma(s[k].py, S, s[k].y);
      s[k].z = s[k].z + __builtin_fma(c[k].B, S, -Lbtot);
      pz[k] = pzn;
    }
    return;
  } else {
    // Consecutive steps meet at two sub-bodies of length L1 with no kick between them.
    // Forest's body is the exact flow of its length, so the pair is one body of length 2 L1
    // (sub-step angle doubled: cos = 1 - 2 s1^2, sin = 2 c1 s1) -- one square root and one
    // arcsine difference instead of two.  The cosine is formed from s1 alone: c1^2 - s1^2
    // carries c1's rounding twice, and the body's x, which cancels O(1) terms, is sensitive
    // to the pair's norm (that form gave 2-3x the x error of a 10-turn DA scan).
    const double c11 = __builtin_fma(-2.0 * s1, s1, 1.0), s11 = 2.0 * (c1 * s1);
    const double L11 = L1 + L1, L11b = L1b + L1b;
#define TP_EBODY_STEP(k, KICK1, KICK2, LE, LEB, CE, SE)                            \
  KICK1;                                                                            \
  TP_AFTER_KICK(k);                                                                 \
  ebody(s[k], c[k], irho, rho, L2, L2b, c2, s2, d2[k], id2[k], pz[k]);              \
  KICK2;                                                                            \
  TP_AFTER_KICK(k);                                                                 \
  ebody(s[k], c[k], irho, rho, L2, L2b, c2, s2, d2[k], id2[k], pz[k]);              \
  KICK1;                                                                            \
  TP_AFTER_KICK(k);                                                                 \
  ebody(s[k], c[k], irho, rho, LE, LEB, CE, SE, d2[k], id2[k], pz[k]);
#define TP_EBODY_STEPS(KICK1, KICK2)                                                \
  {                                                                                 \
    _Pragma("unroll") for (int k = 0; k < PPL; ++k)                                 \
        ebody(s[k], c[k], irho, rho, L1, L1b, c1, s1, d2[k], id2[k], pz[k]);        \
    for (int st = 1; st < nsteps; ++st) {                                           \
      _Pragma("unroll") for (int k = 0; k < PPL; ++k) {                             \
        TP_EBODY_STEP(k, KICK1, KICK2, L11, L11b, c11, s11)                         \
      }                                                                             \
    }                                                                               \
    _Pragma("unroll") for (int k = 0; k < PPL; ++k) {                               \
      TP_EBODY_STEP(k, KICK1, KICK2, L1, L1b, c1, s1)                               \
    }                                                                               \
  }
    if constexpr (MM > 0) {
      double cf[2 * MM + 2];
#pragma unroll
      for (int i = 0; i < 2 * MM + 2; ++i) cf[i] = poly[i];
      TP_EBODY_STEPS(kick_cf<MM>(s[k], cf, K1), kick_cf<MM>(s[k], cf, K2))
    } else {
      TP_EBODY_STEPS(kick<KK>(s[k], poly, m, K1), kick<KK>(s[k], poly, m, K2))
    }
#undef TP_EBODY_STEPS
#undef TP_EBODY_STEP
    return;
  }
#endif
  if constexpr (MM > 0) {
    double cf[2 * MM + 2];
#pragma unroll
    for (int i = 0; i < 2 * MM + 2; ++i) cf[i] = poly[i];
    for (int st = 0; st < nsteps; ++st) {
#pragma unroll
      for (int k = 0; k < PPL; ++k) {
        ebody(s[k], c[k], irho, rho, L1, L1b, c1, s1, d2[k], id2[k], pz[k]);
        kick_cf<MM>(s[k], cf, K1);
        TP_AFTER_KICK(k);
        ebody(s[k], c[k], irho, rho, L2, L2b, c2, s2, d2[k], id2[k], pz[k]);
        kick_cf<MM>(s[k], cf, K2);
        TP_AFTER_KICK(k);
        ebody(s[k], c[k], irho, rho, L2, L2b, c2, s2, d2[k], id2[k], pz[k]);
        kick_cf<MM>(s[k], cf, K1);
        TP_AFTER_KICK(k);
        ebody(s[k], c[k], irho, rho, L1, L1b, c1, s1, d2[k], id2[k], pz[k]);
      }
    }
    return;
  }
  for (int st = 0; st < nsteps; ++st) {
#pragma unroll
    for (int k = 0; k < PPL; ++k) {
      ebody(s[k], c[k], irho, rho, L1, L1b, c1, s1, d2[k], id2[k], pz[k]);
      kick<KK>(s[k], poly, m, K1);
      if (KK != TP_KICK_NONE) {
        TP_AFTER_KICK(k);
      }
      ebody(s[k], c[k], irho, rho, L2, L2b, c2, s2, d2[k], id2[k], pz[k]);
      kick<KK>(s[k], poly, m, K2);
      if (KK != TP_KICK_NONE) {
        TP_AFTER_KICK(k);
      }
      ebody(s[k], c[k], irho, rho, L2, L2b, c2, s2, d2[k], id2[k], pz[k]);
      kick<KK>(s[k], poly, m, K1);
      if (KK != TP_KICK_NONE) {
        TP_AFTER_KICK(k);
      }
      ebody(s[k], c[k], irho, rho, L1, L1b, c1, s1, d2[k], id2[k], pz[k]);
    }
  }
#undef TP_AFTER_KICK
}

// Misalignment, src/tracking.jl:60-67 and the r1/r2 products of every magnet pass!.
TP_DEV void add6(St& s, const double* __restrict__ t) {
  s.x = s.x + t[0];
  s.px = s.px + t[1];
  s.y = s.y + t[2];
  s.py = s.py + t[3];
  s.z = s.z + t[4];
  s.d = s.d + t[5];
}
TP_DEV void mat6(St& s, const double* __restrict__ M) {
  const double r[6] = {s.x, s.px, s.y, s.py, s.z, s.d};
  double o[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    double acc = M[i] * r[0];
#pragma unroll
    for (int j = 1; j < 6; ++j) acc = acc + M[i + 6 * j] * r[j];
    o[i] = acc;
  }
  s.x = o[0];
  s.px = o[1];
  s.y = o[2];
  s.py = o[3];
  s.z = o[4];
  s.d = o[5];
}

// Kick dispatch on the record's specialisation; TP_KICK_GEN only in kernels built with
// FT_GENKICK (the host never sends a record a kernel variant does not contain).
// General kicks of order 3..10 run with their coefficients in registers (MM = m); other
// orders take the loop over the table.
#define TP_GEN_ORDERS(CALLM)                                                              \
  switch (m) {                                                                           \
    case 3: CALLM(3); return;                                                            \
    case 4: CALLM(4); return;                                                            \
    case 5: CALLM(5); return;                                                            \
    case 6: CALLM(6); return;                                                            \
    case 7: CALLM(7); return;                                                            \
    case 8: CALLM(8); return;                                                            \
    case 9: CALLM(9); return;                                                            \
    case 10: CALLM(10); return;                                                          \
    default: break;                                                                      \
  }

template <bool EXH, int KK, int PPL, bool SER>
TP_DEV void mpole_dispatch(St* s, const Cx* c, const double* __restrict__ p, int nsteps, int m) {
  if constexpr (KK == TP_KICK_GEN) {
#define TP_CALLM(MM) mpole_steps<EXH, KK, PPL, SER, MM>(s, c, p, nsteps, m)
    TP_GEN_ORDERS(TP_CALLM)
#undef TP_CALLM
  }
  mpole_steps<EXH, KK, PPL, SER>(s, c, p, nsteps, m);
}

template <int KK, int PPL>
TP_DEV void esbend_dispatch(St* s, const Cx* c, const double* __restrict__ p, int nsteps, int m, bool driftlimit) {
  if constexpr (KK == TP_KICK_GEN) {
    if (!driftlimit) {
#define TP_CALLM(MM) esbend_body<KK, PPL, MM>(s, c, p, nsteps, m, false)
      TP_GEN_ORDERS(TP_CALLM)
#undef TP_CALLM
    }
  }
  esbend_body<KK, PPL>(s, c, p, nsteps, m, driftlimit);
}

#define TP_KICK_SWITCH(FEAT, CALL)                                         \
  switch (rec.kick) {                                                      \
    case TP_KICK_NONE: CALL(TP_KICK_NONE); break;                          \
    case TP_KICK_QUAD: CALL(TP_KICK_QUAD); break;                          \
    case TP_KICK_SEXT: CALL(TP_KICK_SEXT); break;                          \
    default:                                                               \
      if constexpr ((FEAT) & FT_GENKICK) { CALL(TP_KICK_GEN); }            \
      break;                                                               \
  }

// SER_OK: the launch may use the series drift (inv_pz), which pays where several waves share
// a SIMD; it holds two more values per particle and two per element, so the variants with
// general kicks (up to 22 coefficients in registers) use the reciprocal square root always.
// DCH: the element may change delta (changes_delta); with DCH = false the code that renews
// the delta hoists (the RF kick, the other families, misalignments) is not compiled in.
template <bool EXH, unsigned FEAT, int PPL, bool SER_OK, bool DCH = true>
TP_DEV void apply_elem(St* s, Cx* c, EdgeMemo* mm, const tp_rec& rec, const double* __restrict__ pool, double beti) {
  constexpr bool SER = SER_OK && (FEAT & FT_GENKICK) == 0;
  const double* __restrict__ p = pool + rec.off;
  const uint32_t fl = rec.flags;
  if constexpr (DCH && (FEAT & FT_MISALIGN) != 0) {
    if (fl & TP_F_MISALIGN) {
      const double* mb = pool + rec.moff;
#pragma unroll
      for (int k = 0; k < PPL; ++k) {
        if (fl & TP_F_T1) add6(s[k], mb);
        if (fl & TP_F_R1) mat6(s[k], mb + 12);
        mkctx(s[k], c[k], beti);
      }
    }
  }
  const int m = rec.max_order;
  const int ns = rec.nsteps;
  switch (rec.tag) {
    case TP_DRIFT:
#pragma unroll
      for (int k = 0; k < PPL; ++k) drift<EXH, SER>(s[k], c[k], p[0], p[1]);
      break;
    case TP_MPOLE: {
#define TP_CALL(KK) mpole_dispatch<EXH, KK, PPL, SER>(s, c, p, ns, m)
      TP_KICK_SWITCH(FEAT, TP_CALL)
#undef TP_CALL
      break;
    }
    case TP_THINMP: {
#define TP_CALL(KK)                                               \
  _Pragma("unroll") for (int k = 0; k < PPL; ++k) kick<KK>(s[k], p, m, 1.0)
      TP_KICK_SWITCH(FEAT, TP_CALL)
#undef TP_CALL
      break;
    }
    case TP_SBEND:
      if constexpr ((FEAT & FT_SBEND) != 0) {
        const int fin = (fl >> TP_F_FIN_SHIFT) & 0xF, fout = (fl >> TP_F_FOUT_SHIFT) & 0xF;
        const double irho = p[6];
        if (fin != 0) {
#pragma unroll
          for (int k = 0; k < PPL; ++k) edge_lin<true>(s[k], fin, p + 9, irho, false, &mm[k]);
        }
#define TP_CALL(KK) sbend_steps<EXH, KK, PPL, SER>(s, c, p, ns, m, beti)
        TP_KICK_SWITCH(FEAT, TP_CALL)
#undef TP_CALL
        if (fout != 0) {
#pragma unroll
          for (int k = 0; k < PPL; ++k) edge_lin<true>(s[k], fout, p + 13, irho, true, &mm[k]);
        }
      }
      break;
    case TP_ESBEND:
      if constexpr ((FEAT & FT_ESBEND) != 0) {
        const bool dl = (fl & TP_F_DRIFTLIMIT) != 0;
        const int fin = (fl >> TP_F_FIN_SHIFT) & 0xF, fout = (fl >> TP_F_FOUT_SHIFT) & 0xF;
        const double irho = p[0];
        if (p[16] != 0.0) {  // yrot returns early for phi == 0 (:728-730); p[16] = -e1
#pragma unroll
          for (int k = 0; k < PPL; ++k) s[k] = yrot(s[k], c[k], p[11], p[12]);
        }
        if (fin != 0) {
#pragma unroll
          for (int k = 0; k < PPL; ++k) s[k] = bend_fringe(s[k], c[k], irho, p[23]);
        }
        if (!dl) {
#pragma unroll
          for (int k = 0; k < PPL; ++k) s[k] = bend_edge(s[k], c[k], irho, p[13], p[14], p[15], p[16]);
        }
#define TP_CALL(KK) esbend_dispatch<KK, PPL>(s, c, p, ns, m, dl)
        TP_KICK_SWITCH(FEAT, TP_CALL)
#undef TP_CALL
        if (!dl) {
#pragma unroll
          for (int k = 0; k < PPL; ++k) s[k] = bend_edge(s[k], c[k], irho, p[17], p[18], p[19], p[20]);
        }
        if (fout != 0) {
#pragma unroll
          for (int k = 0; k < PPL; ++k) s[k] = bend_fringe(s[k], c[k], -irho, p[23]);
        }
        if (p[20] != 0.0) {  // p[20] = -e2
#pragma unroll
          for (int k = 0; k < PPL; ++k) s[k] = yrot(s[k], c[k], p[21], p[22]);
        }
      }
      break;
    case TP_ESBEND_LE:
      if constexpr ((FEAT & FT_ESBEND_LE) != 0) {
        const bool dl = (fl & TP_F_DRIFTLIMIT) != 0;
        const int fin = (fl >> TP_F_FIN_SHIFT) & 0xF, fout = (fl >> TP_F_FOUT_SHIFT) & 0xF;
        const double irho = p[0];
        if (fin != 0) {
#pragma unroll
          for (int k = 0; k < PPL; ++k) edge_lin<true>(s[k], fin, p + 11, irho, false, &mm[k]);
        }
#define TP_CALL(KK) esbend_dispatch<KK, PPL>(s, c, p, ns, m, dl)
        TP_KICK_SWITCH(FEAT, TP_CALL)
#undef TP_CALL
        if (fout != 0) {
#pragma unroll
          for (int k = 0; k < PPL; ++k) edge_lin<true>(s[k], fout, p + 15, irho, true, &mm[k]);
        }
      }
      break;
    case TP_RFCAV:  // src/tracking.jl:502-518
      if constexpr ((FEAT & FT_RF) != 0) {
        const bool dr = (fl & TP_F_RF_DRIFT) != 0;
#pragma unroll
        for (int k = 0; k < PPL; ++k) {
          if (dr) drift<EXH, SER>(s[k], c[k], p[0], p[1]);
          if (DCH && (fl & TP_F_RF_KICK)) {
            const double phase = p[3] * ((s[k].z - p[4]) / 2.99792458e8) - p[5];
            s[k].d = s[k].d - p[2] * sin_lib(phase) / p[6];
            mkctx(s[k], c[k], beti);
          }
          if (dr) drift<EXH, SER>(s[k], c[k], p[0], p[1]);
        }
      }
      break;
    case TP_MARKER:
      break;
    default:  // the other element families, out of line (tp_extra.h)
      if constexpr (DCH && (FEAT & FT_EXTRA) != 0) {
#pragma unroll
        for (int k = 0; k < PPL; ++k) {
          s[k] = extra_map<EXH>(s[k], rec.tag, fl, ns, p, beti);
          mkctx(s[k], c[k], beti);
        }
      }
      break;
  }
  if constexpr (DCH && (FEAT & FT_MISALIGN) != 0) {
    if (fl & TP_F_MISALIGN) {
      const double* mb = pool + rec.moff;
#pragma unroll
      for (int k = 0; k < PPL; ++k) {
        if (fl & TP_F_R2) mat6(s[k], mb + 48);
        if (fl & TP_F_T2) add6(s[k], mb + 6);
        mkctx(s[k], c[k], beti);
      }
    }
  }
}

// An element that changes delta.  Without misalignments only the RF kick and the other
// families do, so only their cases are compiled here.
template <bool EXH, unsigned FEAT, int PPL, bool SER_OK>
TP_DEV void apply_elem_dch(St* s, Cx* c, EdgeMemo* mm, const tp_rec& rec, const double* __restrict__ pool,
                           double beti) {
  if constexpr ((FEAT & FT_MISALIGN) != 0) {
    apply_elem<EXH, FEAT, PPL, SER_OK, true>(s, c, mm, rec, pool, beti);
  } else {
    constexpr bool SER = SER_OK && (FEAT & FT_GENKICK) == 0;
    const double* __restrict__ p = pool + rec.off;
    const uint32_t fl = rec.flags;
    if ((FEAT & FT_RF) != 0 && rec.tag == TP_RFCAV) {  // src/tracking.jl:502-518
      const bool dr = (fl & TP_F_RF_DRIFT) != 0;
#pragma unroll
      for (int k = 0; k < PPL; ++k) {
        if (dr) drift<EXH, SER>(s[k], c[k], p[0], p[1]);
        const double phase = p[3] * ((s[k].z - p[4]) / 2.99792458e8) - p[5];
        s[k].d = s[k].d - p[2] * sin_lib(phase) / p[6];
        mkctx(s[k], c[k], beti);
        if (dr) drift<EXH, SER>(s[k], c[k], p[0], p[1]);
      }
    } else if constexpr ((FEAT & FT_EXTRA) != 0) {  // the other families (tp_extra.h)
#pragma unroll
      for (int k = 0; k < PPL; ++k) {
        s[k] = extra_map<EXH>(s[k], rec.tag, fl, rec.nsteps, p, beti);
        mkctx(s[k], c[k], beti);
      }
    }
  }
}

// check_lost (src/tracking.jl:49-52): isnan(x) || |x| > 1 || |y| > 1 || |px| > 1 || |py| > 1
// (limits from the table, validated positive), and the optional circular aperture (ap: the
// table has one; wave-uniform), as the wave's lane mask.  isnan(x) || |x| > CL is
// !(|x| <= CL); the other tests are ordered comparisons, false for NaN as in the reference.
// Each compare's lane mask is ORed on the scalar unit, so the test costs its four compares.
// The aperture sum is rounded explicitly so both builds take identical loss decisions.
TP_DEV uint64_t lost_mask(const St& s, const TpKArgs& a, bool ap) {
  uint64_t m = __builtin_amdgcn_ballot_w64(!(fabs(s.x) <= a.coord_limit));
  m |= __builtin_amdgcn_ballot_w64(fabs(s.y) > a.coord_limit);
  m |= __builtin_amdgcn_ballot_w64(fabs(s.px) > a.angle_limit);
  m |= __builtin_amdgcn_ballot_w64(fabs(s.py) > a.angle_limit);
  if (ap) {
    // an empty volatile statement keeps this a branch on the uniform flag: otherwise the
    // compiler computes the aperture test for every element of tables without one
    __asm__ volatile("");
    m |= __builtin_amdgcn_ballot_w64(__dadd_rn(__dmul_rn(s.x, s.x), __dmul_rn(s.y, s.y)) > a.ap_r2);
  }
  return m;
}

// Loads PPL particles per lane (idx / valid per particle); lanes past the beam's end and
// particles already lost start dead with a zero state.
template <int PPL>
TP_DEV void load_lanes(St* s, Cx* c, bool* alive, const double* __restrict__ coords,
                       const int32_t* __restrict__ lost_turn, int64_t ld, double beti, const int64_t* idx,
                       const bool* valid) {
#pragma unroll
  for (int k = 0; k < PPL; ++k) {
    alive[k] = false;
    s[k] = St{0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (valid[k]) {
      alive[k] = (lost_turn[idx[k]] == -1);
      if (alive[k]) {
        const double* q = coords + idx[k];
        s[k] = St{q[0], q[ld], q[2 * ld], q[3 * ld], q[4 * ld], q[5 * ld]};
      }
    }
    mkctx(s[k], c[k], beti);
  }
}

TP_DEV void store_state(double* q, int64_t ld, const St& s) {
  q[0] = s.x;
  q[ld] = s.px;
  q[2 * ld] = s.y;
  q[3 * ld] = s.py;
  q[4 * ld] = s.z;
  q[5 * ld] = s.d;
}

// Turns [t_begin, t_end) of the ring for PPL particles per lane.  A particle found lost
// after an element has its entry state, turn (a.turn0 + t) and position written out and
// is dead from then on; the wave leaves early once every particle in it is dead.
template <bool EXH, unsigned FEAT, int PPL, bool SER_OK>
TP_DEV void run_turns(St* s, Cx* c, EdgeMemo* mm, bool* alive, const int32_t* __restrict__ seq, const tp_rec* __restrict__ recs,
                      const double* __restrict__ pool, double* __restrict__ coords, int32_t* __restrict__ lost_turn,
                      int32_t* __restrict__ lost_elem, const TpKArgs& a, const int64_t* idx, int32_t t_begin,
                      int32_t t_end) {
  const bool ap = __builtin_amdgcn_readfirstlane(a.ap_r2 > 0.0 ? 1 : 0) != 0;
  // alive[k] as the wave's lane mask (scalar registers): the loss test of an element is then
  // its compares and scalar mask logic, with no per-lane flag to test
  uint64_t am[PPL];
#pragma unroll
  for (int k = 0; k < PPL; ++k) am[k] = __builtin_amdgcn_ballot_w64(alive[k]);
  for (int32_t t = t_begin; t < t_end; ++t) {
    uint64_t anyl = 0;
#pragma unroll
    for (int k = 0; k < PPL; ++k) anyl |= am[k];
    if (anyl == 0) break;
    // the lattice of this turn: row (turn - row_turn0) of a time-varying table
    const int32_t* __restrict__ seq_t = seq + (a.n_rows > 1 ? (int64_t)(a.turn0 + t - a.row_turn0) * a.n_seq : 0);
    // elements of this turn: a launch may start inside its first turn and end inside its last
    const int32_t eb = t == 0 ? a.e_first : 0;
    const int32_t ee = (t == a.nturns - 1 && a.e_last > 0) ? a.e_last : a.n_seq;
    // Software-pipelined scalar loads: the record of element e+1 and the sequence entry of
    // element e+2 are requested while element e runs, so each element waits for one
    // round trip of scalar loads (its parameters) instead of three.
    int32_t ri_next = eb + 1 < ee ? seq_t[eb + 1] : 0;
    tp_rec rec_next = eb < ee ? recs[seq_t[eb]] : tp_rec{};
    // After an element: a particle lost there has its entry state s0, the turn and the
    // position written out.
    auto settle = [&](const St* s0, int32_t e) {
      uint64_t lm[PPL];
      uint64_t anylost = 0;
#pragma unroll
      for (int k = 0; k < PPL; ++k) {
        lm[k] = lost_mask(s[k], a, ap) & am[k];
        anylost |= lm[k];
      }
      if (anylost != 0) {
        const int lane = lane_id_here();  // not held across the element loop
#pragma unroll
        for (int k = 0; k < PPL; ++k) {
          am[k] &= ~lm[k];
          if ((lm[k] >> lane) & 1) {
            store_state(coords + idx[k], a.ld, s0[k]);
            lost_turn[idx[k]] = a.turn0 + t;
            lost_elem[idx[k]] = e;
            // the dead lane goes on from the zero state: no NaN left in it to send its wave
            // through the slow paths that only NaN or large angles take.  Its delta hoists
            // (c) stay those of the particle's last delta, finite unless its delta was,
            // until the next element that changes delta renews them.  A dead lane's values
            // never reach an output.
            s[k] = St{0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
          }
        }
      }
    };
    int32_t e = eb;
    while (e < ee) {
      // Elements that leave delta unchanged: the delta hoists (c) are invariant in this loop,
      // so they keep their registers; an element that renews them (changes_delta) ends the
      // loop and runs below.
      for (; e < ee; ++e) {
        const tp_rec rec = rec_next;
        if (changes_delta<FEAT>(rec)) break;
        if (e + 1 < ee) {
          rec_next = recs[ri_next];
          if (e + 2 < ee) ri_next = seq_t[e + 2];
        }
        St s0[PPL];
#pragma unroll
        for (int k = 0; k < PPL; ++k) s0[k] = s[k];
        apply_elem<EXH, FEAT, PPL, SER_OK, false>(s, c, mm, rec, pool, a.beti);
        settle(s0, e);
      }
      if (e >= ee) break;
      const tp_rec rec = rec_next;
      if (e + 1 < ee) {
        rec_next = recs[ri_next];
        if (e + 2 < ee) ri_next = seq_t[e + 2];
      }
      St s0[PPL];
#pragma unroll
      for (int k = 0; k < PPL; ++k) s0[k] = s[k];
      apply_elem_dch<EXH, FEAT, PPL, SER_OK>(s, c, mm, rec, pool, a.beti);
      settle(s0, e);
      ++e;
    }
  }
  const int lane = lane_id();
#pragma unroll
  for (int k = 0; k < PPL; ++k) alive[k] = ((am[k] >> lane) & 1) != 0;
}

// Per-wave LDS of the tracking kernels (dynamic, sized by the launch: TP_LDS_PER_WAVE(ppl)
// bytes per wave of the block): for each particle per lane k, the edge memo's 128 doubles
// (64 deltas, 64 values).  Each lane uses only its own slots, so no barrier is needed.
TP_DEV double* wave_lds(double* lds, int ppl) {
  const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x) >> 6;
  return lds + (size_t)wave * ppl * (TP_LDS_PER_WAVE(1) / 8);
}

template <unsigned FEAT, int PPL>
TP_DEV void edge_memo_init(EdgeMemo* mm, double* wl) {
#pragma unroll
  for (int k = 0; k < PPL; ++k) {
    mm[k].ke = 0;
    mm[k].kf = 0;
    mm[k].w = wl + 128 * k;
    if constexpr ((FEAT & (FT_SBEND | FT_ESBEND_LE)) != 0) mm[k].w[lane_id()] = __builtin_nan("");
  }
}

// WPE > 0 asks the register allocator for at least WPE resident waves per SIMD (spilling
// if it must); WPE = 0 leaves the budget to the compiler.
#define TP_WAVES(WPE) __attribute__((amdgpu_waves_per_eu((WPE) > 0 ? (WPE) : 1)))

// SER: the series drift (the caller decides once per call: TpKArgs::ser).
template <bool EXH, unsigned FEAT, int PPL, int WPE, bool SER>
__global__ __launch_bounds__(256) TP_WAVES(WPE) void track_kernel(const int32_t* __restrict__ seq, const tp_rec* __restrict__ recs,
                                                    const double* __restrict__ pool, double* __restrict__ coords,
                                                    int32_t* __restrict__ lost_turn, int32_t* __restrict__ lost_elem,
                                                    TpKArgs a) {
  // Classic launch: the wave's lanes are list entries base + k blockDim.  Persistent launch
  // (a.counter): the wave takes tiles of 64 PPL consecutive entries from the queue until
  // it is empty; waves then start their tiles at different times and drift apart, so the
  // scalar parameter loads of different waves do not all wait at the same element.
  const bool persist = a.counter != nullptr;
  const int lane = threadIdx.x & 63;
  const int64_t n_tiles = (a.n + 64 * PPL - 1) / (64 * PPL);
  extern __shared__ double tp_lds[];
  double* wl = wave_lds(tp_lds, PPL);
  EdgeMemo mm[PPL];
  edge_memo_init<FEAT, PPL>(mm, wl);
  for (;;) {
    int64_t i0, stride;
    if (persist) {
      int32_t it = 0;
      if (lane == 0) it = atomicAdd(a.counter, 1);
      it = __builtin_amdgcn_readfirstlane(it);
      if (it >= n_tiles) break;
      i0 = (int64_t)it * 64 * PPL + lane;
      stride = 64;
    } else {
      i0 = (int64_t)blockIdx.x * blockDim.x * PPL + threadIdx.x;
      stride = blockDim.x;
    }
    int64_t idx[PPL];
    bool valid[PPL];
#pragma unroll
    for (int k = 0; k < PPL; ++k) {
      const int64_t i = i0 + (int64_t)k * stride;
      valid[k] = i < a.n;
      idx[k] = (a.order && valid[k]) ? (int64_t)a.order[i] : i;
    }
    St s[PPL];
    Cx c[PPL];
    bool alive[PPL];
    load_lanes<PPL>(s, c, alive, coords, lost_turn, a.ld, a.beti, idx, valid);
    run_turns<EXH, FEAT, PPL, SER>(s, c, mm, alive, seq, recs, pool, coords, lost_turn, lost_elem, a, idx, 0, a.nturns);
#pragma unroll
    for (int k = 0; k < PPL; ++k)
      if (alive[k]) store_state(coords + idx[k], a.ld, s[k]);
    if (!persist) break;
  }
}

// Batched persistent kernel: every wave takes work items -- (lattice variant, tile of 64
// particles), sorted by decreasing cost -- from an atomic queue until it is empty.  The
// item index is made wave-uniform with readfirstlane, so the variant's descriptor and
// its element table (a slice of the concatenated tables) are read with scalar loads, as
// in the one-beam kernel.  With CKPT the wave also stores its tile's state every
// ckpt_stride turns (and at the last turn) into the variant's checkpoint slab, and the
// tile's loss sum and survivor count, for the batched adjoint.
template <bool EXH, unsigned FEAT, bool CKPT, int WPE>
__global__ __launch_bounds__(256) TP_WAVES(WPE) void batch_kernel(const TpVarDesc* __restrict__ desc, const int2* __restrict__ items,
                                                    int32_t n_items, int32_t* __restrict__ counter,
                                                    const int32_t* __restrict__ seq, const tp_rec* __restrict__ recs,
                                                    const double* __restrict__ pool, double* __restrict__ ckpt,
                                                    double* __restrict__ item_loss) {
  const int lane = threadIdx.x & 63;
  extern __shared__ double tp_lds[];
  double* wl = wave_lds(tp_lds, 1);
  EdgeMemo mm[1];
  edge_memo_init<FEAT, 1>(mm, wl);
  for (;;) {
    int32_t it = 0;
    if (lane == 0) it = atomicAdd(counter, 1);
    it = __builtin_amdgcn_readfirstlane(it);
    if (it >= n_items) break;
    const int2 item = items[it];
    const TpVarDesc& d = desc[item.x];
    TpKArgs a{};
    a.n_seq = d.n_seq;
    a.beti = d.beti;
    a.coord_limit = d.coord_limit;
    a.angle_limit = d.angle_limit;
    a.ap_r2 = d.ap_r2;
    a.ld = d.n;
    a.turn0 = d.turn0;
    a.nturns = d.nturns;
    a.n_rows = d.n_rows;
    a.row_turn0 = d.row_turn0;
    const int32_t* vseq = seq + d.seq_off;
    const tp_rec* vrecs = recs + d.rec_off;
    const double* vpool = pool + d.pool_off;
    const int64_t li = (int64_t)item.y * 64 + lane;
    int64_t idx[1] = {li};
    bool valid[1] = {li < d.n};
    St s[1];
    Cx c[1];
    bool alive[1];
    load_lanes<1>(s, c, alive, d.coords, d.lost_turn, d.n, d.beti, idx, valid);
    if constexpr (CKPT) {
      double* ck = ckpt + d.ck_off + (int64_t)item.y * d.nseg * 384 + lane;
      for (int32_t j = 0; j < d.nseg; ++j) {
        const int32_t tb = j * d.ckpt_stride;
        const int32_t te = (j == d.nseg - 1) ? d.nturns : tb + d.ckpt_stride;
        run_turns<EXH, FEAT, 1, true>(s, c, mm, alive, vseq, vrecs, vpool, d.coords, d.lost_turn, d.lost_elem, a, idx, tb, te);
        store_state(ck + (int64_t)j * 384, 64, s[0]);
      }
      const double l2 = alive[0] ? s[0].x * s[0].x + s[0].y * s[0].y : 0.0;
      const double lsum = wave_sum(l2);
      const double cnt = (double)__popcll(__ballot(alive[0]));
      if (lane == 0) {
        item_loss[2 * (int64_t)it] = lsum;
        item_loss[2 * (int64_t)it + 1] = cnt;
      }
    } else {
      run_turns<EXH, FEAT, 1, true>(s, c, mm, alive, vseq, vrecs, vpool, d.coords, d.lost_turn, d.lost_elem, a, idx, 0,
                              d.nturns);
    }
    if (alive[0]) store_state(d.coords + li, d.n, s[0]);
  }
}

// Clears the work-queue counters inside a captured graph (no memset nodes in graphs).
__global__ void fill_i32(int32_t* __restrict__ p, int32_t v, int32_t n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Survivor list for the next launch: each wave ballots its entries whose particle is alive,
// lane 0 reserves that many slots with one atomicAdd, and every kept entry is written at
// its rank within the wave's ballot.  Lanes map to particles one to one and no particle's
// arithmetic depends on its lane, so tracking with the list gives the same bits.
__global__ void compact_alive(const int32_t* __restrict__ order_in, int32_t n_in,
                              const int32_t* __restrict__ lost_turn, int32_t* __restrict__ order_out,
                              int32_t* __restrict__ count) {
  const int32_t i = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
  const int lane = threadIdx.x & 63;
  int32_t id = 0;
  bool keep = false;
  if (i < n_in) {
    id = order_in ? order_in[i] : i;
    keep = lost_turn[id] == -1;
  }
  const unsigned long long m = __ballot(keep);
  int32_t at = 0;
  if (lane == 0 && m != 0ull) at = atomicAdd(count, (int32_t)__popcll(m));
  at = __shfl(at, 0, 64);
  if (keep) order_out[at + (int32_t)__popcll(m & ((1ull << lane) - 1ull))] = id;
}

}  // namespace TPK_BUILD
}  // namespace tpk

#define TP_KPTRS(a) a.seq, a.recs, a.pool, a.coords, a.lost_turn, a.lost_elem

// Register budgets compiled per kernel: the default, and at least 5 or 6 waves per SIMD
// for the one-particle-per-lane kernels of variants B and C (the others keep the default).
template <bool EXH, unsigned FEAT>
constexpr bool has_budgets() {
  return EXH && (FEAT == TP_FEAT_B || FEAT == TP_FEAT_C);
}

#define TP_WPE_SWITCH(waves, CALL)                    \
  if constexpr (has_budgets<EXH, FEAT>()) {           \
    if ((waves) == 5) { CALL(5); }                    \
    else if ((waves) == 6) { CALL(6); }               \
    else { CALL(0); }                                 \
  } else {                                            \
    CALL(0);                                          \
  }

template <bool EXH, unsigned FEAT, bool SER>
static void launch_vs(const TpKArgs& a, int ppl, unsigned grid, int block, hipStream_t stream) {
  if (ppl == 2) {
    hipLaunchKernelGGL((tpk::TPK_BUILD::track_kernel<EXH, FEAT, 2, 0, SER>), dim3(grid), dim3(block),
                       TP_LDS_PER_WAVE(2) * (block / 64), stream,
                       TP_KPTRS(a), a);
  } else {
#define TP_CALL(W)                                                                                               \
  hipLaunchKernelGGL((tpk::TPK_BUILD::track_kernel<EXH, FEAT, 1, W, SER>), dim3(grid), dim3(block),      \
                     TP_LDS_PER_WAVE(1) * (block / 64), stream,                                            \
                     TP_KPTRS(a), a)
    TP_WPE_SWITCH(a.waves, TP_CALL)
#undef TP_CALL
  }
}

// The series drift exists in the fast build of the variants without general kicks.
template <bool EXH, unsigned FEAT>
static void launch_v(const TpKArgs& a, int ppl, unsigned grid, int block, hipStream_t stream) {
  if constexpr (TP_FAST && (FEAT & FT_GENKICK) == 0) {
    if (a.ser) {
      launch_vs<EXH, FEAT, true>(a, ppl, grid, block, stream);
      return;
    }
  }
  launch_vs<EXH, FEAT, false>(a, ppl, grid, block, stream);
}

template <bool EXH, unsigned FEAT, bool SER>
static const void* track_fn_s(int ppl, int waves) {
  const void* fn = nullptr;
  if (ppl == 2) {
    fn = reinterpret_cast<const void*>(tpk::TPK_BUILD::track_kernel<EXH, FEAT, 2, 0, SER>);
  } else {
#define TP_CALL(W) fn = reinterpret_cast<const void*>(tpk::TPK_BUILD::track_kernel<EXH, FEAT, 1, W, SER>)
    TP_WPE_SWITCH(waves, TP_CALL)
#undef TP_CALL
  }
  return fn;
}

// The kernel a launch with these arguments runs (launch_v's choice), for occupancy queries.
template <bool EXH, unsigned FEAT>
static const void* track_fn(int ppl, int waves, int ser) {
  if constexpr (TP_FAST && (FEAT & FT_GENKICK) == 0) {
    if (ser) return track_fn_s<EXH, FEAT, true>(ppl, waves);
  }
  return track_fn_s<EXH, FEAT, false>(ppl, waves);
}

template <bool EXH, unsigned FEAT, bool CKPT>
static const void* batch_fn(int waves) {
  const void* fn = nullptr;
#define TP_CALL(W) fn = reinterpret_cast<const void*>(tpk::TPK_BUILD::batch_kernel<EXH, FEAT, CKPT, W>)
  TP_WPE_SWITCH(waves, TP_CALL)
#undef TP_CALL
  return fn;
}

template <bool EXH, unsigned FEAT, bool CKPT>
static void launch_batch_v(const TpBatchArgs& b, hipStream_t stream) {
#define TP_CALL(W)                                                                                               \
  hipLaunchKernelGGL((tpk::TPK_BUILD::batch_kernel<EXH, FEAT, CKPT, W>), dim3(b.grid), dim3(b.block),      \
                     TP_LDS_PER_WAVE(1) * (b.block / 64), stream,                                             \
                     b.desc, b.items, b.n_items, b.counter, b.seq, b.recs, b.pool, b.ckpt, b.item_loss)
  TP_WPE_SWITCH(b.waves, TP_CALL)
#undef TP_CALL
}

#define TP_BATCH_DISPATCH(CALL)                               \
  switch (variant) {                                          \
    case TP_VARIANT_A: CALL(true, TP_FEAT_A); break;          \
    case TP_VARIANT_B: CALL(true, TP_FEAT_B); break;          \
    case TP_VARIANT_C: CALL(true, TP_FEAT_C); break;          \
    default:                                                  \
      if (exact_h) { CALL(true, TP_FEAT_ALL); }               \
      else { CALL(false, TP_FEAT_ALL); }                      \
      break;                                                  \
  }

hipError_t TP_BATCH_LAUNCH(const TpBatchArgs& b, bool ckpt, hipStream_t stream) {
  if (b.n_items <= 0) return hipSuccess;
  const int variant = b.variant, exact_h = b.exact_h;
  if (ckpt) {
#define TP_CALL(E, F) launch_batch_v<E, F, true>(b, stream)
    TP_BATCH_DISPATCH(TP_CALL)
#undef TP_CALL
  } else {
#define TP_CALL(E, F) launch_batch_v<E, F, false>(b, stream)
    TP_BATCH_DISPATCH(TP_CALL)
#undef TP_CALL
  }
  return hipGetLastError();
}

// Resident blocks per CU of the batched kernel, for sizing the persistent grid.
int TP_BATCH_OCC(int variant, int exact_h, bool ckpt, int block, int waves) {
  const void* fn = nullptr;
  if (ckpt) {
#define TP_CALL(E, F) fn = batch_fn<E, F, true>(waves)
    TP_BATCH_DISPATCH(TP_CALL)
#undef TP_CALL
  } else {
#define TP_CALL(E, F) fn = batch_fn<E, F, false>(waves)
    TP_BATCH_DISPATCH(TP_CALL)
#undef TP_CALL
  }
  int nb = 0;
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, block, TP_LDS_PER_WAVE(1) * (block / 64)) == hipSuccess
             ? nb : 0;
}

int TP_TRACK_OCC(int variant, int exact_h, int block, int ppl, int waves, int ser) {
  const void* fn = nullptr;
#define TP_CALL(E, F) fn = track_fn<E, F>(ppl, waves, ser)
  TP_BATCH_DISPATCH(TP_CALL)
#undef TP_CALL
  int nb = 0;
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, block, TP_LDS_PER_WAVE(ppl) * (block / 64)) ==
                 hipSuccess ? nb : 0;
}

hipError_t TP_LAUNCH(const TpKArgs& a_in, int block, int ppl, int max_grid, hipStream_t stream) {
  if (a_in.n <= 0) return hipSuccess;
  TpKArgs a = a_in;
  const int64_t per_block = (int64_t)block * ppl;
  int64_t grid64 = (a.n + per_block - 1) / per_block;
  if (a.counter && max_grid > 0 && grid64 > max_grid) grid64 = max_grid;  // persistent grid
  if (grid64 > 0x7fffffff) return hipErrorInvalidValue;
  const unsigned grid = (unsigned)grid64;
  switch (a.variant) {
    case TP_VARIANT_A: launch_v<true, TP_FEAT_A>(a, ppl, grid, block, stream); break;
    case TP_VARIANT_B: launch_v<true, TP_FEAT_B>(a, ppl, grid, block, stream); break;
    case TP_VARIANT_C: launch_v<true, TP_FEAT_C>(a, ppl, grid, block, stream); break;
    default:
      if (a.exact_h) launch_v<true, TP_FEAT_ALL>(a, ppl, grid, block, stream);
      else launch_v<false, TP_FEAT_ALL>(a, ppl, grid, block, stream);
      break;
  }
  return hipGetLastError();
}

#if TP_FAST
hipError_t tp_launch_fill_i32(int32_t* p, int32_t v, int32_t n, hipStream_t stream) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(tpk::TPK_BUILD::fill_i32, dim3((n + 255) / 256), dim3(256), 0, stream, p, v, n);
  return hipGetLastError();
}

hipError_t tp_launch_compact_alive(const int32_t* order_in, int32_t n_in, const int32_t* lost_turn, int32_t* order_out,
                                   int32_t* count, hipStream_t stream) {
  if (n_in <= 0) return hipSuccess;
  hipLaunchKernelGGL(tpk::TPK_BUILD::compact_alive, dim3((unsigned)((n_in + 255) / 256)), dim3(256), 0, stream,
                     order_in, n_in, lost_turn, order_out, count);
  return hipGetLastError();
}
#endif
