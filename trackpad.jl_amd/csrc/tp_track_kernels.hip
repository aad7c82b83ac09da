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
#include "tp_maps.h"

#if TP_FAST
#define TP_LAUNCH tp_launch_track_fast
#define TPK_BUILD fast
#else
#define TP_LAUNCH tp_launch_track_exact
#define TPK_BUILD exact
#endif

namespace tpk {
namespace TPK_BUILD {  // kernel names carry the build flavour (profiles tell them apart)

// Thick multipole: Yoshida-4 with strthinkick, src/tracking.jl:317-341 / :386-491
template <bool EXH, int KK, int PPL>
TP_DEV void mpole_steps(St* s, const Cx* c, const double* __restrict__ p, int nsteps, int m) {
  const double L1 = p[0], L2 = p[1], K1 = p[2], K2 = p[3], L1b = p[4], L2b = p[5];
  const double* poly = p + 6;
  for (int st = 0; st < nsteps; ++st) {
#pragma unroll
    for (int k = 0; k < PPL; ++k) {
      drift<EXH>(s[k], c[k], L1, L1b);
      kick<KK>(s[k], poly, m, K1);
      drift<EXH>(s[k], c[k], L2, L2b);
      kick<KK>(s[k], poly, m, K2);
      drift<EXH>(s[k], c[k], L2, L2b);
      kick<KK>(s[k], poly, m, K1);
      drift<EXH>(s[k], c[k], L1, L1b);
    }
  }
}

// Expanded-Hamiltonian bend body: Yoshida-4 with bndthinkick, src/tracking.jl:350-375
template <bool EXH, int KK, int PPL>
TP_DEV void sbend_steps(St* s, const Cx* c, const double* __restrict__ p, int nsteps, int m, double beti) {
  const double L1 = p[0], L2 = p[1], K1 = p[2], K2 = p[3], L1b = p[4], L2b = p[5];
  const double irho = p[6], K1i = p[7], K2i = p[8];
  const double* poly = p + 17;
  for (int st = 0; st < nsteps; ++st) {
#pragma unroll
    for (int k = 0; k < PPL; ++k) {
      drift<EXH>(s[k], c[k], L1, L1b);
      bkick<KK>(s[k], poly, m, K1, K1i, irho, beti);
      drift<EXH>(s[k], c[k], L2, L2b);
      bkick<KK>(s[k], poly, m, K2, K2i, irho, beti);
      drift<EXH>(s[k], c[k], L2, L2b);
      bkick<KK>(s[k], poly, m, K1, K1i, irho, beti);
      drift<EXH>(s[k], c[k], L1, L1b);
    }
  }
}

// Exact bend body with Yoshida-4 sub-kicks, src/tracking.jl:923-953
template <int KK, int PPL>
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
  double d2[PPL], id2[PPL];
#pragma unroll
  for (int k = 0; k < PPL; ++k) {
    d2[k] = pxyz(c[k], 0.0, s[k].py);
#if TP_FAST
    id2[k] = rcp_nr(d2[k]);
#else
    id2[k] = 0.0;
#endif
  }
  if (nsteps == 0) {
#pragma unroll
    for (int k = 0; k < PPL; ++k) ebody(s[k], c[k], irho, rho, p[1], p[5], p[7], p[8], d2[k], id2[k]);
    return;
  }
  const double L1 = p[1], L2 = p[2], K1 = p[3], K2 = p[4], L1b = p[5], L2b = p[6];
  const double c1 = p[7], s1 = p[8], c2 = p[9], s2 = p[10];
  const double* poly = p + 24;
  for (int st = 0; st < nsteps; ++st) {
#pragma unroll
    for (int k = 0; k < PPL; ++k) {
      ebody(s[k], c[k], irho, rho, L1, L1b, c1, s1, d2[k], id2[k]);
      kick<KK>(s[k], poly, m, K1);
      if (KK != TP_KICK_NONE) {
        d2[k] = pxyz(c[k], 0.0, s[k].py);
#if TP_FAST
        id2[k] = rcp_nr(d2[k]);
#endif
      }
      ebody(s[k], c[k], irho, rho, L2, L2b, c2, s2, d2[k], id2[k]);
      kick<KK>(s[k], poly, m, K2);
      if (KK != TP_KICK_NONE) {
        d2[k] = pxyz(c[k], 0.0, s[k].py);
#if TP_FAST
        id2[k] = rcp_nr(d2[k]);
#endif
      }
      ebody(s[k], c[k], irho, rho, L2, L2b, c2, s2, d2[k], id2[k]);
      kick<KK>(s[k], poly, m, K1);
      if (KK != TP_KICK_NONE) {
        d2[k] = pxyz(c[k], 0.0, s[k].py);
#if TP_FAST
        id2[k] = rcp_nr(d2[k]);
#endif
      }
      ebody(s[k], c[k], irho, rho, L1, L1b, c1, s1, d2[k], id2[k]);
    }
  }
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
#define TP_KICK_SWITCH(FEAT, CALL)                                         \
  switch (rec.kick) {                                                      \
    case TP_KICK_NONE: CALL(TP_KICK_NONE); break;                          \
    case TP_KICK_QUAD: CALL(TP_KICK_QUAD); break;                          \
    case TP_KICK_SEXT: CALL(TP_KICK_SEXT); break;                          \
    default:                                                               \
      if constexpr ((FEAT) & FT_GENKICK) { CALL(TP_KICK_GEN); }            \
      break;                                                               \
  }

template <bool EXH, unsigned FEAT, int PPL>
TP_DEV void apply_elem(St* s, Cx* c, const tp_rec& rec, const double* __restrict__ pool, double beti) {
  const double* __restrict__ p = pool + rec.off;
  const uint32_t fl = rec.flags;
  if constexpr ((FEAT & FT_MISALIGN) != 0) {
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
      for (int k = 0; k < PPL; ++k) drift<EXH>(s[k], c[k], p[0], p[1]);
      break;
    case TP_MPOLE: {
#define TP_CALL(KK) mpole_steps<EXH, KK, PPL>(s, c, p, ns, m)
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
          for (int k = 0; k < PPL; ++k) edge_lin(s[k], fin, p + 9, irho, false);
        }
#define TP_CALL(KK) sbend_steps<EXH, KK, PPL>(s, c, p, ns, m, beti)
        TP_KICK_SWITCH(FEAT, TP_CALL)
#undef TP_CALL
        if (fout != 0) {
#pragma unroll
          for (int k = 0; k < PPL; ++k) edge_lin(s[k], fout, p + 13, irho, true);
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
#define TP_CALL(KK) esbend_body<KK, PPL>(s, c, p, ns, m, dl)
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
          for (int k = 0; k < PPL; ++k) edge_lin(s[k], fin, p + 11, irho, false);
        }
#define TP_CALL(KK) esbend_body<KK, PPL>(s, c, p, ns, m, dl)
        TP_KICK_SWITCH(FEAT, TP_CALL)
#undef TP_CALL
        if (fout != 0) {
#pragma unroll
          for (int k = 0; k < PPL; ++k) edge_lin(s[k], fout, p + 15, irho, true);
        }
      }
      break;
    case TP_RFCAV:  // src/tracking.jl:502-518
      if constexpr ((FEAT & FT_RF) != 0) {
        const bool dr = (fl & TP_F_RF_DRIFT) != 0;
#pragma unroll
        for (int k = 0; k < PPL; ++k) {
          if (dr) drift<EXH>(s[k], c[k], p[0], p[1]);
          if (fl & TP_F_RF_KICK) {
            const double phase = p[3] * ((s[k].z - p[4]) / 2.99792458e8) - p[5];
            s[k].d = s[k].d - p[2] * sin_lib(phase) / p[6];
            mkctx(s[k], c[k], beti);
          }
          if (dr) drift<EXH>(s[k], c[k], p[0], p[1]);
        }
      }
      break;
    default:  // TP_MARKER
      break;
  }
  if constexpr ((FEAT & FT_MISALIGN) != 0) {
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

// check_lost (src/tracking.jl:49-52) plus the circular aperture; the aperture sum is
// rounded explicitly so both builds take identical loss decisions.
TP_DEV bool is_lost(const St& s, const TpKArgs& a) {
  bool l = isnan(s.x) || fabs(s.x) > a.coord_limit || fabs(s.y) > a.coord_limit ||
           fabs(s.px) > a.angle_limit || fabs(s.py) > a.angle_limit;
  if (a.ap_r2 > 0.0) l = l || (__dadd_rn(__dmul_rn(s.x, s.x), __dmul_rn(s.y, s.y)) > a.ap_r2);
  return l;
}

template <bool EXH, unsigned FEAT, int PPL>
__global__ __launch_bounds__(256) void track_kernel(const int32_t* __restrict__ seq, const tp_rec* __restrict__ recs,
                                                    const double* __restrict__ pool, double* __restrict__ coords,
                                                    int32_t* __restrict__ lost_turn, int32_t* __restrict__ lost_elem,
                                                    TpKArgs a) {
  const int64_t base = (int64_t)blockIdx.x * blockDim.x * PPL + threadIdx.x;
  St s[PPL];
  Cx c[PPL];
  bool alive[PPL];
  int64_t idx[PPL];
#pragma unroll
  for (int k = 0; k < PPL; ++k) {
    idx[k] = base + (int64_t)k * blockDim.x;
    alive[k] = false;
    s[k] = St{0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (idx[k] < a.n) {
      alive[k] = (lost_turn[idx[k]] == -1);
      if (alive[k]) {
        const double* q = coords + idx[k];
        s[k] = St{q[0], q[a.ld], q[2 * a.ld], q[3 * a.ld], q[4 * a.ld], q[5 * a.ld]};
      }
    }
    mkctx(s[k], c[k], a.beti);
  }
  for (int32_t t = 0; t < a.nturns; ++t) {
    bool anyl = false;
#pragma unroll
    for (int k = 0; k < PPL; ++k) anyl = anyl || alive[k];
    if (!__any(anyl)) break;
    for (int32_t e = 0; e < a.n_seq; ++e) {
      const tp_rec rec = recs[seq[e]];
      St s0[PPL];
#pragma unroll
      for (int k = 0; k < PPL; ++k) s0[k] = s[k];
      apply_elem<EXH, FEAT, PPL>(s, c, rec, pool, a.beti);
      bool lostnow[PPL];
      bool anylost = false;
#pragma unroll
      for (int k = 0; k < PPL; ++k) {
        lostnow[k] = alive[k] && is_lost(s[k], a);
        anylost = anylost || lostnow[k];
      }
      if (__any(anylost)) {
#pragma unroll
        for (int k = 0; k < PPL; ++k) {
          if (lostnow[k]) {
            double* q = coords + idx[k];
            q[0] = s0[k].x;
            q[a.ld] = s0[k].px;
            q[2 * a.ld] = s0[k].y;
            q[3 * a.ld] = s0[k].py;
            q[4 * a.ld] = s0[k].z;
            q[5 * a.ld] = s0[k].d;
            lost_turn[idx[k]] = a.turn0 + t;
            lost_elem[idx[k]] = e;
            alive[k] = false;
          }
        }
      }
    }
  }
#pragma unroll
  for (int k = 0; k < PPL; ++k) {
    if (alive[k]) {
      double* q = coords + idx[k];
      q[0] = s[k].x;
      q[a.ld] = s[k].px;
      q[2 * a.ld] = s[k].y;
      q[3 * a.ld] = s[k].py;
      q[4 * a.ld] = s[k].z;
      q[5 * a.ld] = s[k].d;
    }
  }
}

}  // namespace TPK_BUILD
}  // namespace tpk

#define TP_KPTRS(a) a.seq, a.recs, a.pool, a.coords, a.lost_turn, a.lost_elem

template <bool EXH, unsigned FEAT>
static void launch_v(const TpKArgs& a, int ppl, unsigned grid, int block, hipStream_t stream) {
  if (ppl == 2)
    hipLaunchKernelGGL((tpk::TPK_BUILD::track_kernel<EXH, FEAT, 2>), dim3(grid), dim3(block), 0, stream, TP_KPTRS(a), a);
  else
    hipLaunchKernelGGL((tpk::TPK_BUILD::track_kernel<EXH, FEAT, 1>), dim3(grid), dim3(block), 0, stream, TP_KPTRS(a), a);
}

hipError_t TP_LAUNCH(const TpKArgs& a, int block, int ppl, hipStream_t stream) {
  if (a.n <= 0) return hipSuccess;
  const int64_t per_block = (int64_t)block * ppl;
  const int64_t grid64 = (a.n + per_block - 1) / per_block;
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
