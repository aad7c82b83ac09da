// tp_adjoint_kernels.hip -- reverse-mode adjoint of multi-turn tracking for gfx950.
//
// Loss: L = (1/N_alive) sum over particles alive after the last turn of (x^2 + y^2).
// The forward pass (the tracking kernel, run segment by segment) stores the beam at
// segment boundaries.  adjoint_kernel processes one segment backwards: each lane starts
// from the stored state at the segment's end, undoes every sub-step with its inverse map
// and applies the sub-step's VJP to its adjoint (tp_adjoint.h), so no intermediate state
// is stored.  dL/d(strength) of every lattice position with a gradient slot is reduced
// across the wave with DPP shuffles, accumulated per wave in LDS and added once per launch
// to the block's row of a [blocks x slots] partial array; reduce_kernel sums the rows in a
// fixed order, so gradients are bit-reproducible.
//
// Compiled twice like the forward kernels (TP_FAST = 0 / 1).
#include "tp_adjoint.h"

#if TP_FAST
#define TP_ADJ_LAUNCH tp_launch_adjoint_fast
#define TP_BATCH_BACK_LAUNCH tp_launch_batch_back_fast
#define TP_BATCH_BACK_OCC tp_batch_back_occupancy_fast
#define TPK_BUILD adj_fast
#else
#define TP_ADJ_LAUNCH tp_launch_adjoint_exact
#define TP_BATCH_BACK_LAUNCH tp_launch_batch_back_exact
#define TP_BATCH_BACK_OCC tp_batch_back_occupancy_exact
#define TPK_BUILD adj_exact
#endif

namespace tpk {
namespace TPK_BUILD {

#define TP_KICK_SWITCH(FEAT, CALL)                                         \
  switch (rec.kick) {                                                      \
    case TP_KICK_NONE: CALL(TP_KICK_NONE); break;                          \
    case TP_KICK_QUAD: CALL(TP_KICK_QUAD); break;                          \
    case TP_KICK_SEXT: CALL(TP_KICK_SEXT); break;                          \
    default:                                                               \
      if constexpr ((FEAT) & FT_GENKICK) { CALL(TP_KICK_GEN); }            \
      break;                                                               \
  }

template <int KK, bool SER>
TP_DEV void mpole_back(St& s, Ad& a, const Cx& c, const double* __restrict__ p, int ns, int m, int g, double& gacc,
                       bool lin) {
#if TP_FAST
  if constexpr (KK != TP_KICK_GEN) {
    // the forward's step forms (mpole_steps) undone as such; a strength gradient of another
    // order than the kick's own coefficient (a magnet whose main strength is zero) takes the
    // general path below
    constexpr int gown = KK == TP_KICK_QUAD ? 1 : (KK == TP_KICK_SEXT ? 2 : -1);
    if (!lin && (g < 0 || g == gown)) {
      mpole_back_fast<KK, SER>(s, a, c, p, ns, g >= 0, gacc);
      return;
    }
  }
#endif
  const double L1 = p[0], L2 = p[1], K1 = p[2], K2 = p[3], L1b = p[4], L2b = p[5];
  const double* poly = p + 6;
  // the Yoshida step D1 K1 D2 K2 D2 K1 D1 is a palindrome: undo it in the same order
  for (int st = 0; st < ns; ++st) {
    drift_back<SER>(s, a, c, L1, L1b, lin);
    kick_back<KK>(s, a, poly, m, K1, g, gacc);
    drift_back<SER>(s, a, c, L2, L2b, lin);
    kick_back<KK>(s, a, poly, m, K2, g, gacc);
    drift_back<SER>(s, a, c, L2, L2b, lin);
    kick_back<KK>(s, a, poly, m, K1, g, gacc);
    drift_back<SER>(s, a, c, L1, L1b, lin);
  }
}

template <int KK, bool SER>
TP_DEV void sbend_back(St& s, Ad& a, const Cx& c, const double* __restrict__ p, int ns, int m, double beti, bool lin) {
  const double L1 = p[0], L2 = p[1], K1 = p[2], K2 = p[3], L1b = p[4], L2b = p[5];
  const double irho = p[6], K1i = p[7], K2i = p[8];
  const double* poly = p + 17;
  for (int st = 0; st < ns; ++st) {
    drift_back<SER>(s, a, c, L1, L1b, lin);
    bkick_back<KK>(s, a, poly, m, K1, K1i, irho, beti);
    drift_back<SER>(s, a, c, L2, L2b, lin);
    bkick_back<KK>(s, a, poly, m, K2, K2i, irho, beti);
    drift_back<SER>(s, a, c, L2, L2b, lin);
    bkick_back<KK>(s, a, poly, m, K1, K1i, irho, beti);
    drift_back<SER>(s, a, c, L1, L1b, lin);
  }
}

template <int KK>
TP_DEV void esbend_body_back(St& s, Ad& a, const Cx& c, const double* __restrict__ p, int ns, int m) {
  const double irho = p[0];
  double gdummy = 0.0;
  const double rho = 1.0 / irho;
#if TP_FAST
#define TP_ID2(d) rcp_nr(d)
#else
#define TP_ID2(d) (1.0 / (d))
#endif
  double d2 = pxyz(c, 0.0, s.py);
  double id2 = TP_ID2(d2);
  double ipz;
  double pz = pxyz_r(c, s.px, s.py, ipz);
  if (ns == 0) {
    ebody_back(s, a, c, irho, rho, p[1], p[5], p[7], p[8], d2, id2, pz, ipz);
    return;
  }
  const double L1 = p[1], L2 = p[2], K1 = p[3], K2 = p[4], L1b = p[5], L2b = p[6];
  const double c1 = p[7], s1 = p[8], c2 = p[9], s2 = p[10];
  const double* poly = p + 24;
#if TP_FAST
  if constexpr (KK == TP_KICK_NONE) {
    // Pure dipole: the forward kernel applies the body as one rotation of (px, pzmx),
    // pzmx = pz - 1 - x irho (its sub-bodies' rotations composed), with x formed at the end
    // and the path terms telescoped (see esbend_body).  Undone as a whole: the inverse
    // rotation brings (px, pzmx) back to the element's entry, and, the rotation being
    // orthogonal, the same matrix carries the adjoint of (px, pzmx) back.  The element map, with B = beti + delta, dd = B^2 - 1 - py^2 and
    // pz(a) = sqrt(B^2 - a^2 - py^2):
    //   b0 = pz(px) - 1 - x irho, (aN, bN) = R (px, b0), x' = (pz(aN) - 1 - bN) rho,
    //   S = Ltot + rho (asin(px/d2) - asin(aN/d2)), y' = y + py S, z' = z + B S - Lbtot,
    //   px' = aN, d2 = sqrt(B^2 - py^2), d asin(a/d2)/da = 1/pz(a).
    const double xo = s.x, aN = s.px, py = s.py, B = c.B;
    const double Ltot = (double)ns * (2.0 * L1 + 2.0 * L2);
    const double Lbtot = (double)ns * (2.0 * L1b + 2.0 * L2b);
    const double dd = __builtin_fma(-py, py, (B - 1.0) * (B + 1.0));
    const double pzN = pz, ipzN = ipz;  // pxyz of the output px (the caller's)
    const double bN = __builtin_fma(-xo, irho, __builtin_fma(-aN, aN, dd) * rcp_nr(pzN + 1.0));
    if (__builtin_expect(__builtin_fma(aN, aN, bN * bN) < 0.999999 * (d2 * d2), 1)) {
      // adjoint parts that need only the output side
      const double aS = a.y * py + a.z * B;   // d/dS
      const double adif = aS * rho;           // d/d(asin difference)
      const double apzN = a.x * rho;          // x' = (pzN - 1 - bN) rho
      double ga = a.px - adif * ipzN - apzN * aN * ipzN;  // d/d aN
      double gb = -apzN;                                   // d/d bN
      // the body's one rotation (rot_body, as the forward applies it) undone
      const Rot R = rot_body(c1, s1, c2, s2, ns);
      const double ra = __builtin_fma(aN, R.c, -(bN * R.s));
      const double rb = __builtin_fma(aN, R.s, bN * R.c);
      {
        const double ga_ = __builtin_fma(ga, R.c, -(gb * R.s));
        gb = __builtin_fma(ga, R.s, gb * R.c);
        ga = ga_;
      }
      double ipz0;
      const double pz0 = pxyz_r(c, ra, py, ipz0);
      const double S = __builtin_fma(asin_diff(ra, pz0, aN, pzN, d2, id2), rho, Ltot);
      const double ad2 = adif * id2 * (aN * ipzN - ra * ipz0);
      // b0 = pz0 - 1 - x irho
      const double apz0 = gb;
      const double ax = -gb * irho;
      ga = ga + adif * ipz0 - apz0 * ra * ipz0;
      const double apy = a.py + a.y * S - apzN * py * ipzN - apz0 * py * ipz0 - ad2 * py * id2;
      const double aB = a.z * S + (apzN * ipzN + apz0 * ipz0 + ad2 * id2) * B;
      s.x = (__builtin_fma(-ra, ra, dd) * rcp_nr(pz0 + 1.0) - rb) * rho;
      s.px = ra;
      s.y = __builtin_fma(-py, S, s.y);
      s.z = s.z - __builtin_fma(B, S, -Lbtot);
      a.x = ax;
      a.px = ga;
      a.py = apy;
      a.d = a.d + aB;
      return;
    }
    // a grazing particle: the forward kernel took the reference's sub-bodies; so does this
  }
#endif
  // The four bodies of a step (lengths L1 L2 L2 L1, a palindrome) one at a time: a rolled
  // loop keeps one body's reverse sweep live instead of four interleaved ones.
  for (int st = 0; st < ns; ++st) {
#pragma nounroll
    for (int j = 0; j < 4; ++j) {
      const bool outer = (j == 0) || (j == 3);
      ebody_back(s, a, c, irho, rho, outer ? L1 : L2, outer ? L1b : L2b, outer ? c1 : c2, outer ? s1 : s2, d2, id2,
                 pz, ipz);
      if (j < 3) {
        kick_back<KK>(s, a, poly, m, j == 1 ? K2 : K1, -1, gdummy);
        if (KK != TP_KICK_NONE) {
          d2 = pxyz(c, 0.0, s.py);
          id2 = TP_ID2(d2);
          pz = pxyz_r(c, s.px, s.py, ipz);
        }
      }
    }
  }
#undef TP_ID2
}

TP_DEV void sub6(St& s, const double* __restrict__ t) {
  s.x = s.x - t[0];
  s.px = s.px - t[1];
  s.y = s.y - t[2];
  s.py = s.py - t[3];
  s.z = s.z - t[4];
  s.d = s.d - t[5];
}

// One element backwards: s is the element's output on entry, its input on return.  minv
// holds the inverses of the r1 / r2 matrices at the pool offsets of the matrices (variant
// with FT_MISALIGN); lin selects the linearised drift (variant with FT_LINH).
// DCH: the element may change delta (changes_delta); with DCH = false the code that renews
// the delta hoists (the RF kick, the other families, misalignments) is not compiled in.
// slot >= 0: the element's strength-gradient slot; the wave's sum of the lanes' gradients is
// added to row[slot] inside the magnet's case, so the other element types reach the loop's
// back-edge straight from their maps.  Every lane adds the sum: one read and one write of the
// same LDS word by the whole wave, so the row gains it once.
template <unsigned FEAT, bool SER_OK, bool DCH = true>
TP_DEV void apply_elem_back(St& s, Ad& a, Cx& c, EdgeMemo* mm, const tp_rec& rec, const double* __restrict__ pool,
                            const double* __restrict__ minv, double beti, bool lin, int32_t slot, double* row) {
  // the series drift needs two more hoisted values; kernels with general kicks never take it
  constexpr bool SER = SER_OK && (FEAT & FT_GENKICK) == 0;
  const double* __restrict__ p = pool + rec.off;
  const uint32_t fl = rec.flags;
  const int m = rec.max_order, ns = rec.nsteps, g = rec.gorder;
  if (DCH && (fl & (TP_F_T2 | TP_F_R2))) {
    if (fl & TP_F_T2) sub6(s, pool + rec.moff + 6);
    if constexpr ((FEAT & FT_MISALIGN) != 0) {
      if (fl & TP_F_R2) mat6_back(s, a, pool + rec.moff + 48, minv + rec.moff + 48);
    }
    mkctx(s, c, beti);
  }
  switch (rec.tag) {
    case TP_DRIFT:
      drift_back<SER>(s, a, c, p[0], p[1], lin);
      break;
    case TP_MPOLE: {
      double gacc = 0.0;
#define TP_CALL(KK) mpole_back<KK, SER>(s, a, c, p, ns, m, g, gacc, lin)
      TP_KICK_SWITCH(FEAT, TP_CALL)
#undef TP_CALL
      if (slot >= 0) row[slot] += wave_sum(gacc);
      break;
    }
    case TP_THINMP: {
      double gacc = 0.0;
#define TP_CALL(KK) kick_back<KK>(s, a, p, m, 1.0, -1, gacc)
      TP_KICK_SWITCH(FEAT, TP_CALL)
#undef TP_CALL
      break;
    }
    case TP_SBEND:
      if constexpr ((FEAT & FT_SBEND) != 0) {
        const int fin = (fl >> TP_F_FIN_SHIFT) & 0xF, fout = (fl >> TP_F_FOUT_SHIFT) & 0xF;
        const double irho = p[6];
        if (fout != 0) edge_back<true>(s, a, fout, p + 13, irho, true, mm);
#define TP_CALL(KK) sbend_back<KK, SER>(s, a, c, p, ns, m, beti, lin)
        TP_KICK_SWITCH(FEAT, TP_CALL)
#undef TP_CALL
        if (fin != 0) edge_back<true>(s, a, fin, p + 9, irho, false, mm);
      }
      break;
    case TP_ESBEND_LE:
      if constexpr ((FEAT & FT_ESBEND_LE) != 0) {
        const int fin = (fl >> TP_F_FIN_SHIFT) & 0xF, fout = (fl >> TP_F_FOUT_SHIFT) & 0xF;
        const double irho = p[0];
        if (fout != 0) edge_back<true>(s, a, fout, p + 15, irho, true, mm);
#define TP_CALL(KK) esbend_body_back<KK>(s, a, c, p, ns, m)
        TP_KICK_SWITCH(FEAT, TP_CALL)
#undef TP_CALL
        if (fin != 0) edge_back<true>(s, a, fin, p + 11, irho, false, mm);
      }
      break;
    case TP_ESBEND:
      // ExactSBend with Forest's yrot / fringe / wedge maps, src/tracking.jl:901-973, undone
      // in reverse order (the host refuses the drift-limit case |irho| < 1e-6)
      if constexpr ((FEAT & FT_ESBEND) != 0) {
        const int fin = (fl >> TP_F_FIN_SHIFT) & 0xF, fout = (fl >> TP_F_FOUT_SHIFT) & 0xF;
        const double irho = p[0];
        SA r{s, a};
        if (p[20] != 0.0) r = yrot_back(r.s, r.a, c, p[21], p[22]);  // p[20] = -e2
        if (fout != 0) r = fringe_back(r.s, r.a, c, -irho, p[23]);
        r = wedge_back(r.s, r.a, c, irho, p[17], p[18], p[19], p[20]);
        s = r.s;
        a = r.a;
#define TP_CALL(KK) esbend_body_back<KK>(s, a, c, p, ns, m)
        TP_KICK_SWITCH(FEAT, TP_CALL)
#undef TP_CALL
        r = wedge_back(s, a, c, irho, p[13], p[14], p[15], p[16]);
        if (fin != 0) r = fringe_back(r.s, r.a, c, irho, p[23]);
        if (p[16] != 0.0) r = yrot_back(r.s, r.a, c, p[11], p[12]);  // p[16] = -e1
        s = r.s;
        a = r.a;
      }
      break;
    case TP_RFCAV:
      if constexpr ((FEAT & FT_RF) != 0) {
        const bool dr = (fl & TP_F_RF_DRIFT) != 0;
        if (dr) drift_back<SER>(s, a, c, p[0], p[1], lin);
        if (DCH && (fl & TP_F_RF_KICK)) {
          rf_kick_back(s, a, p);
          mkctx(s, c, beti);
        }
        if (dr) drift_back<SER>(s, a, c, p[0], p[1], lin);
      }
      break;
    case TP_MARKER:
      break;
    default:  // the other families (tags TP_SOLENOID ..; the host refuses the wiggler)
      if constexpr (DCH && (FEAT & FT_EXTRA) != 0) {
        const SA r = extra_back(s, a, rec.tag, fl, ns, p, beti, lin);
        s = r.s;
        a = r.a;
        mkctx(s, c, beti);
      }
      break;
  }
  if (DCH && (fl & (TP_F_T1 | TP_F_R1))) {
    if constexpr ((FEAT & FT_MISALIGN) != 0) {
      if (fl & TP_F_R1) mat6_back(s, a, pool + rec.moff + 12, minv + rec.moff + 12);
    }
    if (fl & TP_F_T1) sub6(s, pool + rec.moff);
    mkctx(s, c, beti);
  }
}

// An element that changes delta, backwards.  Without misalignments only the RF kick and the
// other families do, so only their cases are compiled here.
template <unsigned FEAT, bool SER_OK>
TP_DEV void apply_elem_back_dch(St& s, Ad& a, Cx& c, EdgeMemo* mm, const tp_rec& rec,
                                const double* __restrict__ pool, const double* __restrict__ minv, double beti,
                                bool lin, int32_t slot, double* row) {
  if constexpr ((FEAT & FT_MISALIGN) != 0) {
    apply_elem_back<FEAT, SER_OK, true>(s, a, c, mm, rec, pool, minv, beti, lin, slot, row);
  } else {
    constexpr bool SER = SER_OK && (FEAT & FT_GENKICK) == 0;
    const double* __restrict__ p = pool + rec.off;
    const uint32_t fl = rec.flags;
    if ((FEAT & FT_RF) != 0 && rec.tag == TP_RFCAV) {
      const bool dr = (fl & TP_F_RF_DRIFT) != 0;
      if (dr) drift_back<SER>(s, a, c, p[0], p[1], lin);
      rf_kick_back(s, a, p);
      mkctx(s, c, beti);
      if (dr) drift_back<SER>(s, a, c, p[0], p[1], lin);
    } else if constexpr ((FEAT & FT_EXTRA) != 0) {
      const SA r = extra_back(s, a, rec.tag, fl, rec.nsteps, p, beti, lin);
      s = r.s;
      a = r.a;
      mkctx(s, c, beti);
    }
  }
}

// nturns whole turns backwards for one lane: s / l enter as the state and adjoint at the
// end and leave as those at the start.  The strength gradient of every position with a
// slot is summed over the wave and added to the wave's LDS row by lane 0 (no strength
// gradients when slot_of_pos is null).  A time-varying table (n_rows > 1) holds the
// lattice of turn r + row_turn0 in sequence row r; the turns undone here are turn_last,
// turn_last - 1, ..., each with its own row.
template <unsigned FEAT, bool SER_OK>
TP_DEV void back_turns(St& s, Ad& l, Cx& c, EdgeMemo* mm, const int32_t* __restrict__ seq_rows, const tp_rec* __restrict__ recs,
                       const double* __restrict__ pool, const double* __restrict__ minv,
                       const int32_t* __restrict__ slot_of_pos, int32_t n_seq, int32_t nturns, int32_t turn_last,
                       int32_t n_rows, int32_t row_turn0, double beti, bool lin, double* row) {
  for (int32_t t = 0; t < nturns; ++t) {
    const int32_t* __restrict__ seq =
        seq_rows + (n_rows > 1 ? (int64_t)(turn_last - t - row_turn0) * n_seq : 0);
    // scalar loads pipelined as in the forward element loop: the record and slot of
    // element e-1 and the sequence entry of e-2 are requested while element e runs
    int32_t ri_next = n_seq > 1 ? seq[n_seq - 2] : 0;
    tp_rec rec_next = n_seq > 0 ? recs[seq[n_seq - 1]] : tp_rec{};
    int32_t slot_next = (slot_of_pos && n_seq > 0) ? slot_of_pos[n_seq - 1] : -1;
    int32_t e = n_seq - 1;
    auto advance = [&]() {
      if (e >= 1) {
        rec_next = recs[ri_next];
        slot_next = slot_of_pos ? slot_of_pos[e - 1] : -1;
        if (e >= 2) ri_next = seq[e - 2];
      }
    };
    while (e >= 0) {
      // elements that leave delta unchanged: the delta hoists are invariant in this loop
      // (changes_delta); an element that renews them ends it and runs below
      for (; e >= 0; --e) {
        const tp_rec rec = rec_next;
        const int32_t slot = slot_next;
        if (changes_delta<FEAT>(rec)) break;
        advance();
        apply_elem_back<FEAT, SER_OK, false>(s, l, c, mm, rec, pool, minv, beti, lin, slot, row);
      }
      if (e < 0) break;
      const tp_rec rec = rec_next;
      const int32_t slot = slot_next;
      advance();
      apply_elem_back_dch<FEAT, SER_OK>(s, l, c, mm, rec, pool, minv, beti, lin, slot, row);
      --e;
    }
  }
}

// WPE > 0 asks the register allocator for at least WPE resident waves per SIMD (spilling
// if it must); WPE = 0 leaves the budget to the compiler.
#define TP_WAVES(WPE) __attribute__((amdgpu_waves_per_eu((WPE) > 0 ? (WPE) : 1)))

template <unsigned FEAT, int WPE>
__global__ __launch_bounds__(256) TP_WAVES(WPE) void adjoint_kernel(const int32_t* __restrict__ seq, const tp_rec* __restrict__ recs,
                                                      const double* __restrict__ pool,
                                                      const int32_t* __restrict__ slot_of_pos,
                                                      const double* __restrict__ end_state, double* __restrict__ lam,
                                                      const int32_t* __restrict__ lost_turn,
                                                      double* __restrict__ partial, TpAdjArgs a) {
  extern __shared__ double lds[];  // [waves per block][n_slots], then the edge memo's [waves][128]
  const int nw = blockDim.x / 64, wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  for (int k = threadIdx.x; k < nw * a.n_slots; k += blockDim.x) lds[k] = 0.0;
  EdgeMemo mm{0, 0, lds + nw * a.n_slots + 128 * __builtin_amdgcn_readfirstlane(wave)};
  mm.w[lane] = __builtin_nan("");
  __syncthreads();
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool alive = (i < a.n) && (lost_turn[i] == -1);
  St s{0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  Ad l{0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  if (alive) {
    const double* q = end_state + i;
    s = St{q[0], q[a.ld], q[2 * a.ld], q[3 * a.ld], q[4 * a.ld], q[5 * a.ld]};
    if (a.init_lambda == 1) {
      l.x = 2.0 * s.x;
      l.y = 2.0 * s.y;
    } else if (a.init_lambda >= 2) {  // unit seed: row init_lambda - 2 of the Jacobian
      switch (a.init_lambda - 2) {
        case 0: l.x = 1.0; break;
        case 1: l.px = 1.0; break;
        case 2: l.y = 1.0; break;
        case 3: l.py = 1.0; break;
        case 4: l.z = 1.0; break;
        default: l.d = 1.0; break;
      }
    } else {
      const double* r = lam + i;
      l = Ad{r[0], r[a.ld], r[2 * a.ld], r[3 * a.ld], r[4 * a.ld], r[5 * a.ld]};
    }
  }
  Cx c;
  mkctx(s, c, a.beti);
  const bool lin = (FEAT & FT_LINH) != 0 && a.exact_h == 0;
  back_turns<FEAT, false>(s, l, c, &mm, seq, recs, pool, a.minv, slot_of_pos, a.n_seq, a.nturns, a.turn_last, a.n_rows,
                   a.row_turn0, a.beti, lin, lds + wave * a.n_slots);
  if (i < a.n) {
    double* r = lam + i;
    r[0] = l.x;
    r[a.ld] = l.px;
    r[2 * a.ld] = l.y;
    r[3 * a.ld] = l.py;
    r[4 * a.ld] = l.z;
    r[5 * a.ld] = l.d;
  }
  __syncthreads();
  for (int k = threadIdx.x; k < a.n_slots; k += blockDim.x) {
    double acc = 0.0;
    for (int w = 0; w < nw; ++w) acc += lds[w * a.n_slots + k];
    partial[(int64_t)blockIdx.x * a.n_slots + k] += acc;
  }
}

// Batched persistent reverse pass: waves take (variant, tile) items from the queue, as in
// the batched forward kernel that stored the checkpoints.  Each item walks its segments
// backwards from the stored states, in the same per-lane order as adjoint_kernel, keeps
// its strength gradient in the wave's own LDS row and writes the row to its item slot;
// the per-variant sums are formed afterwards in tile order (batch_reduce_kernel).
template <unsigned FEAT, int WPE>
__global__ __launch_bounds__(256) TP_WAVES(WPE) void batch_back_kernel(const TpVarDesc* __restrict__ desc,
                                                         const int2* __restrict__ items, int32_t n_items,
                                                         int32_t* __restrict__ counter, const int32_t* __restrict__ seq,
                                                         const tp_rec* __restrict__ recs,
                                                         const double* __restrict__ pool,
                                                         const double* __restrict__ minv,
                                                         const int32_t* __restrict__ slots,
                                                         const double* __restrict__ ckpt,
                                                         double* __restrict__ item_partial, int32_t max_slots,
                                                         int32_t exact_h) {
  extern __shared__ double lds[];  // [waves per block][max_slots], then the edge memo's [waves][128]
  const int wave = threadIdx.x / 64, lane = threadIdx.x & 63;
  const bool lin = (FEAT & FT_LINH) != 0 && exact_h == 0;
  double* row = lds + wave * max_slots;
  EdgeMemo mm{0, 0, lds + (blockDim.x / 64) * max_slots + 128 * __builtin_amdgcn_readfirstlane(wave)};
  mm.w[lane] = __builtin_nan("");
  for (;;) {
    int32_t it = 0;
    if (lane == 0) it = atomicAdd(counter, 1);
    it = __builtin_amdgcn_readfirstlane(it);
    if (it >= n_items) break;
    const int2 item = items[it];
    const TpVarDesc& d = desc[item.x];
    for (int k = lane; k < d.n_slots; k += 64) row[k] = 0.0;
    const int64_t li = (int64_t)item.y * 64 + lane;
    const bool alive = (li < d.n) && (d.lost_turn[li] == -1);
    const double* ck = ckpt + d.ck_off + (int64_t)item.y * d.nseg * 384 + lane;
    St s{0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    Ad l{0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    Cx c;
    for (int32_t j = d.nseg - 1; j >= 0; --j) {
      if (alive) {
        const double* q = ck + (int64_t)j * 384;
        s = St{q[0], q[64], q[128], q[192], q[256], q[320]};
        if (j == d.nseg - 1) {
          l.x = 2.0 * s.x;
          l.y = 2.0 * s.y;
        }
      }
      mkctx(s, c, d.beti);
      const int32_t nt = (j == d.nseg - 1) ? d.nturns - j * d.ckpt_stride : d.ckpt_stride;
      back_turns<FEAT, true>(s, l, c, &mm, seq + d.seq_off, recs + d.rec_off, pool + d.pool_off,
                       minv ? minv + d.pool_off : nullptr, slots + d.pos_off, d.n_seq, nt,
                       d.turn0 + j * d.ckpt_stride + nt - 1, d.n_rows, d.row_turn0, d.beti, lin, row);
    }
    if (li < d.n) {
      double* r = d.lam + li;
      r[0] = l.x;
      r[d.n] = l.px;
      r[2 * d.n] = l.y;
      r[3 * d.n] = l.py;
      r[4 * d.n] = l.z;
      r[5 * d.n] = l.d;
    }
    for (int k = lane; k < d.n_slots; k += 64) item_partial[(int64_t)it * max_slots + k] = row[k];
  }
}

// One block per variant: its gradient slots summed over its tiles in tile order, and its
// loss sum and survivor count likewise (the items of a variant are contiguous).
__global__ void batch_reduce_kernel(const TpVarDesc* __restrict__ desc, const int32_t* __restrict__ vfirst,
                                    const int32_t* __restrict__ vtiles, const double* __restrict__ item_partial,
                                    int32_t max_slots, const double* __restrict__ item_loss, double* __restrict__ grad,
                                    double* __restrict__ vloss) {
  const int v = blockIdx.x;
  const TpVarDesc& d = desc[v];
  const int64_t first = vfirst[v];
  const int32_t nt = vtiles[v];
  for (int k = threadIdx.x; k < d.n_slots; k += blockDim.x) {
    double acc = 0.0;
    for (int32_t i = 0; i < nt; ++i) acc += item_partial[(first + i) * max_slots + k];
    grad[d.slot_off + k] = acc;
  }
  if (threadIdx.x == 0) {
    double ls = 0.0, na = 0.0;
    for (int32_t i = 0; i < nt; ++i) {
      ls += item_loss[2 * (first + i)];
      na += item_loss[2 * (first + i) + 1];
    }
    vloss[2 * v] = ls;
    vloss[2 * v + 1] = na;
  }
}

// grad[k] = sum over blocks b of partial[b][k], in block order.
__global__ void reduce_kernel(const double* __restrict__ partial, int nblocks, int n_slots, double* __restrict__ grad) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n_slots) return;
  double acc = 0.0;
  for (int b = 0; b < nblocks; ++b) acc += partial[(int64_t)b * n_slots + k];
  grad[k] = acc;
}

// Register budgets compiled: the default, and at least 3 or 4 waves per SIMD for variants
// B and C (the default leaves their reverse kernels at 2); the all-families variant at 3
// (its out-of-line maps would otherwise take up to 256 VGPRs and set the kernel's count).
#define TP_WPE_SWITCH(waves, CALL)                            \
  if constexpr (FEAT == TP_FEAT_B || FEAT == TP_FEAT_C) {     \
    if ((waves) == 3) { CALL(3); }                            \
    else if ((waves) == 4) { CALL(4); }                       \
    else { CALL(0); }                                         \
  } else if constexpr (FEAT == TP_FEAT_ADJ_ALL) {            \
    CALL(3);                                                  \
  } else {                                                    \
    CALL(0);                                                  \
  }

template <unsigned FEAT>
static void launch_v(const TpAdjArgs& a, unsigned grid, hipStream_t stream) {
  const size_t lds = (size_t)4 * ((size_t)(a.n_slots > 0 ? a.n_slots : 1) + 128) * sizeof(double);
#define TP_CALL(W)                                                                                          \
  hipLaunchKernelGGL((adjoint_kernel<FEAT, W>), dim3(grid), dim3(256), lds, stream, a.seq, a.recs, a.pool, \
                     a.slot_of_pos, a.end_state, a.lam, a.lost_turn, a.partial, a)
  TP_WPE_SWITCH(a.waves, TP_CALL)
#undef TP_CALL
}

template <unsigned FEAT>
static void launch_back_v(const TpBatchArgs& b, hipStream_t stream) {
  const size_t lds = (size_t)(b.block / 64) * ((size_t)(b.max_slots > 0 ? b.max_slots : 1) + 128) * sizeof(double);
#define TP_CALL(W)                                                                                             \
  hipLaunchKernelGGL((batch_back_kernel<FEAT, W>), dim3(b.grid), dim3(b.block), lds, stream, b.desc, b.items, \
                     b.n_items, b.counter, b.seq, b.recs, b.pool, b.minv, b.slots, b.ckpt, b.item_partial,       \
                     b.max_slots, b.exact_h)
  TP_WPE_SWITCH(b.waves, TP_CALL)
#undef TP_CALL
}

template <unsigned FEAT>
static const void* back_fn(int waves) {
  const void* fn = nullptr;
#define TP_CALL(W) fn = reinterpret_cast<const void*>(batch_back_kernel<FEAT, W>)
  TP_WPE_SWITCH(waves, TP_CALL)
#undef TP_CALL
  return fn;
}

}  // namespace TPK_BUILD
}  // namespace tpk

hipError_t TP_BATCH_BACK_LAUNCH(const TpBatchArgs& b, hipStream_t stream) {
  if (b.n_items <= 0) return hipSuccess;
  switch (b.back_variant) {
    case TP_VARIANT_A: tpk::TPK_BUILD::launch_back_v<TP_FEAT_A>(b, stream); break;
    case TP_VARIANT_B: tpk::TPK_BUILD::launch_back_v<TP_FEAT_B>(b, stream); break;
    case TP_VARIANT_C: tpk::TPK_BUILD::launch_back_v<TP_FEAT_C>(b, stream); break;
    default: tpk::TPK_BUILD::launch_back_v<TP_FEAT_ADJ_ALL>(b, stream); break;
  }
  return hipGetLastError();
}

int TP_BATCH_BACK_OCC(int variant, int block, int max_slots, int waves) {
  const size_t lds = (size_t)(block / 64) * ((size_t)(max_slots > 0 ? max_slots : 1) + 128) * sizeof(double);
  const void* fn = nullptr;
  switch (variant) {
    case TP_VARIANT_A: fn = tpk::TPK_BUILD::back_fn<TP_FEAT_A>(waves); break;
    case TP_VARIANT_B: fn = tpk::TPK_BUILD::back_fn<TP_FEAT_B>(waves); break;
    case TP_VARIANT_C: fn = tpk::TPK_BUILD::back_fn<TP_FEAT_C>(waves); break;
    default: fn = tpk::TPK_BUILD::back_fn<TP_FEAT_ADJ_ALL>(waves); break;
  }
  int nb = 0;
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, block, lds) == hipSuccess ? nb : 0;
}

hipError_t TP_ADJ_LAUNCH(const TpAdjArgs& a, hipStream_t stream) {
  if (a.n <= 0) return hipSuccess;
  const int64_t grid64 = (a.n + 255) / 256;
  if (grid64 > 0x7fffffff) return hipErrorInvalidValue;
  const unsigned grid = (unsigned)grid64;
  switch (a.variant) {
    case TP_VARIANT_A: tpk::TPK_BUILD::launch_v<TP_FEAT_A>(a, grid, stream); break;
    case TP_VARIANT_B: tpk::TPK_BUILD::launch_v<TP_FEAT_B>(a, grid, stream); break;
    case TP_VARIANT_C: tpk::TPK_BUILD::launch_v<TP_FEAT_C>(a, grid, stream); break;
    default: tpk::TPK_BUILD::launch_v<TP_FEAT_ADJ_ALL>(a, grid, stream); break;
  }
  return hipGetLastError();
}

#if TP_FAST
hipError_t tp_launch_grad_reduce(const double* partial, int nblocks, int n_slots, double* grad, hipStream_t stream) {
  if (n_slots <= 0) return hipSuccess;
  hipLaunchKernelGGL(tpk::TPK_BUILD::reduce_kernel, dim3((n_slots + 255) / 256), dim3(256), 0, stream, partial,
                     nblocks, n_slots, grad);
  return hipGetLastError();
}

hipError_t tp_launch_batch_reduce(const TpVarDesc* desc, const int32_t* vfirst, const int32_t* vtiles, int32_t n_var,
                                  const double* item_partial, int32_t max_slots, const double* item_loss, double* grad,
                                  double* vloss, hipStream_t stream) {
  if (n_var <= 0) return hipSuccess;
  hipLaunchKernelGGL(tpk::TPK_BUILD::batch_reduce_kernel, dim3(n_var), dim3(256), 0, stream, desc, vfirst, vtiles,
                     item_partial, max_slots, item_loss, grad, vloss);
  return hipGetLastError();
}
#endif
