/*
 * trackpad_table.h -- flat device element table ("lattice table") for libtrackpad_hip.
 *
 * A host writer (trackpad_hip.table.compile_table in Python, TrackPadHIP.compile_table in
 * Julia) turns a static TrackPad Lattice + Beam into one little-endian byte blob:
 *
 *   tp_table_header                      64 bytes
 *   int32_t seq[n_rows][n_seq]           record index of each lattice position, one row per
 *                                        turn of a time-varying lattice (n_rows = 1 for a
 *                                        static one), padded with zeros to a multiple of 8
 *                                        bytes
 *   tp_rec  recs[n_rec]                  32 bytes each; identical elements share a record
 *   double  pool[pool_len]               per-element parameter blocks (see below)
 *
 * It replaces the reference's per-element structs (the files under src/elements/) and their
 * Adapt.adapt_structure GPU hooks (e.g. src/elements/quad.jl:113-118): every
 * particle-independent quantity the reference recomputes inside pass! on each call is
 * computed once by the writer, with the same floating-point operations, so the device
 * reproduces the reference's values bit for bit.
 *
 * Parameter blocks (offsets in doubles from rec.off; m = rec.max_order, 0 <= m <= 10;
 * "poly" = interleaved (b0, a0, b1, a1, ..., bm, am) after the reference's substitutions
 * b1 <- k1, b2 <- k2/2, b3 <- k3/6 and the kick-angle correction, src/tracking.jl:395-414,
 * 446, 478, 670-683, 935-942):
 *
 *   TP_MARKER    (none)
 *   TP_DRIFT     [0] L  [1] L*beti
 *   TP_MPOLE     [0] L1 [1] L2 [2] K1 [3] K2 [4] L1*beti [5] L2*beti [6..] poly
 *                with SL = L/nsteps, L1 = SL*DRIFT1, L2 = SL*DRIFT2, K1 = SL*KICK1,
 *                K2 = SL*KICK2 (src/tracking.jl:324-328)
 *   TP_THINMP    [0..] poly (kick length 1, src/tracking.jl:624)
 *   TP_SBEND     [0..5] as TP_MPOLE  [6] irho  [7] K1*irho  [8] K2*irho
 *                [9] fx_in [10] fy_in [11] e1 [12] fc_in
 *                [13] fx_out [14] fy_out [15] e2 [16] fc_out   [17..] poly
 *                fx = irho*tan(e), fc = the fringecorr of src/tracking.jl:162-168,
 *                fy = irho*tan(e) (used when fc == 0 and the method is Brown/fallback)
 *   TP_ESBEND / TP_ESBEND_LE
 *                [0] irho [1] L1 [2] L2 [3] K1 [4] K2 [5] L1*beti [6] L2*beti
 *                [7] cos(irho*L1) [8] sin(irho*L1) [9] cos(irho*L2) [10] sin(irho*L2)
 *                (nsteps == 0: [1] = L, [5] = L*beti, [7],[8] = cos/sin(irho*L))
 *                TP_ESBEND:    [11] cos(e1) [12] sin(e1)                  yrot entrance
 *                              [13] cos(-e1) [14] sin(-e1) [15] sin(-2e1) [16] -e1  edge in
 *                              [17] cos(-e2) [18] sin(-e2) [19] sin(-2e2) [20] -e2  edge out
 *                              [21] cos(e2) [22] sin(e2)                  yrot exit
 *                              [23] gk
 *                TP_ESBEND_LE: [11] fx_in [12] fy_in [13] e1 [14] fc_in
 *                              [15] fx_out [16] fy_out [17] e2 [18] fc_out
 *                [24..] poly
 *   TP_RFCAV     [0] L/2 [1] (L/2)*beti [2] volt/energy [3] 2pi*freq [4] lag [5] philag
 *                [6] beta*beta
 *   TP_SOLENOID  [0] L [1] ks  (ks = 0 is written as TP_DRIFT)            src/tracking.jl:529-566
 *   TP_CORRECTOR [0] L [1] xkick [2] ykick [3] xkick^2/3 + ykick^2/3 [4] xkick/2 [5] ykick/2
 *                                                                         src/tracking.jl:579-605
 *   TP_TRANSLATION [0] dx [1] dy [2] ds                                   src/tracking.jl:1182-1192
 *   TP_YROTATION [0] cos(a) [1] sin(a) [2] tan(a), a = -angle (0 is written as TP_MARKER)
 *                                                                         src/tracking.jl:1194-1218
 *   TP_CRAB      [0] L/2 [1] (L/2)*beti [2] volt'/E [3] k [4] phi' [5] k*volt'/E*beti
 *                (volt' = volt(1+err1), phi' = phi+err2, E = max(|energy|, eps))
 *                                                                         src/tracking.jl:1531-1547
 *   TP_ACCEL     [0] volt/(beta^2 E) [1] k [2] phis [3] sin(phis)        src/tracking.jl:1549-1559
 *   TP_LONGRF    [0] 2pi*h*eta/k (eta = alphac - (1 - beta^2))            src/tracking.jl:1575-1585
 *   TP_BOOST     [0] tan(angle) [1] 1/cos(angle)                          src/tracking.jl:1587-1595
 *   TP_INVBOOST  [0] sin(angle) [1] cos(angle)                            src/tracking.jl:1597-1604
 *   TP_GAUSSBB   [0] sx^2 [1] sy^2 [2] factor, then (weight, xoff, yoff) for each of
 *                rec.nsteps slices (StrongThinGaussianBeam: one slice, factor 1)
 *                                                                         src/tracking.jl:1610-1641
 *   TP_RLCWAKE   [0] scale [1] Rshunt*w0/Q0 [2] w0' [3] 2*Q0' [4] w0 [5] 2*Q0
 *                (w0 = 2pi*freq, Q0' = sqrt(Q0^2 - 1/4), w0' = w0/Q0*Q0')  src/tracking.jl:1647-1654
 *   TP_TABWAKE   [0] scale [1] fliphalf [2] times[0]*fliphalf, then times[n] and wakes[n],
 *                n = rec.nsteps                                           src/tracking.jl:1656-1663
 *   TP_LBEND     [0] L [1] K [2] angle [3] by_error [4] angle/L, then the entrance and exit
 *                edge blocks (fx, fy, e, fc) of TP_SBEND at [5] and [9]   src/tracking.jl:1076-1170
 *   TP_WIGGLER   [0] Nstep [1] dl*1.3512.. [2] dl*(-1.7024..) [3] NH [4] NV [5] rad
 *                [6] Po*m/c [7] srCoef [8] dl (dl = lw/Nsteps, Nstep = rec.nsteps), then 11
 *                doubles per harmonic, NH of By then NV of Bx: kx ky kz tz, the leading
 *                factors of the ax, axpy, ay, aypx terms and of Bx, By, and 1 / 0 for the
 *                sin(k u)/k branch                                        src/tracking.jl:1224-1525
 *   TP_NAN       (none): every coordinate becomes NaN -- the reference's _nan6 for SBendSC and
 *                LBend of zero length (src/tracking.jl:1047-1049, 1158-1160)
 *
 *   Misalignment block at rec.moff (present when any TP_F_T1/T2/R1/R2 flag is set):
 *                [0..6) t1  [6..12) t2  [12..48) r1  [48..84) r2   (r1, r2 column-major)
 */
#ifndef TRACKPAD_TABLE_H
#define TRACKPAD_TABLE_H

#include <stdint.h>

#define TP_TABLE_MAGIC 0x54504854u /* bytes 'T','H','P','T' */
#define TP_TABLE_VERSION 1u
#define TP_MAXORD 10

enum tp_tag {
  TP_MARKER = 0,
  TP_DRIFT = 1,
  TP_MPOLE = 2,      /* Quadrupole / Sextupole / Octupole: Yoshida-4 + strthinkick */
  TP_THINMP = 3,     /* ThinMultipole */
  TP_SBEND = 4,      /* SBend / RBend: linear edges + Yoshida-4 + bndthinkick */
  TP_ESBEND = 5,     /* ExactSBend / ERBend: yrot, Forest fringe, wedge, exact body */
  TP_ESBEND_LE = 6,  /* exact body with the linear edge maps of SBend */
  TP_RFCAV = 7,      /* RFCavity */
  TP_SOLENOID = 8,   /* Solenoid with ks != 0 */
  TP_CORRECTOR = 9,  /* Corrector, HKicker, VKicker */
  TP_TRANSLATION = 10,
  TP_YROTATION = 11,
  TP_CRAB = 12,      /* CrabCavity */
  TP_ACCEL = 13,     /* AccelCavity */
  TP_LONGRF = 14,    /* LongitudinalRFMap */
  TP_BOOST = 15,     /* LorentzBoost (mode 0) */
  TP_INVBOOST = 16,  /* InvLorentzBoost (mode 0) */
  TP_GAUSSBB = 17,   /* StrongThinGaussianBeam, StrongGaussianBeam */
  TP_RLCWAKE = 18,   /* LongitudinalRLCWake */
  TP_TABWAKE = 19,   /* LongitudinalWake (tabulated) */
  TP_LBEND = 20,     /* LBend */
  TP_NAN = 21,       /* the reference's NaN map */
  TP_WIGGLER = 22,   /* Wiggler */
  TP_NTAGS = 23
};
/* Also written with the tags above: DriftSC (TP_DRIFT), QuadrupoleSC / SextupoleSC /
 * OctupoleSC (TP_MPOLE), SBendSC / RBendSC (TP_SBEND), SpaceCharge (TP_MARKER: its
 * single-particle map is the identity, src/tracking.jl:1176-1180). */

/* rec.flags */
#define TP_F_T1 0x1u
#define TP_F_T2 0x2u
#define TP_F_R1 0x4u
#define TP_F_R2 0x8u
#define TP_F_MISALIGN (TP_F_T1 | TP_F_T2 | TP_F_R1 | TP_F_R2)
#define TP_F_FIN_SHIFT 4  /* 4 bits: entrance edge method (SBEND, ESBEND_LE) or fringe on (ESBEND) */
#define TP_F_FOUT_SHIFT 8 /* 4 bits: exit edge method / fringe on */
#define TP_F_DRIFTLIMIT 0x1000u /* exact bend with |irho| < 1e-6: body is a drift, edges skipped */
#define TP_F_RF_DRIFT 0x2000u   /* RFCavity with L > 0 */
#define TP_F_RF_KICK 0x4000u    /* RFCavity with energy > 0 */

/* header.flags */
#define TP_H_EXACT_HAMILTONIAN 0x1u /* USE_EXACT_HAMILTONIAN (src/tracking.jl:38) */

typedef struct tp_table_header {
  uint32_t magic;
  uint32_t version;
  uint32_t n_seq;    /* lattice positions */
  uint32_t n_rec;    /* distinct element records */
  uint32_t pool_len; /* doubles in the parameter pool */
  uint32_t flags;    /* TP_H_* */
  uint32_t n_rows;   /* sequence rows: 0 or 1 = a static lattice; R > 1 = row r is the
                        lattice resolved at turn row_turn0 + r (time-varying elements
                        materialised per turn, src/lattice.jl:267, src/time_dependence.jl) */
  int32_t row_turn0; /* turn of row 0 (time-varying tables only) */
  double beti;        /* 1/beta of the reference particle (src/lattice.jl:61) */
  double coord_limit; /* |x|,|y| loss limit, COORD_LIMIT (src/tracking.jl:25) */
  double angle_limit; /* |px|,|py| loss limit, ANGLE_LIMIT (src/tracking.jl:26) */
  double aperture_r2; /* circular aperture R^2 checked at element exit; 0 = none */
} tp_table_header;

typedef struct tp_rec {
  int32_t tag;       /* enum tp_tag */
  int32_t nsteps;    /* num_int_steps */
  int32_t max_order; /* polynomial order m */
  int32_t off;       /* pool offset of the parameter block */
  uint32_t flags;    /* TP_F_* */
  int32_t moff;      /* pool offset of the misalignment block, or -1 */
  int32_t gorder;    /* polynomial index of the main strength (1 quad, 2 sext, 3 oct), -1 none */
  int32_t kick;      /* written by the library at load time (kick specialisation); writers put 0 */
} tp_rec;

#endif /* TRACKPAD_TABLE_H */
