// tp_table_parse.h -- host-side reading of the element table blob (include/trackpad_table.h):
// bounds and consistency checks, the per-record kick specialisation and the feature mask that
// selects a kernel variant.  Plain C++ with no HIP dependency, so that the parser that guards
// every device read of a record's block also builds alone, with the host sanitizers, outside
// the library (tests/sanitize).
#ifndef TP_TABLE_PARSE_H
#define TP_TABLE_PARSE_H

#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <vector>

#include "trackpad_hip.h"
#include "trackpad_table.h"

// Kick specialisations, chosen per record at table load from the structural zeros of
// the polynomial.  Every specialisation evaluates the same products and sums as the
// full Horner recurrence of src/tracking.jl:290-298 minus the terms that are a product
// with an exact zero, so finite nonzero results are bit-identical; only the sign of an
// exactly-zero result can differ.
enum tp_kick {
  TP_KICK_NONE = 0,  // all coefficients zero: the kick is the identity
  TP_KICK_QUAD = 1,  // m = 1, only b1 nonzero
  TP_KICK_SEXT = 2,  // m = 2, only b2 nonzero
  TP_KICK_GEN = 3    // anything else: full Horner over max_order
};

// Kernel feature bits: a kernel variant contains only the element maps its mask names.
enum : unsigned {
  FT_SBEND = 1u,      // expanded-Hamiltonian bends (TP_SBEND)
  FT_ESBEND = 2u,     // Forest exact bends with yrot / fringe / wedge (TP_ESBEND)
  FT_ESBEND_LE = 4u,  // exact bend body with linear edges (TP_ESBEND_LE)
  FT_RF = 8u,         // RF cavities
  FT_MISALIGN = 16u,  // t1/t2/r1/r2 misalignments
  FT_GENKICK = 32u,   // general Horner kicks up to order 10
  FT_LINH = 64u,      // linearised drift (USE_EXACT_HAMILTONIAN = false)
  FT_EXTRA = 128u     // the other element families (tags TP_SOLENOID .. TP_NAN, tp_extra.h)
};
// Compiled variants, smallest first; the host picks the first whose mask covers the table.
#define TP_FEAT_A 0u                                              // drift/multipole lattices
#define TP_FEAT_B (FT_ESBEND_LE | FT_RF)                          // rings of exact bends + RF
#define TP_FEAT_C (FT_ESBEND_LE | FT_RF | FT_SBEND | FT_GENKICK)  // + high-order multipoles
#define TP_FEAT_ALL 255u
// The reverse kernels' largest variant: every family (the host refuses the wiggler); it is
// selected for Forest exact bends, r1/r2 rotations, the linearised drift and tags 8-21.
#define TP_FEAT_ADJ_ALL TP_FEAT_ALL
enum { TP_VARIANT_A = 0, TP_VARIANT_B = 1, TP_VARIANT_C = 2, TP_VARIANT_ALL = 3 };

namespace tpr {

// Sets the thread's error message (tp_last_error) and returns code.
int fail(int code, const char* fmt, ...);

// A parsed table: host copies of its parts, with the kick field of every record written.
struct tp_table_host {
  tp_table_header hdr{};
  unsigned features = 0;
  int variant = TP_VARIANT_ALL;
  std::vector<tp_rec> recs;   // with the kick specialisation filled in
  std::vector<int32_t> seq;   // n_rows x n_seq
  std::vector<double> pool;
};

// Length in doubles of a record's parameter block (trackpad_table.h); -2 for the wiggler,
// whose length depends on the harmonic counts in its block; -1 for an unknown tag.
inline int64_t block_len(int tag, int m, int nsteps) {
  const int64_t poly = 2 * (int64_t)(m + 1);
  switch (tag) {
    case TP_MARKER: return 0;
    case TP_DRIFT: return 2;
    case TP_MPOLE: return 6 + poly;
    case TP_THINMP: return poly;
    case TP_SBEND: return 17 + poly;
    case TP_ESBEND:
    case TP_ESBEND_LE: return 24 + poly;
    case TP_RFCAV: return 7;
    case TP_SOLENOID: return 2;
    case TP_CORRECTOR: return 6;
    case TP_TRANSLATION: return 3;
    case TP_YROTATION: return 3;
    case TP_CRAB: return 6;
    case TP_ACCEL: return 4;
    case TP_LONGRF: return 1;
    case TP_BOOST:
    case TP_INVBOOST: return 2;
    case TP_GAUSSBB: return 3 + 3 * (int64_t)nsteps;
    case TP_RLCWAKE: return 6;
    case TP_TABWAKE: return 3 + 2 * (int64_t)nsteps;
    case TP_LBEND: return 13;
    case TP_NAN: return 0;
    case TP_WIGGLER: return -2;
    default: return -1;
  }
}

// Offset of the interleaved (b_n, a_n) polynomial inside a record's block, or -1.
inline int64_t poly_off(int tag) {
  switch (tag) {
    case TP_MPOLE: return 6;
    case TP_THINMP: return 0;
    case TP_SBEND: return 17;
    case TP_ESBEND:
    case TP_ESBEND_LE: return 24;
    default: return -1;
  }
}

// Kick specialisation from the structural zeros of the polynomial (indices <= m).
inline int classify_kick(const double* poly, int m) {
  bool nz_b1 = false, nz_b2 = false, nz_other = false;
  for (int i = 0; i <= m; ++i) {
    const double b = poly[2 * i], a = poly[2 * i + 1];
    if (a != 0.0) nz_other = true;
    if (b != 0.0) {
      if (i == 1) nz_b1 = true;
      else if (i == 2) nz_b2 = true;
      else nz_other = true;
    }
  }
  if (nz_other || (nz_b1 && nz_b2)) return TP_KICK_GEN;
  if (nz_b1) return TP_KICK_QUAD;
  if (nz_b2) return TP_KICK_SEXT;
  return TP_KICK_NONE;
}

// Parses and validates a table blob into t; fills the kick specialisation and the feature
// mask.  Returns TP_OK, or a TP_E* code with the message set.  Every sequence entry, parameter
// block and misalignment block of an accepted blob lies inside it: the device kernels read
// them unchecked.
inline int parse_table(const void* bytes, size_t nbytes, tp_table_host* t) {
  if (!bytes) return fail(TP_EINVAL, "table: bytes is NULL");
  if (nbytes < sizeof(tp_table_header)) return fail(TP_ETABLE, "table: %zu bytes is shorter than the header", nbytes);
  tp_table_header h;
  std::memcpy(&h, bytes, sizeof h);
  if (h.magic != TP_TABLE_MAGIC) return fail(TP_ETABLE, "table: bad magic 0x%08x", h.magic);
  if (h.version != TP_TABLE_VERSION) return fail(TP_ETABLE, "table: version %u, library reads %u", h.version, TP_TABLE_VERSION);
  if (h.n_rows > 1000000u) return fail(TP_ETABLE, "table: %u sequence rows", h.n_rows);
  const size_t rows = h.n_rows > 1 ? h.n_rows : 1;
  if (rows > 1 && (int64_t)h.row_turn0 + (int64_t)rows > 0x7fffffffLL)
    return fail(TP_ETABLE, "table: sequence rows overflow the turn range");
  const size_t n_pos = rows * (size_t)h.n_seq;
  const size_t seq_bytes = (n_pos * 4 + 7) / 8 * 8;
  const size_t need = sizeof h + seq_bytes + (size_t)h.n_rec * sizeof(tp_rec) + (size_t)h.pool_len * 8;
  if (need != nbytes) return fail(TP_ETABLE, "table: size %zu, header implies %zu", nbytes, need);
  if (h.n_seq > 0 && h.n_rec == 0) return fail(TP_ETABLE, "table: %u positions but no records", h.n_seq);
  if (!(std::isfinite(h.beti) && h.beti >= 1.0)) return fail(TP_ETABLE, "table: beti %g is not >= 1", h.beti);
  if (!(h.coord_limit > 0.0) || !(h.angle_limit > 0.0) || !(h.aperture_r2 >= 0.0))
    return fail(TP_ETABLE, "table: bad loss limits");
  t->hdr = h;
  const char* b = (const char*)bytes;
  t->seq.resize(n_pos);
  if (n_pos) std::memcpy(t->seq.data(), b + sizeof h, n_pos * 4);
  t->recs.resize(h.n_rec);
  if (h.n_rec) std::memcpy(t->recs.data(), b + sizeof h + seq_bytes, (size_t)h.n_rec * sizeof(tp_rec));
  t->pool.resize(h.pool_len);
  if (h.pool_len)
    std::memcpy(t->pool.data(), b + sizeof h + seq_bytes + (size_t)h.n_rec * sizeof(tp_rec), (size_t)h.pool_len * 8);
  for (size_t i = 0; i < n_pos; ++i)
    if (t->seq[i] < 0 || (uint32_t)t->seq[i] >= h.n_rec)
      return fail(TP_ETABLE, "table: seq[%zu] = %d out of range", i, t->seq[i]);
  unsigned feat = 0;
  if (!(h.flags & TP_H_EXACT_HAMILTONIAN)) feat |= FT_LINH;
  for (uint32_t i = 0; i < h.n_rec; ++i) {
    tp_rec& r = t->recs[i];
    if (r.tag < 0 || r.tag >= TP_NTAGS) return fail(TP_EUNSUPPORTED, "table: record %u: unknown tag %d", i, r.tag);
    if (r.max_order < 0 || r.max_order > TP_MAXORD)
      return fail(TP_EUNSUPPORTED, "table: record %u: max_order %d outside [0, 10]", i, r.max_order);
    if (r.nsteps < 0 || r.nsteps > 1000000) return fail(TP_ETABLE, "table: record %u: nsteps %d", i, r.nsteps);
    if (r.tag == TP_GAUSSBB && r.nsteps < 1) return fail(TP_ETABLE, "table: record %u: beam-beam without slices", i);
    if (r.tag == TP_TABWAKE && r.nsteps < 2) return fail(TP_ETABLE, "table: record %u: wake table shorter than 2", i);
    int64_t len = block_len(r.tag, r.max_order, r.nsteps);
    if (len == -2) {  // TP_WIGGLER: 9 + 11 per harmonic
      if (r.off < 0 || (int64_t)r.off + 9 > (int64_t)h.pool_len)
        return fail(TP_ETABLE, "table: record %u: parameter block outside the pool", i);
      const double nh = t->pool[(size_t)r.off + 3], nv = t->pool[(size_t)r.off + 4];
      if (!(nh >= 0.0 && nv >= 0.0 && nh + nv <= 64.0) || r.nsteps < 0)
        return fail(TP_ETABLE, "table: record %u: wiggler with %g + %g harmonics", i, nh, nv);
      len = 9 + 11 * ((int64_t)nh + (int64_t)nv);
    }
    if (r.off < 0 || (int64_t)r.off + len > (int64_t)h.pool_len)
      return fail(TP_ETABLE, "table: record %u: parameter block outside the pool", i);
    if (r.flags & TP_F_MISALIGN) {
      if (r.tag == TP_MARKER || r.tag == TP_RFCAV ||
          (r.tag >= TP_TRANSLATION && r.tag != TP_LBEND && r.tag != TP_WIGGLER))
        return fail(TP_ETABLE, "table: record %u: misalignment on an element without one", i);
      if (r.moff < 0 || (int64_t)r.moff + 84 > (int64_t)h.pool_len)
        return fail(TP_ETABLE, "table: record %u: misalignment block outside the pool", i);
      feat |= FT_MISALIGN;
    }
    if (r.gorder < -1 || r.gorder > r.max_order) return fail(TP_ETABLE, "table: record %u: gorder %d", i, r.gorder);
    r.kick = TP_KICK_NONE;
    const int64_t po = poly_off(r.tag);
    if (po >= 0) {
      r.kick = classify_kick(t->pool.data() + r.off + po, r.max_order);
      if (r.kick == TP_KICK_GEN) feat |= FT_GENKICK;
    }
    if (r.tag == TP_SBEND) feat |= FT_SBEND;
    if (r.tag == TP_ESBEND) feat |= FT_ESBEND;
    if (r.tag == TP_ESBEND_LE) feat |= FT_ESBEND_LE;
    if (r.tag == TP_RFCAV) feat |= FT_RF;
    if (r.tag >= TP_SOLENOID) feat |= FT_EXTRA;
  }
  t->features = feat;
  if ((feat & ~TP_FEAT_A) == 0) t->variant = TP_VARIANT_A;
  else if ((feat & ~TP_FEAT_B) == 0) t->variant = TP_VARIANT_B;
  else if ((feat & ~TP_FEAT_C) == 0) t->variant = TP_VARIANT_C;
  else t->variant = TP_VARIANT_ALL;
  return TP_OK;
}

}  // namespace tpr

#endif  // TP_TABLE_PARSE_H
