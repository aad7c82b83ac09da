// tp_device.h -- declarations shared by the kernel translation units and the C ABI.
#ifndef TP_DEVICE_H
#define TP_DEVICE_H

#include <stdint.h>
#include <hip/hip_runtime.h>

#include "trackpad_table.h"
#include "tp_table_parse.h"

// tp_kick, the FT_* feature bits, TP_FEAT_* and TP_VARIANT_* live beside the host-side table
// parser that assigns them.

// Dynamic LDS of a tracking launch per wave of its block (tp_track_kernels.hip: the edge memo
// of each particle a lane holds).
#define TP_LDS_PER_WAVE(ppl) ((size_t)(ppl) * 128 * 8)

struct TpKArgs {
  const int32_t* seq;
  const tp_rec* recs;
  const double* pool;
  int32_t n_seq;
  int32_t exact_h;  // USE_EXACT_HAMILTONIAN
  double beti;
  double coord_limit;
  double angle_limit;
  double ap_r2;
  double* coords;  // 6 planes of ld doubles
  int64_t n;
  int64_t ld;
  int32_t* lost_turn;
  int32_t* lost_elem;
  int32_t turn0;
  int32_t nturns;
  // Element range of a launch: the first turn starts at element e_first, the last ends before
  // element e_last (0 = at the end of the turn); 0 / 0 = whole turns.
  int32_t e_first;
  int32_t e_last;
  int32_t variant;  // TP_VARIANT_*
  int32_t waves;    // register budget of the kernel: 0 (compiler's), 5 or 6 waves per SIMD
  int32_t n_rows;   // sequence rows (1: static lattice; > 1: one row per turn from row_turn0)
  int32_t row_turn0;
  // Particle list of the launch: lane i tracks particle order[i] (i < n); null = particle i.
  // The host compacts the list of survivors between launches (compact_alive).
  const int32_t* order;
  // Persistent launch: waves take tiles of 64 x PPL list entries from this queue counter
  // (cleared before the launch) until it runs dry; null = one tile per wave of the grid.
  int32_t* counter;
  // Fast build: series drift (1) or the refined reciprocal square root (0), chosen once per
  // call by the host (tp_set_drift); variants with general kicks always take the latter.
  int32_t ser;
};

// One lattice variant of a batch: where its element table sits in the concatenated
// tables, its particle set, and its turn range.
struct TpVarDesc {
  int64_t seq_off;   // first sequence entry (rows x n_seq entries)
  int64_t pos_off;   // first entry of the variant's slot_of_pos (n_seq entries)
  int64_t rec_off;   // first record
  int64_t pool_off;  // first pool double
  int64_t slot_off;  // first gradient slot of the variant in the concatenated gradient
  double* coords;    // the variant's particle set: 6 planes of n doubles
  int32_t* lost_turn;
  int32_t* lost_elem;
  double* lam;       // adjoint output (dL/dx0 before the 1/N_alive scale), 6 planes of n
  int64_t n;
  int64_t ck_off;    // first double of the variant's checkpoint slab (nseg x 6 x 64 per tile)
  int32_t n_seq;
  int32_t nturns;
  int32_t turn0;
  int32_t n_slots;
  int32_t nseg;         // ceil(nturns / ckpt_stride)
  int32_t ckpt_stride;
  int32_t n_rows;        // sequence rows of a time-varying lattice (1 = static)
  int32_t row_turn0;
  double beti;
  double coord_limit;
  double angle_limit;
  double ap_r2;
};

struct TpBatchArgs {
  const TpVarDesc* desc;
  const int2* items;  // (variant, tile of 64 particles), sorted by decreasing cost
  int32_t n_items;
  int32_t* counter;   // work-queue head, zeroed before the launch
  const int32_t* seq;  // concatenated tables of all variants
  const tp_rec* recs;
  const double* pool;
  const int32_t* slots;  // concatenated slot_of_pos
  const double* minv;    // concatenated r1/r2 inverses at pool offsets (adjoint; null without rotations)
  double* ckpt;          // checkpoint slabs (adjoint)
  double* item_loss;     // per item: loss sum, survivors (adjoint)
  double* item_partial;  // per item: max_slots gradient sums (adjoint)
  int32_t max_slots;
  int32_t variant;       // forward kernel variant
  int32_t back_variant;  // reverse kernel variant
  int32_t exact_h;
  int32_t grid;
  int32_t block;
  int32_t waves;  // register budget of the kernel launched (0 = the compiler's)
};

// Arguments of one backward segment of the adjoint (tp_adjoint_kernels.hip).
struct TpAdjArgs {
  const int32_t* seq;
  const tp_rec* recs;
  const double* pool;
  const int32_t* slot_of_pos;  // gradient slot of each lattice position, -1 for none (null: no
                               // strength gradients)
  const double* end_state;     // beam at the end of the segment (6 planes of ld)
  double* lam;                 // adjoint: at the segment end on entry, at its start on return
  const int32_t* lost_turn;    // loss record after the last turn (alive == -1)
  double* partial;             // [blocks][n_slots] strength-gradient partial sums (accumulated)
  int64_t n;
  int64_t ld;
  int32_t n_seq;
  int32_t n_slots;
  int32_t nturns;              // turns in this segment
  int32_t init_lambda;         // 0: continue from lam; 1: start from the gradient of the
                               // loss sum, 2x / 2y; 2 + i: start from the unit vector e_i
  int32_t variant;
  int32_t waves;               // register budget: 0 (compiler's), 3 or 4 waves per SIMD
  const double* minv;          // inverses of the r1/r2 matrices at their pool offsets (or null)
  int32_t exact_h;             // USE_EXACT_HAMILTONIAN (0: linearised drift)
  int32_t turn_last;           // turn number of the segment's last turn
  int32_t n_rows;              // sequence rows of a time-varying table (1 = static)
  int32_t row_turn0;           // turn of row 0
  double beti;
};

// Host launchers exported by the kernel translation units (one per build flavour).
// max_grid: blocks of a persistent launch (a.counter set; ignored otherwise)
hipError_t tp_launch_track_exact(const TpKArgs& a, int block, int ppl, int max_grid, hipStream_t stream);
hipError_t tp_launch_track_fast(const TpKArgs& a, int block, int ppl, int max_grid, hipStream_t stream);
hipError_t tp_launch_batch_exact(const TpBatchArgs& b, bool ckpt, hipStream_t stream);
hipError_t tp_launch_batch_fast(const TpBatchArgs& b, bool ckpt, hipStream_t stream);
int tp_batch_occupancy_exact(int variant, int exact_h, bool ckpt, int block, int waves);
// Resident blocks per CU of the tracking kernel that a launch shape selects.
// ser: the occupancy of the series-drift kernel (fast build, variants without general kicks).
int tp_track_occupancy_exact(int variant, int exact_h, int block, int ppl, int waves, int ser);
int tp_track_occupancy_fast(int variant, int exact_h, int block, int ppl, int waves, int ser);
int tp_batch_occupancy_fast(int variant, int exact_h, bool ckpt, int block, int waves);
hipError_t tp_launch_batch_back_exact(const TpBatchArgs& b, hipStream_t stream);
hipError_t tp_launch_batch_back_fast(const TpBatchArgs& b, hipStream_t stream);
int tp_batch_back_occupancy_exact(int variant, int block, int max_slots, int waves);
int tp_batch_back_occupancy_fast(int variant, int block, int max_slots, int waves);
// Per variant: gradient sums over its tiles in tile order, loss sum and survivors.
hipError_t tp_launch_batch_reduce(const TpVarDesc* desc, const int32_t* vfirst, const int32_t* vtiles, int32_t n_var,
                                  const double* item_partial, int32_t max_slots, const double* item_loss, double* grad,
                                  double* vloss, hipStream_t stream);
hipError_t tp_launch_fill_i32(int32_t* p, int32_t v, int32_t n, hipStream_t stream);
// order_out <- the entries of order_in (null: 0..n_in-1) whose particle is still alive
// (lost_turn == -1), in no particular order; *count (cleared by the caller) += their number.
hipError_t tp_launch_compact_alive(const int32_t* order_in, int32_t n_in, const int32_t* lost_turn, int32_t* order_out,
                                   int32_t* count, hipStream_t stream);
hipError_t tp_launch_adjoint_exact(const TpAdjArgs& a, hipStream_t stream);
hipError_t tp_launch_adjoint_fast(const TpAdjArgs& a, hipStream_t stream);
hipError_t tp_launch_grad_reduce(const double* partial, int nblocks, int n_slots, double* grad, hipStream_t stream);

#endif  // TP_DEVICE_H
