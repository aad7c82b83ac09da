// tp_runtime.h -- host-side objects of libtrackpad_hip shared by its translation units
// (tp_abi.hip: the C ABI; tp_autotune.hip: the launch-shape autotuner).
#ifndef TP_RUNTIME_H
#define TP_RUNTIME_H

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <string>
#include <vector>

#include "trackpad_hip.h"
#include "tp_device.h"

// A context is reference counted: every table and particle set created on it holds a
// reference, so destroy calls may come in any order (Julia and Python finalizers run in
// arbitrary order) and the stream outlives everything enqueued on it.
struct tp_ctx {
  std::atomic<int> refs{1};
  std::atomic<bool> closed{false};
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  int block = 256;
  int ppl = 1;
  int tpl = 0;  // turns per launch, 0 = all
  int fwd_waves = 0, back_waves = 0;  // register budgets (0 = the compiler's)
  int drift = TP_DRIFT_AUTO;           // drift form of the fast build (tp_set_drift)
  int cus = 0;                         // compute units of the device
  int32_t* d_counter = nullptr;        // work-queue counter of persistent tracking launches
  int32_t* d_scratch = nullptr;        // device address of the static scratch (tp_track_noalloc)
  tp_timing_t timing{};
};

// The parsed host copy (hdr, features, variant, recs, seq, pool) and the device copy.
struct tp_table : tpr::tp_table_host {
  tp_ctx* ctx = nullptr;
  void* dev = nullptr;
  size_t nbytes = 0;
  const int32_t* d_seq = nullptr;
  const tp_rec* d_recs = nullptr;
  const double* d_pool = nullptr;
  int32_t* d_slots = nullptr;    // gradient slot of each position (adjoint), built on first use
  int32_t n_slots = 0;
  double* d_minv = nullptr;      // r1/r2 inverses at the matrices' pool offsets (adjoint), on first use
};

// A particle set is reference counted too: a batch that binds it holds a reference.
struct tp_particles {
  std::atomic<int> refs{1};
  tp_ctx* ctx = nullptr;
  int64_t n = 0;
  double* d_coords = nullptr;
  int32_t* d_lt = nullptr;
  int32_t* d_le = nullptr;
  int32_t* d_order = nullptr;  // 2 x n survivor lists + 1 counter, allocated on first use (tp_track)
};

// ---------------------------------------------------------------------------------------
// Batches of lattice variants
// ---------------------------------------------------------------------------------------
struct tp_batch {
  tp_ctx* ctx = nullptr;
  int32_t n_var = 0;
  std::vector<tp_particles*> parts;
  std::vector<int32_t> n_seq, n_slots;
  std::vector<int64_t> seq_off, pos_off, rec_off, pool_off, slot_off;  // pos_off: into slots / gorder
  std::vector<tp_table_header> hdr;
  std::vector<int32_t> gorder;  // per concatenated position: gorder of its record
  unsigned features = 0;
  int variant = TP_VARIANT_ALL;
  int exact_h = 1;
  std::string no_adjoint;  // why the adjoint is unsupported ("" = supported)
  int32_t total_slots = 0, max_slots = 0;
  int32_t* d_seq = nullptr;
  tp_rec* d_recs = nullptr;
  double* d_pool = nullptr;
  int32_t* d_slots = nullptr;
  double* d_minv = nullptr;  // r1/r2 inverses of every variant at the concatenated pool offsets (or null)
  int block = 256, fwd_waves = 0, back_waves = 0, cus = 0;
  // the planned launch (rebuilt when its key changes)
  int mode = 0;  // 0 none, 1 track, 2 adjoint
  uint32_t pflags = 0;
  int32_t pstride = 0;
  int pblock = 0, pfw = 0, pbw = 0;
  std::vector<int32_t> pturn0, pnturns;
  int32_t n_items = 0, grid_f = 0, grid_b = 0;
  int64_t ckpt_bytes = 0;
  std::vector<int64_t> lam_off;
  void *d_desc = nullptr, *d_items = nullptr, *d_counter = nullptr, *d_vfirst = nullptr, *d_vtiles = nullptr;
  void *d_ckpt = nullptr, *d_item_loss = nullptr, *d_item_partial = nullptr, *d_grad = nullptr, *d_vloss = nullptr;
  void* d_lam = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  uint64_t hash = 0;  // of the concatenated tables and the particle counts (autotune cache key)
};

#define TP_HIP(call)                                                                       \
  do {                                                                                     \
    hipError_t e_ = (call);                                                                \
    if (e_ != hipSuccess) return fail(TP_EHIP, "%s failed: %s (%s)", #call, hipGetErrorName(e_), \
                                      hipGetErrorString(e_));                             \
  } while (0)

namespace tpr {

// Drift form of a fast-build tracking call (TpKArgs::ser): forced by tp_set_drift, or (auto)
// the series when the particle set -- its size n, not its survivors -- fills at least half of
// the resident wave slots of the series kernel.  Fixed per particle set, so every launch of a
// call, and every call on the set, takes the same form.
int drift_form(const tp_ctx* ctx, int variant, int exact_h, int64_t n, bool fast);

int set_dev(tp_ctx* ctx);
void ctx_release(tp_ctx* ctx);
void particles_release(tp_particles* p);
// Builds (or keeps, if its key is unchanged) the launch plan and graph of a batch;
// mode 1 = track, 2 = adjoint.
int batch_plan(tp_batch* b, int mode, const int32_t* turn0, const int32_t* nturns, int32_t stride, uint32_t flags);
// Launches the planned graph on the context's stream (TP_TRACK_TIMED: events around it).
int batch_launch(tp_batch* b, uint32_t flags);
// Validates the arguments of a batched call and makes the context's device current.
int batch_check(tp_ctx* ctx, tp_batch* b, const int32_t* turn0, const int32_t* nturns, int32_t min_turns,
                const char* who);
// 64-bit FNV-1a of n bytes, continuing from h (start with FNV_OFFSET).
constexpr uint64_t FNV_OFFSET = 1469598103934665603ull;
uint64_t fnv1a(uint64_t h, const void* p, size_t n);

}  // namespace tpr

#endif  // TP_RUNTIME_H
