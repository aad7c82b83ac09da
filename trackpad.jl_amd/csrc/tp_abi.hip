// tp_abi.hip -- the C ABI of libtrackpad_hip (declared in include/trackpad_hip.h).
//
// Host-side runtime: contexts (device + stream), table validation and upload, device
// particle sets, turn chunking and kernel dispatch.  No exception crosses the ABI; every
// failure returns a TP_E* code and leaves a message in tp_last_error().
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include <mutex>

#include "tp_runtime.h"

// Scratch of tp_track_noalloc: word 0 is the persistent launches' queue counter, words
// 64 .. 64 + 2 TP_NOALLOC_MAX_N the two survivor lists and their count.  One copy per device,
// held until the process ends (8 MiB).  Host code never names it directly: its device address
// comes from hipGetSymbolAddress (scratch_ptr below).
static __device__ __align__(256) int32_t scratch[64 + 2 * (size_t)TP_NOALLOC_MAX_N + 64];

namespace tpr {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}


uint64_t fnv1a(uint64_t h, const void* p, size_t n) {
  const unsigned char* c = static_cast<const unsigned char*>(p);
  for (size_t i = 0; i < n; ++i) {
    h ^= c[i];
    h *= 1099511628211ull;
  }
  return h;
}

int drift_form(const tp_ctx* ctx, int variant, int exact_h, int64_t n, bool fast) {
  if (!fast || ctx->drift == TP_DRIFT_RSQRT) return 0;
  if (ctx->drift == TP_DRIFT_SERIES) return 1;
  // measured on the DBA ring: the series is +6 % at 1e6 particles and -17 % at 4,096
  // (64 waves, each alone on its SIMD, pay its per-drift lane test in full)
  const int occ = tp_track_occupancy_fast(variant, exact_h, ctx->block, ctx->ppl, ctx->fwd_waves, 1);
  const int64_t tiles = (n + 64 * ctx->ppl - 1) / (64 * ctx->ppl);
  return (occ > 0 && 2 * tiles >= (int64_t)occ * ctx->cus * (ctx->block / 64)) ? 1 : 0;
}

int set_dev(tp_ctx* ctx) {
  TP_HIP(hipSetDevice(ctx->device));
  return TP_OK;
}

// Device address of `scratch` on `device` (which must be current), looked up once per device
// and kept: each device holds its own copy of the array at its own address.
int scratch_ptr(int device, int32_t** out) {
  static std::mutex mu;
  static std::vector<int32_t*> addr;
  std::lock_guard<std::mutex> lock(mu);
  if ((size_t)device >= addr.size()) addr.resize((size_t)device + 1, nullptr);
  if (!addr[(size_t)device]) {
    void* ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return fail(TP_EHIP, "hipGetSymbolAddress(scratch) on device %d failed: %s", device, hipGetErrorName(e));
    addr[(size_t)device] = static_cast<int32_t*>(ptr);
  }
  *out = addr[(size_t)device];
  return TP_OK;
}

void ctx_release(tp_ctx* ctx) {
  if (ctx->refs.fetch_sub(1) != 1) return;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  (void)hipEventDestroy(ctx->ev0);
  (void)hipEventDestroy(ctx->ev1);
  (void)hipStreamDestroy(ctx->stream);
  (void)hipFree(ctx->d_counter);
  delete ctx;
}

void particles_release(tp_particles* p) {
  if (p->refs.fetch_sub(1) != 1) return;
  tp_ctx* ctx = p->ctx;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  (void)hipFree(p->d_coords);
  (void)hipFree(p->d_lt);
  (void)hipFree(p->d_le);
  (void)hipFree(p->d_order);
  delete p;
  ctx_release(ctx);
}

}  // namespace tpr

using namespace tpr;

extern "C" {

int tp_version(void) { return TP_ABI_VERSION; }

const char* tp_last_error(void) { return g_err.c_str(); }

int tp_device_count(int* n) {
  if (!n) return fail(TP_EINVAL, "tp_device_count: n is NULL");
  TP_HIP(hipGetDeviceCount(n));
  return TP_OK;
}

int tp_ctx_create(int device, tp_ctx** out) {
  if (!out) return fail(TP_EINVAL, "tp_ctx_create: out is NULL");
  *out = nullptr;
  int nd = 0;
  TP_HIP(hipGetDeviceCount(&nd));
  if (device < 0 || device >= nd) return fail(TP_EINVAL, "tp_ctx_create: device %d out of range [0, %d)", device, nd);
  TP_HIP(hipSetDevice(device));
  tp_ctx* c = new (std::nothrow) tp_ctx();
  if (!c) return fail(TP_ENOMEM, "tp_ctx_create: out of host memory");
  c->device = device;
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&c->ev0);
  if (e == hipSuccess) e = hipEventCreate(&c->ev1);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&c->cus, hipDeviceAttributeMultiprocessorCount, device);
  if (e == hipSuccess) e = hipMalloc(&c->d_counter, 256);
  if (e != hipSuccess) {
    if (c->d_counter) (void)hipFree(c->d_counter);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return fail(TP_EHIP, "tp_ctx_create: %s", hipGetErrorName(e));
  }
  // the static scratch of tp_track_noalloc, looked up once per device at set-up; a failure
  // here leaves it unset and tp_track_noalloc reports it (the other entry points do not need it)
  if (scratch_ptr(device, &c->d_scratch) != TP_OK) c->d_scratch = nullptr;
  *out = c;
  return TP_OK;
}

int tp_ctx_destroy(tp_ctx* ctx) {
  if (!ctx) return TP_OK;
  ctx->closed = true;  // a second destroy of the same handle is undefined, as for free()
  ctx_release(ctx);
  return TP_OK;
}

int tp_set_launch(tp_ctx* ctx, int block, int ppl, int turns_per_launch) {
  if (!ctx || ctx->closed) return fail(TP_EINVAL, "tp_set_launch: ctx is NULL or destroyed");
  if (block == 0) block = 256;
  if (ppl == 0) ppl = 1;
  if (block < 64 || block > 256 || block % 64 != 0)
    return fail(TP_EINVAL, "tp_set_launch: block %d not in {64, 128, 192, 256}", block);
  if (ppl != 1 && ppl != 2) return fail(TP_EINVAL, "tp_set_launch: ppl %d not 1 or 2", ppl);
  if (turns_per_launch < 0) return fail(TP_EINVAL, "tp_set_launch: turns_per_launch %d < 0", turns_per_launch);
  ctx->block = block;
  ctx->ppl = ppl;
  ctx->tpl = turns_per_launch;
  return TP_OK;
}

int tp_set_waves(tp_ctx* ctx, int fwd_waves, int back_waves) {
  if (!ctx || ctx->closed) return fail(TP_EINVAL, "tp_set_waves: ctx is NULL or destroyed");
  if (fwd_waves != 0 && fwd_waves != 5 && fwd_waves != 6)
    return fail(TP_EINVAL, "tp_set_waves: fwd_waves %d not in {0, 5, 6}", fwd_waves);
  if (back_waves != 0 && back_waves != 3 && back_waves != 4)
    return fail(TP_EINVAL, "tp_set_waves: back_waves %d not in {0, 3, 4}", back_waves);
  ctx->fwd_waves = fwd_waves;
  ctx->back_waves = back_waves;
  return TP_OK;
}

int tp_set_drift(tp_ctx* ctx, int form) {
  if (!ctx || ctx->closed) return fail(TP_EINVAL, "tp_set_drift: ctx is NULL or destroyed");
  if (form != TP_DRIFT_AUTO && form != TP_DRIFT_SERIES && form != TP_DRIFT_RSQRT)
    return fail(TP_EINVAL, "tp_set_drift: form %d not TP_DRIFT_AUTO, TP_DRIFT_SERIES or TP_DRIFT_RSQRT", form);
  ctx->drift = form;
  return TP_OK;
}

int tp_timing(tp_ctx* ctx, tp_timing_t* out) {
  if (!ctx || !out) return fail(TP_EINVAL, "tp_timing: NULL argument");
  *out = ctx->timing;
  return TP_OK;
}

}  // extern "C"


extern "C" {

int tp_table_check(const void* bytes, size_t nbytes, uint32_t* features) {
  tp_table_host t;
  int rc = parse_table(bytes, nbytes, &t);
  if (rc) return rc;
  if (features) *features = t.features;
  return TP_OK;
}

int tp_table_create(tp_ctx* ctx, const void* bytes, size_t nbytes, tp_table** out) {
  if (!ctx || !bytes || !out) return fail(TP_EINVAL, "tp_table_create: NULL argument");
  if (ctx->closed) return fail(TP_EINVAL, "tp_table_create: context destroyed");
  *out = nullptr;
  tp_table* t = new (std::nothrow) tp_table();
  if (!t) return fail(TP_ENOMEM, "tp_table_create: out of host memory");
  int rc = parse_table(bytes, nbytes, t);
  if (rc) {
    delete t;
    return rc;
  }
  t->ctx = ctx;
  const tp_table_header& h = t->hdr;
  const size_t seq_bytes = (t->seq.size() * 4 + 7) / 8 * 8;
  // Device copy: the same layout with the library's kick field written.
  const char* b = (const char*)bytes;
  std::vector<char> blob(b, b + nbytes);
  std::memcpy(blob.data() + sizeof h + seq_bytes, t->recs.data(), (size_t)h.n_rec * sizeof(tp_rec));
  hipError_t e = hipSetDevice(ctx->device);
  if (e == hipSuccess) e = hipMalloc(&t->dev, nbytes);
  if (e == hipSuccess) e = hipMemcpyAsync(t->dev, blob.data(), nbytes, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) {
    if (t->dev) (void)hipFree(t->dev);
    delete t;
    return fail(e == hipErrorOutOfMemory ? TP_ENOMEM : TP_EHIP, "tp_table_create: %s", hipGetErrorName(e));
  }
  t->nbytes = nbytes;
  ctx->refs.fetch_add(1);
  const char* d = (const char*)t->dev;
  t->d_seq = (const int32_t*)(d + sizeof h);
  t->d_recs = (const tp_rec*)(d + sizeof h + seq_bytes);
  t->d_pool = (const double*)(d + sizeof h + seq_bytes + (size_t)h.n_rec * sizeof(tp_rec));
  *out = t;
  return TP_OK;
}

int tp_table_destroy(tp_table* t) {
  if (!t) return TP_OK;
  tp_ctx* ctx = t->ctx;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  if (t->dev) (void)hipFree(t->dev);
  if (t->d_slots) (void)hipFree(t->d_slots);
  if (t->d_minv) (void)hipFree(t->d_minv);
  delete t;
  ctx_release(ctx);
  return TP_OK;
}

int tp_table_info(const tp_table* t, int32_t* n_seq, int32_t* n_rec, double* beti) {
  if (!t) return fail(TP_EINVAL, "tp_table_info: table is NULL");
  if (n_seq) *n_seq = (int32_t)t->hdr.n_seq;
  if (n_rec) *n_rec = (int32_t)t->hdr.n_rec;
  if (beti) *beti = t->hdr.beti;
  return TP_OK;
}

int tp_particles_create(tp_ctx* ctx, int64_t n, tp_particles** out) {
  if (!ctx || !out) return fail(TP_EINVAL, "tp_particles_create: NULL argument");
  if (ctx->closed) return fail(TP_EINVAL, "tp_particles_create: context destroyed");
  *out = nullptr;
  if (n < 0 || n > ((int64_t)1 << 40)) return fail(TP_EINVAL, "tp_particles_create: n = %lld", (long long)n);
  int rc = set_dev(ctx);
  if (rc) return rc;
  tp_particles* p = new (std::nothrow) tp_particles();
  if (!p) return fail(TP_ENOMEM, "tp_particles_create: out of host memory");
  p->ctx = ctx;
  p->n = n;
  const size_t m = (size_t)(n > 0 ? n : 1);
  hipError_t e = hipMalloc(&p->d_coords, 6 * m * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&p->d_lt, m * sizeof(int32_t));
  if (e == hipSuccess) e = hipMalloc(&p->d_le, m * sizeof(int32_t));
  if (e != hipSuccess) {
    if (p->d_coords) (void)hipFree(p->d_coords);
    if (p->d_lt) (void)hipFree(p->d_lt);
    delete p;
    return fail(e == hipErrorOutOfMemory ? TP_ENOMEM : TP_EHIP, "tp_particles_create(%lld): %s", (long long)n,
                hipGetErrorName(e));
  }
  ctx->refs.fetch_add(1);
  *out = p;
  return TP_OK;
}

int tp_particles_destroy(tp_particles* p) {
  if (!p) return TP_OK;
  particles_release(p);
  return TP_OK;
}

int tp_particles_upload(tp_particles* p, const double* coords, const int64_t* lost_flags01,
                        const int32_t* lost_turn, const int32_t* lost_elem) {
  if (!p || !coords) return fail(TP_EINVAL, "tp_particles_upload: NULL argument");
  int rc = set_dev(p->ctx);
  if (rc) return rc;
  const int64_t n = p->n;
  if (n == 0) return TP_OK;
  std::vector<int32_t> lt((size_t)n), le((size_t)n);
  for (int64_t i = 0; i < n; ++i) {
    int32_t t = -1, e = -1;
    if (lost_turn) {
      t = lost_turn[i];
      e = lost_elem ? lost_elem[i] : (t == -1 ? -1 : -2);
    } else if (lost_flags01 && lost_flags01[i] == 1) {
      t = -2;
      e = -2;
    }
    lt[(size_t)i] = t;
    le[(size_t)i] = e;
  }
  hipStream_t s = p->ctx->stream;
  TP_HIP(hipMemcpyAsync(p->d_coords, coords, 6 * (size_t)n * sizeof(double), hipMemcpyHostToDevice, s));
  TP_HIP(hipMemcpyAsync(p->d_lt, lt.data(), (size_t)n * sizeof(int32_t), hipMemcpyHostToDevice, s));
  TP_HIP(hipMemcpyAsync(p->d_le, le.data(), (size_t)n * sizeof(int32_t), hipMemcpyHostToDevice, s));
  TP_HIP(hipStreamSynchronize(s));
  return TP_OK;
}

int tp_particles_download(tp_particles* p, double* coords, int64_t* lost_flags01, int32_t* lost_turn,
                          int32_t* lost_elem) {
  if (!p) return fail(TP_EINVAL, "tp_particles_download: particles is NULL");
  int rc = set_dev(p->ctx);
  if (rc) return rc;
  const int64_t n = p->n;
  if (n == 0) return TP_OK;
  hipStream_t s = p->ctx->stream;
  std::vector<int32_t> lt;
  int32_t* ltp = lost_turn;
  if (lost_flags01 && !ltp) {
    lt.resize((size_t)n);
    ltp = lt.data();
  }
  if (coords) TP_HIP(hipMemcpyAsync(coords, p->d_coords, 6 * (size_t)n * sizeof(double), hipMemcpyDeviceToHost, s));
  if (ltp) TP_HIP(hipMemcpyAsync(ltp, p->d_lt, (size_t)n * sizeof(int32_t), hipMemcpyDeviceToHost, s));
  if (lost_elem) TP_HIP(hipMemcpyAsync(lost_elem, p->d_le, (size_t)n * sizeof(int32_t), hipMemcpyDeviceToHost, s));
  TP_HIP(hipStreamSynchronize(s));
  if (lost_flags01)
    for (int64_t i = 0; i < n; ++i) lost_flags01[i] = (ltp[i] != -1) ? 1 : 0;
  return TP_OK;
}

int tp_particles_copy(tp_particles* dst, const tp_particles* src) {
  if (!dst || !src) return fail(TP_EINVAL, "tp_particles_copy: NULL argument");
  if (dst->n != src->n) return fail(TP_EINVAL, "tp_particles_copy: sizes %lld and %lld differ", (long long)dst->n,
                                    (long long)src->n);
  if (dst->ctx->device != src->ctx->device) return fail(TP_EINVAL, "tp_particles_copy: different devices");
  int rc = set_dev(dst->ctx);
  if (rc) return rc;
  const size_t n = (size_t)dst->n;
  if (n == 0) return TP_OK;
  hipStream_t s = dst->ctx->stream;
  TP_HIP(hipMemcpyAsync(dst->d_coords, src->d_coords, 6 * n * sizeof(double), hipMemcpyDeviceToDevice, s));
  TP_HIP(hipMemcpyAsync(dst->d_lt, src->d_lt, n * sizeof(int32_t), hipMemcpyDeviceToDevice, s));
  TP_HIP(hipMemcpyAsync(dst->d_le, src->d_le, n * sizeof(int32_t), hipMemcpyDeviceToDevice, s));
  return TP_OK;
}

int tp_particles_device_ptrs(tp_particles* p, double** coords, int32_t** lost_turn, int32_t** lost_elem,
                             int64_t* n) {
  if (!p) return fail(TP_EINVAL, "tp_particles_device_ptrs: particles is NULL");
  if (coords) *coords = p->d_coords;
  if (lost_turn) *lost_turn = p->d_lt;
  if (lost_elem) *lost_elem = p->d_le;
  if (n) *n = p->n;
  return TP_OK;
}

}  // extern "C"

namespace {

// tp_track and tp_track_noalloc.  lists: 2 n + 1 int32 of device memory for the survivor
// lists and their count, or null to use (and on first need allocate) the particle set's own;
// counter: the persistent launches' queue counter, or null for the context's.
int track_impl(tp_ctx* ctx, const tp_table* t, tp_particles* p, int32_t turn0, int32_t nturns, uint32_t flags,
               int32_t* lists_buf, int32_t* counter, const char* who) {
  if (!ctx || !t || !p) return fail(TP_EINVAL, "%s: NULL argument", who);
  if (t->ctx != ctx || p->ctx != ctx) return fail(TP_EINVAL, "%s: table and particles belong to another context", who);
  if (ctx->closed) return fail(TP_EINVAL, "%s: context destroyed", who);
  if (nturns < 0) return fail(TP_EINVAL, "%s: nturns = %d", who, nturns);
  if (flags & ~(TP_TRACK_FAST | TP_TRACK_ASYNC | TP_TRACK_TIMED)) return fail(TP_EINVAL, "%s: unknown flags 0x%x", who, flags);
  if ((int64_t)turn0 + (int64_t)nturns > 0x7fffffffLL) return fail(TP_EINVAL, "%s: turn range overflows int32", who);
  int rc = set_dev(ctx);
  if (rc) return rc;
  TpKArgs a{};
  a.seq = t->d_seq;
  a.recs = t->d_recs;
  a.pool = t->d_pool;
  a.n_seq = (int32_t)t->hdr.n_seq;
  a.exact_h = (t->hdr.flags & TP_H_EXACT_HAMILTONIAN) ? 1 : 0;
  a.beti = t->hdr.beti;
  a.coord_limit = t->hdr.coord_limit;
  a.angle_limit = t->hdr.angle_limit;
  a.ap_r2 = t->hdr.aperture_r2;
  a.coords = p->d_coords;
  a.n = p->n;
  a.ld = p->n;
  a.lost_turn = p->d_lt;
  a.lost_elem = p->d_le;
  a.variant = t->variant;
  a.waves = ctx->fwd_waves;
  a.n_rows = t->hdr.n_rows > 1 ? (int32_t)t->hdr.n_rows : 1;
  a.row_turn0 = t->hdr.row_turn0;
  if (a.n_rows > 1 && nturns > 0 &&
      ((int64_t)turn0 < a.row_turn0 || (int64_t)turn0 + nturns > (int64_t)a.row_turn0 + a.n_rows))
    return fail(TP_EINVAL, "%s: turns [%d, %lld) outside the table's time-varying rows [%d, %lld)", who, turn0,
                (long long)turn0 + nturns, a.row_turn0, (long long)a.row_turn0 + a.n_rows);
  const bool timed = (flags & TP_TRACK_TIMED) != 0;
  const int chunk = ctx->tpl > 0 ? ctx->tpl : (nturns > 0 ? nturns : 1);
  // Before the first launch and between launches the survivors are compacted into a
  // particle list (compact_alive), so lanes of lost particles stop costing issue slots in the
  // following launches.  It needs a count read back per launch, so asynchronous calls and
  // single launches skip it.
  const bool compact = !(flags & TP_TRACK_ASYNC) && nturns > chunk && p->n <= 0x7fffffffLL;
  if (compact && !lists_buf && !p->d_order)
    TP_HIP(hipMalloc(&p->d_order, (size_t)(2 * p->n + 1) * sizeof(int32_t)));
  int32_t* lbuf = lists_buf ? lists_buf : p->d_order;
  int32_t* lists[2] = {lbuf, lbuf ? lbuf + p->n : nullptr};
  int32_t* d_count = lbuf ? lbuf + 2 * p->n : nullptr;
  int32_t* queue = counter ? counter : ctx->d_counter;
  int cur = -1;  // list in use (-1: all particles)
  // A launch with more waves than the device holds at once runs persistent: a resident
  // grid whose waves take 64-particle tiles from a queue (TpKArgs::counter).
  const bool fast = (flags & TP_TRACK_FAST) != 0;
  a.ser = drift_form(ctx, a.variant, a.exact_h, p->n, fast);
  const int occ = fast ? tp_track_occupancy_fast(a.variant, a.exact_h, ctx->block, ctx->ppl, ctx->fwd_waves, a.ser)
                       : tp_track_occupancy_exact(a.variant, a.exact_h, ctx->block, ctx->ppl, ctx->fwd_waves, 0);
  const int max_grid = occ > 0 ? occ * ctx->cus : 0;
  double ms_total = 0.0;
  int launches = 0;
  // With compaction the launches follow the losses.  The first covers a quarter of a turn;
  // while launches lose 8 % of their particles per turn or more (a dynamic-aperture grid
  // loses most of its particles in its first turns) they stay a quarter turn long, so lost
  // particles leave the lanes within the turn; at 2 % per turn or more they are one turn
  // long, below that they double, and once a launch loses under 1 % the rest runs in
  // launches of `chunk` turns (a beam that loses nothing pays one short extra launch).
  int32_t len = compact ? 1 : chunk;
  int32_t parts = compact ? 4 : 1;  // > 1: the next launch covers 1/parts of a turn
  bool any_alive = true;
  if (compact && p->n > 0 && a.n_seq > 0) {
    // particles lost in earlier calls (a scan continued in chunks) leave the list before the
    // first launch
    hipError_t e = tp_launch_fill_i32(d_count, 0, 1, ctx->stream);
    if (e == hipSuccess) e = tp_launch_compact_alive(nullptr, (int32_t)a.n, p->d_lt, lists[0], d_count, ctx->stream);
    if (e != hipSuccess) return fail(TP_EHIP, "%s: compaction launch failed: %s", who, hipGetErrorName(e));
    int32_t alive = 0;
    TP_HIP(hipMemcpyAsync(&alive, d_count, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
    TP_HIP(hipStreamSynchronize(ctx->stream));
    any_alive = alive > 0;
    if (alive > 0 && alive < a.n) {
      cur = 0;
      a.order = lists[cur];
      a.n = alive;
    }
  }
  if (p->n > 0 && a.n_seq > 0 && any_alive) {
    int32_t done = 0, epos = 0;  // whole turns tracked, and elements of turn `done`
    while (done < nturns) {
      a.turn0 = turn0 + done;
      a.e_first = epos;
      int32_t e_end = a.n_seq;  // the launch ends before this element of its last turn
      if (parts > 1) {
        a.nturns = 1;
        const int32_t step = (a.n_seq + parts - 1) / parts;
        e_end = (a.n_seq - epos > step) ? epos + step : a.n_seq;
      } else {
        a.nturns = (nturns - done < len) ? (nturns - done) : len;
      }
      a.e_last = e_end < a.n_seq ? e_end : 0;
      const int64_t blocks = (a.n + (int64_t)ctx->block * ctx->ppl - 1) / ((int64_t)ctx->block * ctx->ppl);
      a.counter = (max_grid > 0 && blocks > max_grid) ? queue : nullptr;
      hipError_t e = hipSuccess;
      if (a.counter) e = tp_launch_fill_i32(a.counter, 0, 1, ctx->stream);
      if (e != hipSuccess) return fail(TP_EHIP, "%s: queue reset failed: %s", who, hipGetErrorName(e));
      if (timed) TP_HIP(hipEventRecord(ctx->ev0, ctx->stream));
      e = fast ? tp_launch_track_fast(a, ctx->block, ctx->ppl, max_grid, ctx->stream)
               : tp_launch_track_exact(a, ctx->block, ctx->ppl, max_grid, ctx->stream);
      if (e != hipSuccess) return fail(TP_EHIP, "%s: kernel launch failed: %s", who, hipGetErrorName(e));
      ++launches;
      if (timed) {
        TP_HIP(hipEventRecord(ctx->ev1, ctx->stream));
        TP_HIP(hipEventSynchronize(ctx->ev1));
        float ms = 0.f;
        TP_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        ms_total += ms;
      }
      // turns the launch covered, then the position after it
      const double covered = (double)(a.nturns - 1) + (double)(e_end - epos) / (double)a.n_seq;
      if (e_end == a.n_seq) {
        done += a.nturns;
        epos = 0;
      } else {
        done += a.nturns - 1;
        epos = e_end;
      }
      if (compact && done < nturns) {
        const int nxt = cur == 1 ? 0 : 1;
        e = tp_launch_fill_i32(d_count, 0, 1, ctx->stream);
        if (e == hipSuccess)
          e = tp_launch_compact_alive(cur < 0 ? nullptr : lists[cur], (int32_t)a.n, p->d_lt, lists[nxt], d_count,
                                      ctx->stream);
        if (e != hipSuccess) return fail(TP_EHIP, "%s: compaction launch failed: %s", who, hipGetErrorName(e));
        int32_t alive = 0;
        TP_HIP(hipMemcpyAsync(&alive, d_count, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
        TP_HIP(hipStreamSynchronize(ctx->stream));
        if (alive == 0) break;  // nothing left to track
        const double lost = (double)(a.n - alive) / (double)a.n;
        const double per_turn = lost / (covered > 0.0 ? covered : 1.0);
        parts = 1;
        if (per_turn >= 0.08) parts = 4;
        else if (per_turn >= 0.02) len = 1;
        else if (lost < 0.01) len = chunk;
        else len = (int32_t)std::min<int64_t>(2 * (int64_t)a.nturns, chunk);
        if (alive < a.n) {      // switch to the shorter list (else keep the current one)
          cur = nxt;
          a.order = lists[cur];
          a.n = alive;
        }
      }
    }
  }
  if (timed) {
    ctx->timing.kernel_ms = ms_total;
    ctx->timing.launches = launches;
    ctx->timing.block = ctx->block;
    ctx->timing.ppl = ctx->ppl;
    ctx->timing.turns_per_launch = chunk;
  }
  if (!(flags & TP_TRACK_ASYNC)) TP_HIP(hipStreamSynchronize(ctx->stream));
  return TP_OK;
}

}  // namespace

extern "C" {

int tp_track(tp_ctx* ctx, const tp_table* t, tp_particles* p, int32_t turn0, int32_t nturns, uint32_t flags) {
  return track_impl(ctx, t, p, turn0, nturns, flags, nullptr, nullptr, "tp_track");
}

int tp_track_noalloc(tp_ctx* ctx, const tp_table* t, tp_particles* p, int32_t turn0, int32_t nturns,
                     uint32_t flags) {
  if (!ctx || !t || !p) return fail(TP_EINVAL, "tp_track_noalloc: NULL argument");
  if (p->n > (int64_t)TP_NOALLOC_MAX_N)
    return fail(TP_EINVAL, "tp_track_noalloc: %lld particles exceed the static scratch maximum of %d",
                (long long)p->n, (int)TP_NOALLOC_MAX_N);
  if (!ctx->d_scratch) {
    const int rc = set_dev(ctx);
    if (rc) return rc;
    if (scratch_ptr(ctx->device, &ctx->d_scratch) != TP_OK) return TP_EHIP;  // message set by scratch_ptr
  }
  // word 0: queue counter; from word 64: list 0 | list 1 | count (each call clears what it
  // uses: fill_i32 before every persistent launch and every compaction)
  return track_impl(ctx, t, p, turn0, nturns, flags, ctx->d_scratch + 64, ctx->d_scratch, "tp_track_noalloc");
}

int tp_sync(tp_ctx* ctx) {
  if (!ctx) return fail(TP_EINVAL, "tp_sync: ctx is NULL");
  int rc = set_dev(ctx);
  if (rc) return rc;
  TP_HIP(hipStreamSynchronize(ctx->stream));
  return TP_OK;
}

int tp_track_host(tp_ctx* ctx, const tp_table* t, double* coords, int64_t n, int64_t* lost_flags, int32_t* lost_turn,
                  int32_t* lost_elem, int32_t turn0, int32_t nturns, uint32_t flags) {
  if (!ctx || !t || (!coords && n > 0) || !lost_flags) return fail(TP_EINVAL, "tp_track_host: NULL argument");
  tp_particles* p = nullptr;
  int rc = tp_particles_create(ctx, n, &p);
  if (rc) return rc;
  std::vector<int32_t> lt((size_t)(n > 0 ? n : 1)), le((size_t)(n > 0 ? n : 1));
  rc = tp_particles_upload(p, coords, lost_flags, nullptr, nullptr);
  if (!rc) rc = tp_track(ctx, t, p, turn0, nturns, flags & ~TP_TRACK_ASYNC);
  if (!rc) rc = tp_particles_download(p, coords, nullptr, lt.data(), le.data());
  tp_particles_destroy(p);
  if (rc) return rc;
  for (int64_t i = 0; i < n; ++i) {
    // reference semantics: flags equal to 1 were skipped; newly lost particles get 1;
    // every other entry is left as it was (src/lattice.jl:269-284).
    if (lt[(size_t)i] >= 0) lost_flags[i] = 1;
    if (lost_turn) lost_turn[i] = lt[(size_t)i];
    if (lost_elem) lost_elem[i] = le[(size_t)i];
  }
  return TP_OK;
}

namespace {

// RAII device buffer for the adjoint's scratch arrays.
struct DevBuf {
  void* p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
};

}  // namespace

}  // extern "C"

namespace {

// Variant of the reverse kernels: the largest one (TP_FEAT_ADJ_ALL) for Forest exact bends,
// the linearised drift, the families of tags 8-22 and r1/r2 rotations; a table whose forward
// variant is ALL for its t1/t2 translations alone runs the C reverse kernel, which undoes
// t1/t2 for every variant at run time.
int adj_variant(unsigned features, int variant, bool rotations) {
  if ((features & (FT_ESBEND | FT_LINH | FT_EXTRA)) || rotations) return TP_VARIANT_ALL;
  return variant > TP_VARIANT_C ? TP_VARIANT_C : variant;
}

bool has_rotations(const std::vector<tp_rec>& recs) {
  for (const tp_rec& r : recs)
    if (r.flags & (TP_F_R1 | TP_F_R2)) return true;
  return false;
}

// Inverse of a 6 x 6 column-major matrix by Gauss-Jordan elimination with partial pivoting in
// long double; false if a pivot vanishes.
bool invert6(const double* M, double* out) {
  long double a[6][12];
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) {
      a[i][j] = M[i + 6 * j];
      a[i][6 + j] = (i == j) ? 1.0L : 0.0L;
    }
  for (int c = 0; c < 6; ++c) {
    int piv = c;
    for (int i = c + 1; i < 6; ++i)
      if (fabsl(a[i][c]) > fabsl(a[piv][c])) piv = i;
    if (!(fabsl(a[piv][c]) > 0.0L)) return false;
    if (piv != c)
      for (int j = 0; j < 12; ++j) std::swap(a[c][j], a[piv][j]);
    const long double d = a[c][c];
    for (int j = 0; j < 12; ++j) a[c][j] /= d;
    for (int i = 0; i < 6; ++i) {
      if (i == c) continue;
      const long double f = a[i][c];
      if (f != 0.0L)
        for (int j = 0; j < 12; ++j) a[i][j] -= f * a[c][j];
    }
  }
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) out[i + 6 * j] = (double)a[i][6 + j];
  return true;
}

// The reverse pass undoes r' = R r with R^-1: minv (shaped like the pool, zero elsewhere)
// receives the inverses of the r1 / r2 matrices of recs, whose pool starts at pool[base],
// at the same offsets.
bool rotation_inverses(const std::vector<tp_rec>& recs, const std::vector<double>& pool, size_t base,
                       std::vector<double>& minv) {
  for (const tp_rec& r : recs) {
    if (r.flags & TP_F_R1)
      if (!invert6(pool.data() + base + r.moff + 12, minv.data() + base + r.moff + 12)) return false;
    if (r.flags & TP_F_R2)
      if (!invert6(pool.data() + base + r.moff + 48, minv.data() + base + r.moff + 48)) return false;
  }
  return true;
}

// The device copy of a table's rotation inverses, built on the first reverse-mode call.
int ensure_minv(tp_ctx* ctx, const tp_table* t, const char* who) {
  tp_table* tm = const_cast<tp_table*>(t);
  if (tm->d_minv || !has_rotations(t->recs)) return TP_OK;
  std::vector<double> minv(t->pool.size(), 0.0);
  if (!rotation_inverses(t->recs, t->pool, 0, minv))
    return fail(TP_EUNSUPPORTED, "%s: a singular r1/r2 matrix has no inverse", who);
  DevBuf buf;
  TP_HIP(hipMalloc(&buf.p, minv.size() * sizeof(double)));
  TP_HIP(hipMemcpyAsync(buf.p, minv.data(), minv.size() * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
  TP_HIP(hipStreamSynchronize(ctx->stream));
  tm->d_minv = (double*)buf.p;  // kept by the table from here on
  buf.p = nullptr;
  return TP_OK;
}

// A time-varying table's strength slots are those of row 0 (per position); the reverse pass
// adds each row's kick derivative to its position's slot, i.e. the gradient is that of an
// offset added to the element's strength at every turn.  That needs the main polynomial
// order of every position to be the same in every row.
bool rows_share_slots(const tp_table* t) {
  const size_t n_seq = t->hdr.n_seq, rows = t->hdr.n_rows > 1 ? t->hdr.n_rows : 1;
  for (size_t r = 1; r < rows; ++r)
    for (size_t e = 0; e < n_seq; ++e)
      if (t->recs[t->seq[r * n_seq + e]].gorder != t->recs[t->seq[e]].gorder) return false;
  return true;
}

// Checks shared by the reverse-mode entry points.
int adj_check(tp_ctx* ctx, const tp_table* t, tp_particles* p, int32_t turn0, int32_t nturns, int32_t ckpt_stride,
              uint32_t flags, const char* who) {
  if (!ctx || !t || !p) return fail(TP_EINVAL, "%s: NULL argument", who);
  if (t->ctx != ctx || p->ctx != ctx) return fail(TP_EINVAL, "%s: table and particles belong to another context", who);
  if (ctx->closed) return fail(TP_EINVAL, "%s: context destroyed", who);
  if (nturns < 1) return fail(TP_EINVAL, "%s: nturns = %d", who, nturns);
  if (ckpt_stride < 0) return fail(TP_EINVAL, "%s: ckpt_stride = %d", who, ckpt_stride);
  if (flags & ~(TP_TRACK_FAST | TP_TRACK_TIMED)) return fail(TP_EINVAL, "%s: unsupported flags 0x%x", who, flags);
  for (const tp_rec& r : t->recs) {
    if (r.tag == TP_ESBEND && (r.flags & TP_F_DRIFTLIMIT))
      return fail(TP_EUNSUPPORTED, "%s: exact bends with |angle/L| < 1e-6 (drift limit) have no adjoint here", who);
  }
  if (t->hdr.n_rows > 1) {
    const int64_t r0 = t->hdr.row_turn0, r1 = r0 + (int64_t)t->hdr.n_rows;
    if ((int64_t)turn0 < r0 || (int64_t)turn0 + nturns > r1)
      return fail(TP_EINVAL, "%s: turns [%d, %lld) outside the table's time-varying rows [%lld, %lld)", who, turn0,
                  (long long)turn0 + nturns, (long long)r0, (long long)r1);
    if (!rows_share_slots(t))
      return fail(TP_EUNSUPPORTED, "%s: a position changes its main multipole order between rows", who);
  }
  return set_dev(ctx);
}

// Forward pass of the reverse-mode entry points: tracks p segment by segment (K turns) and
// stores the beam at the end of every segment into ckpt ([nseg][6][n] doubles).
int adj_forward(tp_ctx* ctx, const tp_table* t, tp_particles* p, int32_t turn0, int32_t nturns, int32_t K, bool fast,
                double* ckpt) {
  const int32_t nseg = (nturns + K - 1) / K;
  const size_t plane = (size_t)p->n;
  for (int32_t j = 0; j < nseg; ++j) {
    const int32_t kt = (j == nseg - 1) ? (nturns - j * K) : K;
    int rc = tp_track(ctx, t, p, turn0 + j * K, kt, (fast ? TP_TRACK_FAST : 0) | TP_TRACK_ASYNC);
    if (rc) return rc;
    if (plane > 0)
      TP_HIP(hipMemcpyAsync(ckpt + (size_t)j * 6 * plane, p->d_coords, 6 * plane * sizeof(double),
                            hipMemcpyDeviceToDevice, ctx->stream));
  }
  return TP_OK;
}

}  // namespace

extern "C" {

int tp_adjoint(tp_ctx* ctx, const tp_table* t, tp_particles* p, int32_t turn0, int32_t nturns, int32_t ckpt_stride,
               uint32_t flags, double* loss, int64_t* n_alive, double* dL_dk, double* dL_dx0) {
  int rc = adj_check(ctx, t, p, turn0, nturns, ckpt_stride, flags, "tp_adjoint");
  if (rc) return rc;
  rc = ensure_minv(ctx, t, "tp_adjoint");
  if (rc) return rc;
  tp_table* tm = const_cast<tp_table*>(t);
  if (!tm->d_slots) {
    std::vector<int32_t> slots(t->hdr.n_seq > 0 ? t->hdr.n_seq : 1, -1);
    int32_t ns = 0;
    for (uint32_t e = 0; e < t->hdr.n_seq; ++e)
      if (t->recs[t->seq[e]].gorder >= 0) slots[e] = ns++;
    DevBuf buf;
    TP_HIP(hipMalloc(&buf.p, slots.size() * sizeof(int32_t)));
    TP_HIP(hipMemcpyAsync(buf.p, slots.data(), slots.size() * sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream));
    TP_HIP(hipStreamSynchronize(ctx->stream));
    tm->d_slots = (int32_t*)buf.p;  // kept by the table from here on
    buf.p = nullptr;
    tm->n_slots = ns;
  }
  const int32_t n_slots = tm->n_slots;
  // LDS of the reverse kernel: a row of slots per wave and the waves' edge memo (128 doubles)
  if ((size_t)4 * ((size_t)n_slots + 128) * sizeof(double) > 64 * 1024)
    return fail(TP_EUNSUPPORTED, "tp_adjoint: %d strength slots exceed the 64 KiB LDS budget (max 1920)", n_slots);
  const int64_t n = p->n;
  const int32_t K = ckpt_stride > 0 ? ckpt_stride : 1;
  const int32_t nseg = (nturns + K - 1) / K;
  const size_t plane = (size_t)(n > 0 ? n : 1);
  const int nblocks = (int)((n + 255) / 256);
  DevBuf ckpt, lam, partial, grad;
  hipError_t e = hipMalloc(&ckpt.p, (size_t)nseg * 6 * plane * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&lam.p, 6 * plane * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&partial.p, (size_t)(nblocks > 0 ? nblocks : 1) * (n_slots > 0 ? n_slots : 1) * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&grad.p, (size_t)(n_slots > 0 ? n_slots : 1) * sizeof(double));
  if (e != hipSuccess)
    return fail(e == hipErrorOutOfMemory ? TP_ENOMEM : TP_EHIP, "tp_adjoint: checkpoint allocation (%d x %lld particles): %s",
                nseg, (long long)n, hipGetErrorName(e));
  const bool fast = (flags & TP_TRACK_FAST) != 0;
  hipStream_t st = ctx->stream;
  TP_HIP(hipEventRecord(ctx->ev0, st));
  // forward pass, storing the beam at the end of every segment
  rc = adj_forward(ctx, t, p, turn0, nturns, K, fast, (double*)ckpt.p);
  if (rc) return rc;
  TP_HIP(hipEventRecord(ctx->ev1, st));
  TP_HIP(hipEventSynchronize(ctx->ev1));
  float fwd_ms = 0.f;
  TP_HIP(hipEventElapsedTime(&fwd_ms, ctx->ev0, ctx->ev1));
  // survivors and loss (host side, from the final beam)
  std::vector<int32_t> lt((size_t)plane);
  std::vector<double> fin(6 * plane);
  if (n > 0) {
    TP_HIP(hipMemcpyAsync(lt.data(), p->d_lt, (size_t)n * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    TP_HIP(hipMemcpyAsync(fin.data(), p->d_coords, 6 * (size_t)n * sizeof(double), hipMemcpyDeviceToHost, st));
    TP_HIP(hipStreamSynchronize(st));
  }
  int64_t na = 0;
  double lsum = 0.0;
  for (int64_t i = 0; i < n; ++i)
    if (lt[(size_t)i] == -1) {
      ++na;
      lsum += fin[(size_t)i] * fin[(size_t)i] + fin[(size_t)(2 * n + i)] * fin[(size_t)(2 * n + i)];
    }
  if (loss) *loss = na > 0 ? lsum / (double)na : 0.0;
  if (n_alive) *n_alive = na;
  // reverse pass
  TP_HIP(hipEventRecord(ctx->ev0, st));
  if (n > 0 && n_slots > 0)
    TP_HIP(hipMemsetAsync(partial.p, 0, (size_t)nblocks * n_slots * sizeof(double), st));
  if (n > 0) TP_HIP(hipMemsetAsync(lam.p, 0, 6 * plane * sizeof(double), st));
  TpAdjArgs a{};
  a.seq = t->d_seq;
  a.recs = t->d_recs;
  a.pool = t->d_pool;
  a.slot_of_pos = t->d_slots;
  a.lam = (double*)lam.p;
  a.lost_turn = p->d_lt;
  a.partial = (double*)partial.p;
  a.n = n;
  a.ld = n;
  a.n_seq = (int32_t)t->hdr.n_seq;
  a.n_slots = n_slots;
  a.variant = adj_variant(t->features, t->variant, has_rotations(t->recs));
  a.exact_h = (t->hdr.flags & TP_H_EXACT_HAMILTONIAN) ? 1 : 0;
  a.minv = t->d_minv;
  a.waves = ctx->back_waves;
  a.n_rows = t->hdr.n_rows > 1 ? (int32_t)t->hdr.n_rows : 1;
  a.row_turn0 = t->hdr.row_turn0;
  a.beti = t->hdr.beti;
  // the reverse pass seeds dL/dx = 2x (the gradient of the loss sum); the 1/N_alive of the
  // mean is applied to the results below
  const double inv_na = na > 0 ? 1.0 / (double)na : 0.0;
  if (n > 0 && na > 0) {
    for (int32_t j = nseg - 1; j >= 0; --j) {
      a.end_state = (const double*)ckpt.p + (size_t)j * 6 * plane;
      a.nturns = (j == nseg - 1) ? (nturns - j * K) : K;
      a.turn_last = turn0 + j * K + a.nturns - 1;
      a.init_lambda = (j == nseg - 1) ? 1 : 0;
      hipError_t le = fast ? tp_launch_adjoint_fast(a, st) : tp_launch_adjoint_exact(a, st);
      if (le != hipSuccess) return fail(TP_EHIP, "tp_adjoint: reverse kernel launch failed: %s", hipGetErrorName(le));
    }
    hipError_t le = tp_launch_grad_reduce((const double*)partial.p, nblocks, n_slots, (double*)grad.p, st);
    if (le != hipSuccess) return fail(TP_EHIP, "tp_adjoint: reduction launch failed: %s", hipGetErrorName(le));
  } else if (n_slots > 0) {
    TP_HIP(hipMemsetAsync(grad.p, 0, (size_t)n_slots * sizeof(double), st));
  }
  TP_HIP(hipEventRecord(ctx->ev1, st));
  TP_HIP(hipEventSynchronize(ctx->ev1));
  float bwd_ms = 0.f;
  TP_HIP(hipEventElapsedTime(&bwd_ms, ctx->ev0, ctx->ev1));
  if (flags & TP_TRACK_TIMED) {
    ctx->timing.forward_ms = fwd_ms;
    ctx->timing.backward_ms = bwd_ms;
  }
  if (dL_dk) {
    std::vector<double> g((size_t)(n_slots > 0 ? n_slots : 1), 0.0);
    if (n_slots > 0) TP_HIP(hipMemcpy(g.data(), grad.p, (size_t)n_slots * sizeof(double), hipMemcpyDeviceToHost));
    int32_t k = 0;
    for (uint32_t e2 = 0; e2 < t->hdr.n_seq; ++e2) {
      const tp_rec& r = t->recs[t->seq[e2]];
      if (r.gorder >= 0) {
        // d b_g / d k: b1 = k1, b2 = k2/2, b3 = k3/6 (src/tracking.jl:395, 446, 478)
        const double sc = r.gorder == 1 ? 1.0 : (r.gorder == 2 ? 0.5 : (r.gorder == 3 ? 1.0 / 6.0 : 1.0));
        dL_dk[e2] = g[(size_t)k++] * inv_na * sc;
      } else {
        dL_dk[e2] = 0.0;
      }
    }
  }
  if (dL_dx0 && n > 0) {
    TP_HIP(hipMemcpy(dL_dx0, lam.p, 6 * (size_t)n * sizeof(double), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < 6 * (size_t)n; ++i) dL_dx0[i] *= inv_na;
  }
  return TP_OK;
}

int tp_jacobian(tp_ctx* ctx, const tp_table* t, tp_particles* p, int32_t turn0, int32_t nturns, int32_t ckpt_stride,
                uint32_t flags, double* jac) {
  int rc = adj_check(ctx, t, p, turn0, nturns, ckpt_stride, flags, "tp_jacobian");
  if (rc) return rc;
  rc = ensure_minv(ctx, t, "tp_jacobian");
  if (rc) return rc;
  const int64_t n = p->n;
  const int32_t K = ckpt_stride > 0 ? ckpt_stride : 1;
  const int32_t nseg = (nturns + K - 1) / K;
  const size_t plane = (size_t)(n > 0 ? n : 1);
  DevBuf ckpt, lam, partial;
  hipError_t e = hipMalloc(&ckpt.p, (size_t)nseg * 6 * plane * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&lam.p, 6 * plane * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&partial.p, sizeof(double));
  if (e != hipSuccess)
    return fail(e == hipErrorOutOfMemory ? TP_ENOMEM : TP_EHIP, "tp_jacobian: checkpoint allocation (%d x %lld particles): %s",
                nseg, (long long)n, hipGetErrorName(e));
  const bool fast = (flags & TP_TRACK_FAST) != 0;
  hipStream_t st = ctx->stream;
  TP_HIP(hipEventRecord(ctx->ev0, st));
  rc = adj_forward(ctx, t, p, turn0, nturns, K, fast, (double*)ckpt.p);
  if (rc) return rc;
  TP_HIP(hipEventRecord(ctx->ev1, st));
  TP_HIP(hipEventSynchronize(ctx->ev1));
  float fwd_ms = 0.f;
  TP_HIP(hipEventElapsedTime(&fwd_ms, ctx->ev0, ctx->ev1));
  TpAdjArgs a{};
  a.seq = t->d_seq;
  a.recs = t->d_recs;
  a.pool = t->d_pool;
  a.slot_of_pos = nullptr;  // no strength gradients
  a.lam = (double*)lam.p;
  a.lost_turn = p->d_lt;
  a.partial = (double*)partial.p;
  a.n = n;
  a.ld = n;
  a.n_seq = (int32_t)t->hdr.n_seq;
  a.n_slots = 0;
  a.variant = adj_variant(t->features, t->variant, has_rotations(t->recs));
  a.exact_h = (t->hdr.flags & TP_H_EXACT_HAMILTONIAN) ? 1 : 0;
  a.minv = t->d_minv;
  a.waves = ctx->back_waves;
  a.n_rows = t->hdr.n_rows > 1 ? (int32_t)t->hdr.n_rows : 1;
  a.row_turn0 = t->hdr.row_turn0;
  a.beti = t->hdr.beti;
  std::vector<double> row(6 * plane);
  float bwd_ms = 0.f;
  // row i of every particle's Jacobian: one reverse pass from the unit seed e_i
  for (int i = 0; i < 6 && n > 0; ++i) {
    TP_HIP(hipEventRecord(ctx->ev0, st));
    for (int32_t j = nseg - 1; j >= 0; --j) {
      a.end_state = (const double*)ckpt.p + (size_t)j * 6 * plane;
      a.nturns = (j == nseg - 1) ? (nturns - j * K) : K;
      a.turn_last = turn0 + j * K + a.nturns - 1;
      a.init_lambda = (j == nseg - 1) ? 2 + i : 0;
      hipError_t le = fast ? tp_launch_adjoint_fast(a, st) : tp_launch_adjoint_exact(a, st);
      if (le != hipSuccess) return fail(TP_EHIP, "tp_jacobian: reverse kernel launch failed: %s", hipGetErrorName(le));
    }
    TP_HIP(hipEventRecord(ctx->ev1, st));
    TP_HIP(hipMemcpyAsync(row.data(), lam.p, 6 * (size_t)n * sizeof(double), hipMemcpyDeviceToHost, st));
    TP_HIP(hipStreamSynchronize(st));
    float ms = 0.f;
    TP_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    bwd_ms += ms;
    if (jac)
      for (int64_t q = 0; q < n; ++q)
        for (int c = 0; c < 6; ++c) jac[(size_t)q * 36 + (size_t)i * 6 + c] = row[(size_t)c * n + q];
  }
  if (flags & TP_TRACK_TIMED) {
    ctx->timing.forward_ms = fwd_ms;
    ctx->timing.backward_ms = bwd_ms;
  }
  return TP_OK;
}

}  // extern "C"

namespace tpr {

void dfree(void*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

void batch_free_plan(tp_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  for (void** p : {&b->d_desc, &b->d_items, &b->d_counter, &b->d_vfirst, &b->d_vtiles, &b->d_ckpt, &b->d_item_loss,
                   &b->d_item_partial, &b->d_grad, &b->d_vloss, &b->d_lam})
    dfree(*p);
  b->mode = 0;
  b->n_items = b->grid_f = b->grid_b = 0;
  b->ckpt_bytes = 0;
}

hipError_t dalloc(void** p, size_t bytes) { return hipMalloc(p, bytes > 0 ? bytes : 8); }

// Persistent grid: as many blocks as stay resident (occupancy x CUs), but no more than
// the items can feed.
int pick_grid(int occ_per_cu, int cus, int wpb, int32_t n_items) {
  const int per_cu = occ_per_cu < 1 ? 1 : occ_per_cu;
  const int64_t need = ((int64_t)n_items + wpb - 1) / wpb;
  return (int)std::max<int64_t>(1, std::min<int64_t>((int64_t)per_cu * cus, need));
}

// Build the work items, descriptors and buffers for one launch shape and capture the
// launch sequence in a graph: counter reset (a fill kernel: no memset node), the batched
// forward kernel and, for the adjoint, the reverse kernel and the per-variant reduction.
int batch_plan(tp_batch* b, int mode, const int32_t* turn0, const int32_t* nturns, int32_t stride, uint32_t flags) {
  const int32_t nv = b->n_var;
  std::vector<int32_t> t0(nv, 0), nt(nturns, nturns + nv);
  if (turn0) t0.assign(turn0, turn0 + nv);
  if (b->mode == mode && b->pflags == flags && b->pstride == stride && b->pblock == b->block &&
      b->pfw == b->fwd_waves && b->pbw == b->back_waves && b->pturn0 == t0 && b->pnturns == nt)
    return TP_OK;
  batch_free_plan(b);
  struct Item {
    int64_t cost;
    int32_t v, tile;
  };
  std::vector<Item> items;
  for (int32_t v = 0; v < nv; ++v) {
    const int64_t n = b->parts[v]->n;
    if (n <= 0 || nt[v] <= 0) continue;
    const int64_t tiles = (n + 63) / 64;
    if ((int64_t)items.size() + tiles > 0x7fffffffLL) return fail(TP_EINVAL, "batch: more than 2^31 work items");
    for (int64_t k = 0; k < tiles; ++k) items.push_back(Item{(int64_t)b->n_seq[v] * nt[v], v, (int32_t)k});
  }
  // longest items first (LPT order); ties by variant then tile, so the tiles of a
  // variant are contiguous and in tile order
  std::sort(items.begin(), items.end(), [](const Item& x, const Item& y) {
    if (x.cost != y.cost) return x.cost > y.cost;
    if (x.v != y.v) return x.v < y.v;
    return x.tile < y.tile;
  });
  const int32_t n_items = (int32_t)items.size();
  std::vector<int32_t> vfirst(nv > 0 ? nv : 1, 0), vtiles(nv > 0 ? nv : 1, 0);
  std::vector<int32_t> hitems(2 * (size_t)(n_items > 0 ? n_items : 1));
  for (int32_t i = 0; i < n_items; ++i) {
    const Item& it = items[(size_t)i];
    if (vtiles[it.v] == 0) vfirst[it.v] = i;
    ++vtiles[it.v];
    hitems[2 * (size_t)i] = it.v;
    hitems[2 * (size_t)i + 1] = it.tile;
  }
  const bool adj = (mode == 2);
  const int32_t K = stride > 0 ? stride : 1;
  std::vector<TpVarDesc> desc(nv > 0 ? nv : 1);
  std::vector<int64_t> lam_off(nv > 0 ? nv : 1, 0);
  int64_t ck = 0, lam = 0;
  for (int32_t v = 0; v < nv; ++v) {
    TpVarDesc& d = desc[v];
    d = TpVarDesc{};
    tp_particles* p = b->parts[v];
    d.seq_off = b->seq_off[v];
    d.pos_off = b->pos_off[v];
    d.n_rows = b->hdr[v].n_rows > 1 ? (int32_t)b->hdr[v].n_rows : 1;
    d.row_turn0 = b->hdr[v].row_turn0;
    d.rec_off = b->rec_off[v];
    d.pool_off = b->pool_off[v];
    d.slot_off = b->slot_off[v];
    d.coords = p->d_coords;
    d.lost_turn = p->d_lt;
    d.lost_elem = p->d_le;
    d.n = p->n;
    d.n_seq = b->n_seq[v];
    d.nturns = nt[v];
    d.turn0 = t0[v];
    d.n_slots = b->n_slots[v];
    d.nseg = adj ? (nt[v] + K - 1) / K : 0;
    d.ckpt_stride = K;
    d.beti = b->hdr[v].beti;
    d.coord_limit = b->hdr[v].coord_limit;
    d.angle_limit = b->hdr[v].angle_limit;
    d.ap_r2 = b->hdr[v].aperture_r2;
    d.ck_off = ck;
    if (adj) ck += (int64_t)vtiles[v] * d.nseg * 384;
    lam_off[v] = lam;
    lam += 6 * p->n;
  }
  hipError_t e = dalloc(&b->d_desc, desc.size() * sizeof(TpVarDesc));
  if (e == hipSuccess) e = dalloc(&b->d_items, hitems.size() * sizeof(int32_t));
  if (e == hipSuccess) e = dalloc(&b->d_counter, 2 * sizeof(int32_t));
  if (adj) {
    if (e == hipSuccess) e = dalloc(&b->d_vfirst, vfirst.size() * sizeof(int32_t));
    if (e == hipSuccess) e = dalloc(&b->d_vtiles, vtiles.size() * sizeof(int32_t));
    if (e == hipSuccess) e = dalloc(&b->d_ckpt, (size_t)ck * sizeof(double));
    if (e == hipSuccess) e = dalloc(&b->d_item_loss, 2 * (size_t)n_items * sizeof(double));
    if (e == hipSuccess) e = dalloc(&b->d_item_partial, (size_t)n_items * (size_t)b->max_slots * sizeof(double));
    if (e == hipSuccess) e = dalloc(&b->d_grad, (size_t)b->total_slots * sizeof(double));
    if (e == hipSuccess) e = dalloc(&b->d_vloss, 2 * (size_t)(nv > 0 ? nv : 1) * sizeof(double));
    if (e == hipSuccess) e = dalloc(&b->d_lam, (size_t)lam * sizeof(double));
  }
  if (e != hipSuccess) {
    batch_free_plan(b);
    return fail(e == hipErrorOutOfMemory ? TP_ENOMEM : TP_EHIP, "batch: allocation of the launch plan (%lld B of checkpoints): %s",
                (long long)(ck * 8), hipGetErrorName(e));
  }
  if (adj)
    for (int32_t v = 0; v < nv; ++v) desc[v].lam = (double*)b->d_lam + lam_off[v];
  hipStream_t st = b->ctx->stream;
  TP_HIP(hipMemcpyAsync(b->d_desc, desc.data(), desc.size() * sizeof(TpVarDesc), hipMemcpyHostToDevice, st));
  TP_HIP(hipMemcpyAsync(b->d_items, hitems.data(), hitems.size() * sizeof(int32_t), hipMemcpyHostToDevice, st));
  if (adj) {
    TP_HIP(hipMemcpyAsync(b->d_vfirst, vfirst.data(), vfirst.size() * sizeof(int32_t), hipMemcpyHostToDevice, st));
    TP_HIP(hipMemcpyAsync(b->d_vtiles, vtiles.data(), vtiles.size() * sizeof(int32_t), hipMemcpyHostToDevice, st));
  }
  TP_HIP(hipStreamSynchronize(st));
  const bool fast = (flags & TP_TRACK_FAST) != 0;
  const int wpb = b->block / 64;
  TpBatchArgs a{};
  a.desc = (const TpVarDesc*)b->d_desc;
  a.items = (const int2*)b->d_items;
  a.n_items = n_items;
  a.seq = b->d_seq;
  a.recs = b->d_recs;
  a.pool = b->d_pool;
  a.slots = b->d_slots;
  a.ckpt = (double*)b->d_ckpt;
  a.item_loss = (double*)b->d_item_loss;
  a.item_partial = (double*)b->d_item_partial;
  a.max_slots = b->max_slots;
  a.variant = b->variant;
  a.back_variant = adj_variant(b->features, b->variant, b->d_minv != nullptr);
  a.minv = b->d_minv;
  a.exact_h = b->exact_h;
  a.block = b->block;
  const int occ_f = fast ? tp_batch_occupancy_fast(b->variant, b->exact_h, adj, b->block, b->fwd_waves)
                         : tp_batch_occupancy_exact(b->variant, b->exact_h, adj, b->block, b->fwd_waves);
  if (occ_f < 1) return fail(TP_EHIP, "batch: the forward kernel does not fit a CU with %d threads", b->block);
  b->grid_f = pick_grid(occ_f, b->cus, wpb, n_items);
  if (adj) {
    const int occ_b = fast ? tp_batch_back_occupancy_fast(a.back_variant, b->block, b->max_slots, b->back_waves)
                           : tp_batch_back_occupancy_exact(a.back_variant, b->block, b->max_slots, b->back_waves);
    if (occ_b < 1) return fail(TP_EHIP, "batch: the reverse kernel does not fit a CU with %d threads", b->block);
    b->grid_b = pick_grid(occ_b, b->cus, wpb, n_items);
  }
  if (n_items > 0) {
    TP_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
    hipError_t le = tp_launch_fill_i32((int32_t*)b->d_counter, 0, 2, st);
    a.counter = (int32_t*)b->d_counter;
    a.grid = b->grid_f;
    a.waves = b->fwd_waves;
    if (le == hipSuccess) le = fast ? tp_launch_batch_fast(a, adj, st) : tp_launch_batch_exact(a, adj, st);
    if (adj) {
      a.counter = (int32_t*)b->d_counter + 1;
      a.grid = b->grid_b;
      a.waves = b->back_waves;
      if (le == hipSuccess) le = fast ? tp_launch_batch_back_fast(a, st) : tp_launch_batch_back_exact(a, st);
      if (le == hipSuccess)
        le = tp_launch_batch_reduce(a.desc, (const int32_t*)b->d_vfirst, (const int32_t*)b->d_vtiles, nv,
                                    a.item_partial, a.max_slots, a.item_loss, (double*)b->d_grad, (double*)b->d_vloss, st);
    }
    hipGraph_t g = nullptr;
    const hipError_t ee = hipStreamEndCapture(st, &g);
    if (le != hipSuccess || ee != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      batch_free_plan(b);
      return fail(TP_EHIP, "batch: graph capture failed: %s", hipGetErrorName(le != hipSuccess ? le : ee));
    }
    b->graph = g;
    e = hipGraphInstantiateWithFlags(&b->exec, g, 0);
    if (e != hipSuccess) {
      batch_free_plan(b);
      return fail(TP_EHIP, "batch: graph instantiation failed: %s", hipGetErrorName(e));
    }
  }
  b->n_items = n_items;
  b->ckpt_bytes = ck * 8;
  b->lam_off = lam_off;
  b->mode = mode;
  b->pflags = flags;
  b->pstride = stride;
  b->pblock = b->block;
  b->pfw = b->fwd_waves;
  b->pbw = b->back_waves;
  b->pturn0 = t0;
  b->pnturns = nt;
  return TP_OK;
}

int batch_check(tp_ctx* ctx, tp_batch* b, const int32_t* turn0, const int32_t* nturns, int32_t min_turns, const char* who) {
  if (!ctx || !b || (!nturns && b->n_var > 0)) return fail(TP_EINVAL, "%s: NULL argument", who);
  if (b->ctx != ctx) return fail(TP_EINVAL, "%s: batch belongs to another context", who);
  if (ctx->closed) return fail(TP_EINVAL, "%s: context destroyed", who);
  for (int32_t v = 0; v < b->n_var; ++v) {
    const int64_t t0 = turn0 ? turn0[v] : 0;
    if (nturns[v] < min_turns) return fail(TP_EINVAL, "%s: nturns[%d] = %d", who, v, nturns[v]);
    if (t0 < 0 || t0 + nturns[v] > 0x7fffffffLL) return fail(TP_EINVAL, "%s: turn range of variant %d overflows int32", who, v);
    const tp_table_header& h = b->hdr[(size_t)v];
    if (h.n_rows > 1 && nturns[v] > 0 && (t0 < h.row_turn0 || t0 + nturns[v] > (int64_t)h.row_turn0 + h.n_rows))
      return fail(TP_EINVAL, "%s: turns of variant %d outside its table's time-varying rows", who, v);
  }
  return set_dev(ctx);
}

int batch_launch(tp_batch* b, uint32_t flags) {
  tp_ctx* ctx = b->ctx;
  hipStream_t st = ctx->stream;
  const bool timed = (flags & TP_TRACK_TIMED) != 0;
  if (timed) TP_HIP(hipEventRecord(ctx->ev0, st));
  if (b->exec) TP_HIP(hipGraphLaunch(b->exec, st));
  if (timed) {
    TP_HIP(hipEventRecord(ctx->ev1, st));
    TP_HIP(hipEventSynchronize(ctx->ev1));
    float ms = 0.f;
    TP_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    ctx->timing.kernel_ms = ms;
    ctx->timing.launches = b->exec ? 1 : 0;
    ctx->timing.block = b->block;
    ctx->timing.ppl = 1;
    ctx->timing.turns_per_launch = 0;
  }
  return TP_OK;
}

}  // namespace tpr

extern "C" {

int tp_batch_create(tp_ctx* ctx, int32_t n_var, tp_table* const* tables, tp_particles* const* parts, tp_batch** out) {
  if (!ctx || !out || n_var < 0 || (n_var > 0 && (!tables || !parts))) return fail(TP_EINVAL, "tp_batch_create: bad argument");
  if (ctx->closed) return fail(TP_EINVAL, "tp_batch_create: context destroyed");
  *out = nullptr;
  if (n_var > (1 << 20)) return fail(TP_EINVAL, "tp_batch_create: n_var = %d", n_var);
  for (int32_t v = 0; v < n_var; ++v) {
    if (!tables[v] || !parts[v]) return fail(TP_EINVAL, "tp_batch_create: variant %d: NULL table or particle set", v);
    if (tables[v]->ctx != ctx || parts[v]->ctx != ctx)
      return fail(TP_EINVAL, "tp_batch_create: variant %d belongs to another context", v);
    if (((tables[v]->features ^ tables[0]->features) & FT_LINH) != 0)
      return fail(TP_EUNSUPPORTED, "tp_batch_create: variants mix the exact and the expanded Hamiltonian");
  }
  int rc = set_dev(ctx);
  if (rc) return rc;
  std::unique_ptr<tp_batch> b(new (std::nothrow) tp_batch());
  if (!b) return fail(TP_ENOMEM, "tp_batch_create: out of host memory");
  std::vector<int32_t> seq, slots;
  std::vector<tp_rec> recs;
  std::vector<double> pool;
  for (int32_t v = 0; v < n_var; ++v) {
    const tp_table* t = tables[v];
    b->features |= t->features;
    b->seq_off.push_back((int64_t)seq.size());
    b->pos_off.push_back((int64_t)slots.size());
    b->rec_off.push_back((int64_t)recs.size());
    b->pool_off.push_back((int64_t)pool.size());
    b->slot_off.push_back(b->total_slots);
    b->n_seq.push_back((int32_t)t->hdr.n_seq);
    b->hdr.push_back(t->hdr);
    seq.insert(seq.end(), t->seq.begin(), t->seq.end());
    recs.insert(recs.end(), t->recs.begin(), t->recs.end());
    pool.insert(pool.end(), t->pool.begin(), t->pool.end());
    int32_t ns = 0;
    for (uint32_t e = 0; e < t->hdr.n_seq; ++e) {
      const int32_t g = t->recs[t->seq[e]].gorder;
      slots.push_back(g >= 0 ? ns++ : -1);
      b->gorder.push_back(g);
    }
    b->n_slots.push_back(ns);
    b->total_slots += ns;
    b->max_slots = std::max(b->max_slots, ns);
    if (b->no_adjoint.empty()) {
      if (!rows_share_slots(t)) b->no_adjoint = "a position changes its main multipole order between rows";
      for (const tp_rec& r : t->recs) {
        if (r.tag == TP_ESBEND && (r.flags & TP_F_DRIFTLIMIT))
          b->no_adjoint = "exact bends with |angle/L| < 1e-6 (drift limit) have no adjoint here";
      }
    }
  }
  const unsigned feat = b->features;
  if ((feat & ~TP_FEAT_A) == 0) b->variant = TP_VARIANT_A;
  else if ((feat & ~TP_FEAT_B) == 0) b->variant = TP_VARIANT_B;
  else if ((feat & ~TP_FEAT_C) == 0) b->variant = TP_VARIANT_C;
  else b->variant = TP_VARIANT_ALL;
  b->exact_h = (feat & FT_LINH) ? 0 : 1;
  std::vector<double> minv;
  for (int32_t v = 0; v < n_var; ++v) {
    if (!has_rotations(tables[v]->recs)) continue;
    if (minv.empty()) minv.assign(pool.size(), 0.0);
    if (!rotation_inverses(tables[v]->recs, pool, (size_t)b->pool_off[v], minv) && b->no_adjoint.empty())
      b->no_adjoint = "a singular r1/r2 matrix has no inverse";
  }
  TP_HIP(hipDeviceGetAttribute(&b->cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
  hipError_t e = dalloc((void**)&b->d_seq, seq.size() * sizeof(int32_t));
  if (e == hipSuccess) e = dalloc((void**)&b->d_recs, recs.size() * sizeof(tp_rec));
  if (e == hipSuccess) e = dalloc((void**)&b->d_pool, pool.size() * sizeof(double));
  if (e == hipSuccess) e = dalloc((void**)&b->d_slots, slots.size() * sizeof(int32_t));
  if (e == hipSuccess && !minv.empty()) e = dalloc((void**)&b->d_minv, minv.size() * sizeof(double));
  hipStream_t st = ctx->stream;
  if (e == hipSuccess && !seq.empty()) e = hipMemcpyAsync(b->d_seq, seq.data(), seq.size() * 4, hipMemcpyHostToDevice, st);
  if (e == hipSuccess && !recs.empty())
    e = hipMemcpyAsync(b->d_recs, recs.data(), recs.size() * sizeof(tp_rec), hipMemcpyHostToDevice, st);
  if (e == hipSuccess && !pool.empty()) e = hipMemcpyAsync(b->d_pool, pool.data(), pool.size() * 8, hipMemcpyHostToDevice, st);
  if (e == hipSuccess && !slots.empty())
    e = hipMemcpyAsync(b->d_slots, slots.data(), slots.size() * 4, hipMemcpyHostToDevice, st);
  if (e == hipSuccess && !minv.empty())
    e = hipMemcpyAsync(b->d_minv, minv.data(), minv.size() * 8, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) {
    for (void* p : {(void*)b->d_seq, (void*)b->d_recs, (void*)b->d_pool, (void*)b->d_slots, (void*)b->d_minv})
      if (p) (void)hipFree(p);
    return fail(e == hipErrorOutOfMemory ? TP_ENOMEM : TP_EHIP, "tp_batch_create: %s", hipGetErrorName(e));
  }
  uint64_t h = fnv1a(FNV_OFFSET, seq.data(), seq.size() * sizeof(int32_t));
  h = fnv1a(h, recs.data(), recs.size() * sizeof(tp_rec));
  h = fnv1a(h, pool.data(), pool.size() * sizeof(double));
  for (int32_t v = 0; v < n_var; ++v) {
    h = fnv1a(h, &parts[v]->n, sizeof(int64_t));
    h = fnv1a(h, &tables[v]->hdr, sizeof(tp_table_header));
  }
  b->hash = h;
  b->ctx = ctx;
  b->n_var = n_var;
  for (int32_t v = 0; v < n_var; ++v) {
    parts[v]->refs.fetch_add(1);
    b->parts.push_back(parts[v]);
  }
  ctx->refs.fetch_add(1);
  *out = b.release();
  return TP_OK;
}

int tp_batch_destroy(tp_batch* b) {
  if (!b) return TP_OK;
  tp_ctx* ctx = b->ctx;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  batch_free_plan(b);
  for (void* p : {(void*)b->d_seq, (void*)b->d_recs, (void*)b->d_pool, (void*)b->d_slots, (void*)b->d_minv})
    if (p) (void)hipFree(p);
  for (tp_particles* p : b->parts) particles_release(p);
  delete b;
  ctx_release(ctx);
  return TP_OK;
}

int tp_batch_set_launch(tp_batch* b, int block, int fwd_waves, int back_waves) {
  if (!b) return fail(TP_EINVAL, "tp_batch_set_launch: batch is NULL");
  if (block == 0) block = 256;
  if (block < 64 || block > 256 || block % 64 != 0)
    return fail(TP_EINVAL, "tp_batch_set_launch: block %d not in {64, 128, 192, 256}", block);
  if (fwd_waves != 0 && fwd_waves != 5 && fwd_waves != 6)
    return fail(TP_EINVAL, "tp_batch_set_launch: fwd_waves %d not in {0, 5, 6}", fwd_waves);
  if (back_waves != 0 && back_waves != 3 && back_waves != 4)
    return fail(TP_EINVAL, "tp_batch_set_launch: back_waves %d not in {0, 3, 4}", back_waves);
  b->block = block;
  b->fwd_waves = fwd_waves;
  b->back_waves = back_waves;
  return TP_OK;
}

int tp_batch_info(const tp_batch* b, int32_t* n_items, int32_t* grid_fwd, int32_t* grid_back, int64_t* ckpt_bytes) {
  if (!b) return fail(TP_EINVAL, "tp_batch_info: batch is NULL");
  if (n_items) *n_items = b->n_items;
  if (grid_fwd) *grid_fwd = b->grid_f;
  if (grid_back) *grid_back = b->grid_b;
  if (ckpt_bytes) *ckpt_bytes = b->ckpt_bytes;
  return TP_OK;
}

int tp_batch_track(tp_ctx* ctx, tp_batch* b, const int32_t* turn0, const int32_t* nturns, uint32_t flags) {
  int rc = batch_check(ctx, b, turn0, nturns, 0, "tp_batch_track");
  if (rc) return rc;
  if (flags & ~(TP_TRACK_FAST | TP_TRACK_ASYNC | TP_TRACK_TIMED))
    return fail(TP_EINVAL, "tp_batch_track: unknown flags 0x%x", flags);
  rc = batch_plan(b, 1, turn0, nturns, 0, flags & TP_TRACK_FAST);
  if (rc) return rc;
  rc = batch_launch(b, flags);
  if (rc) return rc;
  if (!(flags & TP_TRACK_ASYNC)) TP_HIP(hipStreamSynchronize(ctx->stream));
  return TP_OK;
}

int tp_batch_adjoint(tp_ctx* ctx, tp_batch* b, const int32_t* turn0, const int32_t* nturns, int32_t ckpt_stride,
                     uint32_t flags, double* loss, int64_t* n_alive, double* dL_dk, double* dL_dx0) {
  int rc = batch_check(ctx, b, turn0, nturns, 1, "tp_batch_adjoint");
  if (rc) return rc;
  if (flags & ~(TP_TRACK_FAST | TP_TRACK_TIMED)) return fail(TP_EINVAL, "tp_batch_adjoint: unsupported flags 0x%x", flags);
  if (ckpt_stride < 0) return fail(TP_EINVAL, "tp_batch_adjoint: ckpt_stride = %d", ckpt_stride);
  if (!b->no_adjoint.empty()) return fail(TP_EUNSUPPORTED, "tp_batch_adjoint: %s", b->no_adjoint.c_str());
  if ((size_t)(b->block / 64) * ((size_t)b->max_slots + 128) * sizeof(double) > 64 * 1024)
    return fail(TP_EUNSUPPORTED, "tp_batch_adjoint: %d strength slots exceed the 64 KiB LDS budget", b->max_slots);
  const int32_t K = ckpt_stride > 0 ? ckpt_stride : 1;
  rc = batch_plan(b, 2, turn0, nturns, K, flags & TP_TRACK_FAST);
  if (rc) return rc;
  rc = batch_launch(b, flags);
  if (rc) return rc;
  hipStream_t st = ctx->stream;
  const int32_t nv = b->n_var;
  std::vector<double> vloss(2 * (size_t)(nv > 0 ? nv : 1), 0.0), grad((size_t)(b->total_slots > 0 ? b->total_slots : 1), 0.0);
  if (b->exec) {
    TP_HIP(hipMemcpyAsync(vloss.data(), b->d_vloss, 2 * (size_t)nv * sizeof(double), hipMemcpyDeviceToHost, st));
    if (b->total_slots > 0)
      TP_HIP(hipMemcpyAsync(grad.data(), b->d_grad, (size_t)b->total_slots * sizeof(double), hipMemcpyDeviceToHost, st));
    if (dL_dx0) {
      const int64_t tot = nv > 0 ? b->lam_off[(size_t)nv - 1] + 6 * b->parts[(size_t)nv - 1]->n : 0;
      if (tot > 0) TP_HIP(hipMemcpyAsync(dL_dx0, b->d_lam, (size_t)tot * sizeof(double), hipMemcpyDeviceToHost, st));
    }
  }
  TP_HIP(hipStreamSynchronize(st));
  int64_t pos = 0;
  for (int32_t v = 0; v < nv; ++v) {
    const double na = vloss[2 * (size_t)v + 1];
    const double inv_na = na > 0.0 ? 1.0 / na : 0.0;
    if (loss) loss[v] = na > 0.0 ? vloss[2 * (size_t)v] / na : 0.0;
    if (n_alive) n_alive[v] = (int64_t)na;
    if (dL_dk) {
      int32_t k = 0;
      for (int32_t e = 0; e < b->n_seq[v]; ++e) {
        const int32_t g = b->gorder[(size_t)(b->pos_off[v] + e)];
        double val = 0.0;
        if (g >= 0) {
          // d b_g / d k: b1 = k1, b2 = k2/2, b3 = k3/6 (src/tracking.jl:395, 446, 478)
          const double sc = g == 1 ? 1.0 : (g == 2 ? 0.5 : (g == 3 ? 1.0 / 6.0 : 1.0));
          val = na > 0.0 ? grad[(size_t)(b->slot_off[v] + k)] * inv_na * sc : 0.0;
          ++k;
        }
        dL_dk[pos + e] = val;
      }
      pos += b->n_seq[v];
    }
    if (dL_dx0) {
      double* x = dL_dx0 + b->lam_off[v];
      const int64_t m = 6 * b->parts[v]->n;
      if (!b->exec || na <= 0.0)
        for (int64_t i = 0; i < m; ++i) x[i] = 0.0;
      else
        for (int64_t i = 0; i < m; ++i) x[i] *= inv_na;
    }
  }
  return TP_OK;
}

}  // extern "C"
