// tp_autotune.hip -- launch-shape autotuner of libtrackpad_hip (tp_autotune,
// tp_batch_autotune in include/trackpad_hip.h).
//
// Every candidate shape is timed with HIP events on the context's stream, on a device
// copy of the caller's beam (or, for a batch, on the bound particle sets, restored after
// every probe), so tuning never changes what the caller tracks.  The choice depends on the
// device, the lattice and the beam size, so it is cached in a plain text file, one line
// per key:
//   <key> \t <block> <ppl> <turns_per_launch> <fwd_waves> <back_waves> <rate>
#include <hip/hip_runtime.h>

#include <cinttypes>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "tp_runtime.h"

using namespace tpr;

namespace {

struct Choice {
  int block = 256, ppl = 1, tpl = 0, fwd = 0, back = 0;
  double rate = 0.0;
};

bool has_budgets(int variant) { return variant == TP_VARIANT_B || variant == TP_VARIANT_C; }

std::string device_key(int device) {
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) return "unknown";
  char buf[320];
  snprintf(buf, sizeof buf, "%s:%dcu", prop.gcnArchName, prop.multiProcessorCount);
  return buf;
}

uint64_t table_hash(const tp_table* t) {
  uint64_t h = fnv1a(FNV_OFFSET, &t->hdr, sizeof t->hdr);
  h = fnv1a(h, t->seq.data(), t->seq.size() * sizeof(int32_t));
  h = fnv1a(h, t->recs.data(), t->recs.size() * sizeof(tp_rec));
  return fnv1a(h, t->pool.data(), t->pool.size() * sizeof(double));
}

bool cache_lookup(const char* path, const std::string& key, Choice* c) {
  if (!path) return false;
  FILE* f = fopen(path, "r");
  if (!f) return false;
  char line[1024];
  bool found = false;
  while (fgets(line, sizeof line, f)) {
    char* tab = strchr(line, '\t');
    if (!tab || (size_t)(tab - line) != key.size() || strncmp(line, key.c_str(), key.size()) != 0) continue;
    Choice x;
    if (sscanf(tab + 1, "%d %d %d %d %d %lf", &x.block, &x.ppl, &x.tpl, &x.fwd, &x.back, &x.rate) == 6) {
      *c = x;  // the last entry for a key wins
      found = true;
    }
  }
  fclose(f);
  return found;
}

void cache_store(const char* path, const std::string& key, const Choice& c) {
  if (!path) return;
  FILE* f = fopen(path, "a");
  if (!f) return;  // an unwritable cache only costs a re-tune next time
  fprintf(f, "%s\t%d %d %d %d %d %.6e\n", key.c_str(), c.block, c.ppl, c.tpl, c.fwd, c.back, c.rate);
  fclose(f);
}

void fill_out(tp_launch_cfg* out, const Choice& c, int candidates) {
  if (!out) return;
  out->block = c.block;
  out->ppl = c.ppl;
  out->turns_per_launch = c.tpl;
  out->fwd_waves = c.fwd;
  out->back_waves = c.back;
  out->candidates = candidates;
  out->rate = c.rate;
}

// Device copy of every particle set of a batch, to restore them after each probe.
struct BatchBackup {
  tp_batch* b = nullptr;
  std::vector<void*> bufs;
  ~BatchBackup() {
    for (void* p : bufs)
      if (p) (void)hipFree(p);
  }
  hipError_t save() {
    hipStream_t st = b->ctx->stream;
    for (tp_particles* p : b->parts) {
      void* d = nullptr;
      const size_t n = (size_t)p->n;
      hipError_t e = hipMalloc(&d, n > 0 ? n * (6 * sizeof(double) + 2 * sizeof(int32_t)) : 8);
      if (e != hipSuccess) return e;
      bufs.push_back(d);
      if (n == 0) continue;
      char* c = (char*)d;
      e = hipMemcpyAsync(c, p->d_coords, 6 * n * sizeof(double), hipMemcpyDeviceToDevice, st);
      if (e == hipSuccess) e = hipMemcpyAsync(c + 6 * n * 8, p->d_lt, n * 4, hipMemcpyDeviceToDevice, st);
      if (e == hipSuccess) e = hipMemcpyAsync(c + 6 * n * 8 + n * 4, p->d_le, n * 4, hipMemcpyDeviceToDevice, st);
      if (e != hipSuccess) return e;
    }
    return hipStreamSynchronize(st);
  }
  hipError_t restore() {
    hipStream_t st = b->ctx->stream;
    for (size_t v = 0; v < b->parts.size(); ++v) {
      tp_particles* p = b->parts[v];
      const size_t n = (size_t)p->n;
      if (n == 0) continue;
      const char* c = (const char*)bufs[v];
      hipError_t e = hipMemcpyAsync(p->d_coords, c, 6 * n * sizeof(double), hipMemcpyDeviceToDevice, st);
      if (e == hipSuccess) e = hipMemcpyAsync(p->d_lt, c + 6 * n * 8, n * 4, hipMemcpyDeviceToDevice, st);
      if (e == hipSuccess) e = hipMemcpyAsync(p->d_le, c + 6 * n * 8 + n * 4, n * 4, hipMemcpyDeviceToDevice, st);
      if (e != hipSuccess) return e;
    }
    return hipStreamSynchronize(st);
  }
};

}  // namespace

extern "C" {

int tp_autotune(tp_ctx* ctx, const tp_table* t, const tp_particles* p, int32_t probe_turns, uint32_t flags,
                double target_launch_ms, const char* cache_path, tp_launch_cfg* out) {
  if (!ctx || !t || !p) return fail(TP_EINVAL, "tp_autotune: NULL argument");
  if (t->ctx != ctx || p->ctx != ctx) return fail(TP_EINVAL, "tp_autotune: table and particles belong to another context");
  if (ctx->closed) return fail(TP_EINVAL, "tp_autotune: context destroyed");
  if (probe_turns < 1) return fail(TP_EINVAL, "tp_autotune: probe_turns = %d", probe_turns);
  if (flags & ~TP_TRACK_FAST) return fail(TP_EINVAL, "tp_autotune: unsupported flags 0x%x", flags);
  if (!(target_launch_ms >= 0.0)) return fail(TP_EINVAL, "tp_autotune: target_launch_ms = %g", target_launch_ms);
  int rc = set_dev(ctx);
  if (rc) return rc;
  const int64_t n = p->n;
  char kbuf[512];
  snprintf(kbuf, sizeof kbuf, "track %s tab=%016" PRIx64 " n=%" PRId64 " t=%d f=%u target=%.0f",
           device_key(ctx->device).c_str(), table_hash(t), n, probe_turns, flags, target_launch_ms);
  const std::string key = kbuf;
  Choice best;
  best.back = ctx->back_waves;
  if (cache_lookup(cache_path, key, &best)) {
    rc = tp_set_launch(ctx, best.block, best.ppl, best.tpl);
    if (!rc) rc = tp_set_waves(ctx, best.fwd, ctx->back_waves);
    if (rc) return rc;
    best.back = ctx->back_waves;
    fill_out(out, best, 0);
    return TP_OK;
  }
  if (n == 0 || t->hdr.n_seq == 0) {  // nothing to time: defaults
    rc = tp_set_launch(ctx, 256, 1, 0);
    if (!rc) rc = tp_set_waves(ctx, 0, ctx->back_waves);
    if (rc) return rc;
    fill_out(out, best, 0);
    return TP_OK;
  }
  tp_particles* q = nullptr;
  rc = tp_particles_create(ctx, n, &q);
  if (rc) return rc;
  const int save_block = ctx->block, save_ppl = ctx->ppl, save_tpl = ctx->tpl, save_fwd = ctx->fwd_waves;
  // a time-varying table is probed on its own turn rows
  const int32_t pt0 = t->hdr.n_rows > 1 ? t->hdr.row_turn0 : 0;
  if (t->hdr.n_rows > 1 && (uint32_t)probe_turns > t->hdr.n_rows) probe_turns = (int32_t)t->hdr.n_rows;
  int ncand = 0;
  best.rate = -1.0;
  for (int block : {64, 128, 256}) {
    for (int ppl : {1, 2}) {
      for (int w : {0, 5, 6}) {
        if (w != 0 && (ppl != 1 || !has_budgets(t->variant))) continue;
        ctx->block = block;
        ctx->ppl = ppl;
        ctx->tpl = 0;
        ctx->fwd_waves = w;
        // one untimed turn (code-object load, caches), then the timed probe from p's state
        rc = tp_particles_copy(q, p);
        if (!rc) rc = tp_track(ctx, t, q, pt0, 1, flags);
        if (!rc) rc = tp_particles_copy(q, p);
        if (!rc) rc = tp_track(ctx, t, q, pt0, probe_turns, flags | TP_TRACK_TIMED);
        if (rc) break;
        ++ncand;
        const double s = ctx->timing.kernel_ms * 1e-3;
        const double rate = s > 0.0 ? (double)n * probe_turns / s : 0.0;
        if (rate > best.rate) {
          best.rate = rate;
          best.block = block;
          best.ppl = ppl;
          best.fwd = w;
        }
      }
      if (rc) break;
    }
    if (rc) break;
  }
  tp_particles_destroy(q);
  ctx->block = save_block;
  ctx->ppl = save_ppl;
  ctx->tpl = save_tpl;
  ctx->fwd_waves = save_fwd;
  if (rc) return rc;
  const double target_s = (target_launch_ms > 0.0 ? target_launch_ms : 1000.0) * 1e-3;
  const double turns = target_s * best.rate / (double)n;
  best.tpl = turns >= 1.0 ? (turns < 1e9 ? (int)turns : 1000000000) : 1;
  rc = tp_set_launch(ctx, best.block, best.ppl, best.tpl);
  if (!rc) rc = tp_set_waves(ctx, best.fwd, ctx->back_waves);
  if (rc) return rc;
  best.back = ctx->back_waves;
  cache_store(cache_path, key, best);
  fill_out(out, best, ncand);
  return TP_OK;
}

int tp_batch_autotune(tp_ctx* ctx, tp_batch* b, const int32_t* probe_turns, int32_t ckpt_stride, uint32_t flags,
                      const char* cache_path, tp_launch_cfg* out) {
  if (!b) return fail(TP_EINVAL, "tp_batch_autotune: batch is NULL");
  // time-varying variants are probed on their own turn rows
  std::vector<int32_t> pt0((size_t)(b->n_var > 0 ? b->n_var : 1), 0);
  for (int32_t v = 0; v < b->n_var; ++v)
    if (b->hdr[(size_t)v].n_rows > 1) pt0[(size_t)v] = b->hdr[(size_t)v].row_turn0;
  int rc = batch_check(ctx, b, pt0.data(), probe_turns, 1, "tp_batch_autotune");
  if (rc) return rc;
  if (flags & ~(TP_TRACK_FAST | TP_TUNE_ADJOINT)) return fail(TP_EINVAL, "tp_batch_autotune: unsupported flags 0x%x", flags);
  if (ckpt_stride < 0) return fail(TP_EINVAL, "tp_batch_autotune: ckpt_stride = %d", ckpt_stride);
  const bool adj = (flags & TP_TUNE_ADJOINT) != 0;
  if (adj && !b->no_adjoint.empty()) return fail(TP_EUNSUPPORTED, "tp_batch_autotune: %s", b->no_adjoint.c_str());
  const uint32_t kflags = flags & TP_TRACK_FAST;
  const int32_t K = ckpt_stride > 0 ? ckpt_stride : 1;
  uint64_t th = fnv1a(FNV_OFFSET, probe_turns, (size_t)b->n_var * sizeof(int32_t));
  char kbuf[512];
  snprintf(kbuf, sizeof kbuf, "batch %s b=%016" PRIx64 " t=%016" PRIx64 " K=%d f=%u", device_key(ctx->device).c_str(),
           b->hash, th, K, flags);
  const std::string key = kbuf;
  Choice best;
  best.block = b->block;
  best.fwd = b->fwd_waves;
  best.back = b->back_waves;
  if (cache_lookup(cache_path, key, &best)) {
    rc = tp_batch_set_launch(b, best.block, best.fwd, best.back);
    if (rc) return rc;
    fill_out(out, best, 0);
    return TP_OK;
  }
  double pt = 0.0;
  for (int32_t v = 0; v < b->n_var; ++v) pt += (double)b->parts[v]->n * probe_turns[v];
  BatchBackup bk;
  bk.b = b;
  TP_HIP(bk.save());
  int ncand = 0;
  auto probe = [&](int mode, double* ms) -> int {
    int r = batch_plan(b, mode, pt0.data(), probe_turns, mode == 2 ? K : 0, kflags);
    if (!r) r = batch_launch(b, 0);  // untimed: loads the code object, warms the caches
    if (r) return r;
    hipError_t e = bk.restore();
    if (e != hipSuccess) return fail(TP_EHIP, "tp_batch_autotune: restore: %s", hipGetErrorName(e));
    r = batch_launch(b, TP_TRACK_TIMED);
    if (r) return r;
    *ms = b->ctx->timing.kernel_ms;
    e = bk.restore();
    if (e != hipSuccess) return fail(TP_EHIP, "tp_batch_autotune: restore: %s", hipGetErrorName(e));
    ++ncand;
    return TP_OK;
  };
  best.rate = -1.0;
  for (int block : {64, 128, 256}) {
    for (int w : {0, 5, 6}) {
      if (w != 0 && !has_budgets(b->variant)) continue;
      rc = tp_batch_set_launch(b, block, w, b->back_waves);
      double ms = 0.0;
      if (!rc) rc = probe(1, &ms);
      if (rc) return rc;
      const double rate = ms > 0.0 ? pt / (ms * 1e-3) : 0.0;
      if (rate > best.rate) {
        best.rate = rate;
        best.block = block;
        best.fwd = w;
      }
    }
  }
  if (adj) {
    double best_ms = -1.0;
    for (int w : {0, 3, 4}) {
      if (w != 0 && !has_budgets(b->variant)) continue;
      rc = tp_batch_set_launch(b, best.block, best.fwd, w);
      double ms = 0.0;
      if (!rc) rc = probe(2, &ms);
      if (rc) return rc;
      if (best_ms < 0.0 || ms < best_ms) {
        best_ms = ms;
        best.back = w;
      }
    }
  }
  rc = tp_batch_set_launch(b, best.block, best.fwd, best.back);
  if (rc) return rc;
  cache_store(cache_path, key, best);
  fill_out(out, best, ncand);
  return TP_OK;
}

}  // extern "C"
