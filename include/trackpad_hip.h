/*
 * trackpad_hip.h -- C ABI of libtrackpad_hip, the MI355X (gfx950) backend for the
 * multi-particle tracking path of TrackPad.jl.
 *
 * The reference path this library replaces is
 *   linepass!(coords::Matrix{T}, lat::Lattice, beam::Beam{T}, lost_flags::Vector{Int};
 *             time, turn)                                       src/lattice.jl:260-297
 *   ringpass!(coords, lat, beam, lost_flags, nturns; time, dt_turn, turn)
 *                                                               src/lattice.jl:305-317
 * with the per-element maps pass!(elem, r, beti) of src/tracking.jl:115-973.
 *
 * Conventions (all entry points):
 *   - plain C types only; no exceptions cross the boundary;
 *   - return 0 (TP_OK) on success or a negative TP_E* code; tp_last_error() then holds a
 *     message (thread-local), including the hipError_t name for TP_EHIP;
 *   - coords are N x 6 column-major doubles, i.e. six planes x|px|y|py|z|delta of N
 *     values each -- the memory layout of a Julia Matrix{Float64}(N, 6)
 *     (src/lattice.jl:274-275, 287-292);
 *   - loss: lost_flags are the reference's Int64 0/1 (src/lattice.jl:269-284);
 *     lost_turn / lost_elem are int32, 0-based turn and lattice position at whose exit
 *     the particle was lost, -1 while alive, -2 for a particle that arrived already
 *     flagged (lost_flags == 1);
 *   - handles are owned by the library and freed by *_destroy; host buffers are owned by
 *     the caller; a context is not re-entrant (serialise calls per context), distinct
 *     contexts may be used from different threads.
 */
#ifndef TRACKPAD_HIP_H
#define TRACKPAD_HIP_H

#include <stddef.h>
#include <stdint.h>

#include "trackpad_table.h"

#ifdef __cplusplus
extern "C" {
#endif

#define TP_ABI_VERSION 1

enum tp_status {
  TP_OK = 0,
  TP_EINVAL = -1,        /* bad argument or handle */
  TP_EUNSUPPORTED = -2,  /* element or option the device path does not implement */
  TP_EHIP = -3,          /* HIP runtime error */
  TP_ENOMEM = -4,        /* device or host allocation failed */
  TP_ETABLE = -5         /* malformed element table */
};

/* tp_track flags */
#define TP_TRACK_FAST 0x1u  /* FMA-contracted kernel with rsqrt drifts: tolerance, not bits */
#define TP_TRACK_ASYNC 0x2u /* return after enqueueing; tp_sync() waits */
#define TP_TRACK_TIMED 0x4u /* record HIP events around every launch (tp_timing) */

typedef struct tp_ctx tp_ctx;
typedef struct tp_table tp_table;
typedef struct tp_particles tp_particles;
typedef struct tp_batch tp_batch;

typedef struct tp_timing_t {
  double kernel_ms;    /* summed duration of the timed launches of the last timed call */
  int32_t launches;    /* kernel launches in that call */
  int32_t block;       /* work-group size used */
  int32_t ppl;         /* particles per lane used */
  int32_t turns_per_launch;
  double forward_ms;   /* tp_adjoint: forward pass with checkpoints */
  double backward_ms;  /* tp_adjoint: reverse pass and gradient reduction */
} tp_timing_t;

/* Library version (TP_ABI_VERSION). No reference counterpart. */
int tp_version(void);

/* Message of the last failed call on this thread. No reference counterpart (the
 * reference raises Julia exceptions, e.g. src/time_dependence.jl:138). */
const char* tp_last_error(void);

/* Number of visible HIP devices. */
int tp_device_count(int* n);

/* One device and one HIP stream per context. Replaces the implicit single-threaded CPU
 * "device" of src/lattice.jl:266-294.  tp_ctx_destroy may be called before the tables and
 * particle sets created on the context are destroyed (finalizers run in any order): the
 * context is released when the last of them goes; no new work is accepted after it. */
int tp_ctx_create(int device, tp_ctx** out);
int tp_ctx_destroy(tp_ctx* ctx);

/* Launch shape for later tp_track calls: work-group size (64, 128, 192 or 256; 0 =
 * default 256), particles per lane (1 or 2, 0 = default 1), turns per kernel launch
 * (0 = all turns in one launch). */
int tp_set_launch(tp_ctx* ctx, int block, int ppl, int turns_per_launch);

/* Register budget of the kernels used by later calls, as resident waves per SIMD the
 * register allocator must allow (it spills to scratch if it must): fwd_waves 0 (the
 * compiler's choice), 5 or 6 for the tracking kernels; back_waves 0, 3 or 4 for the
 * reverse (adjoint) kernels.  Budgets are compiled for the one-particle-per-lane kernels
 * of lattices made of drifts, multipoles, bends and RF cavities; other kernels keep the
 * compiler's budget.  Results do not depend on the budget.  No reference counterpart. */
int tp_set_waves(tp_ctx* ctx, int fwd_waves, int back_waves);

/* Drift form of the fast build (TP_TRACK_FAST) for later tracking calls on ctx (tp_track,
 * and the forward passes of tp_adjoint / tp_jacobian).  The fast build evaluates the exact
 * drift's L / sqrt(pz^2) (src/tracking.jl:87-99) in one of two forms that round
 * differently and are held to the same accuracy: TP_DRIFT_SERIES, a quadratic series about
 * 1 / sqrt(1 + 2 delta beti + delta^2) where (px^2 + py^2) / that < 2^-18, lane by lane,
 * and the refined reciprocal square root elsewhere; TP_DRIFT_RSQRT, the refined reciprocal
 * square root everywhere.  TP_DRIFT_AUTO (the default) takes the series when the particle
 * set -- its size N, not the number of survivors -- fills at least half of the GPU's
 * resident wave slots, where it is faster: the choice is fixed per particle set, so the
 * launches of a call, survivor compaction and a run split into several calls never switch
 * a particle's form.  Lattices with general kicks of order >= 3 always take the
 * reciprocal square root; the batched kernels always take the series.  The parity build
 * ignores the setting.  No reference counterpart. */
#define TP_DRIFT_AUTO 0
#define TP_DRIFT_SERIES 1
#define TP_DRIFT_RSQRT 2
int tp_set_drift(tp_ctx* ctx, int form);

/* Kernel timing of the last call made with TP_TRACK_TIMED. */
int tp_timing(tp_ctx* ctx, tp_timing_t* out);

/* Upload a validated element table (trackpad_table.h). Replaces the reference's
 * Lattice{E} element vector (src/lattice.jl:87-105) and the materialised-lattice
 * precondition (src/lattice.jl:128-132, src/time_dependence.jl:137-139): time-varying
 * elements cannot be expressed in a table, the writer refuses them. */
int tp_table_create(tp_ctx* ctx, const void* bytes, size_t nbytes, tp_table** out);
int tp_table_destroy(tp_table* t);
/* Validate a table blob without a device (the checks tp_table_create makes); on success
 * *features receives the kernel feature mask the table needs. */
int tp_table_check(const void* bytes, size_t nbytes, uint32_t* features);
/* n_seq, n_rec and the table's beti. */
int tp_table_info(const tp_table* t, int32_t* n_seq, int32_t* n_rec, double* beti);

/* Device-resident particle set (6 planes of n doubles + lost_turn + lost_elem). */
int tp_particles_create(tp_ctx* ctx, int64_t n, tp_particles** out);
int tp_particles_destroy(tp_particles* p);
/* coords: N x 6 column-major (required). lost_flags01 (reference Int64 flags, may be
 * NULL): 1 marks a particle lost before the call. lost_turn / lost_elem (may be NULL)
 * restore the library's own loss record and take precedence over lost_flags01. */
int tp_particles_upload(tp_particles* p, const double* coords, const int64_t* lost_flags01,
                        const int32_t* lost_turn, const int32_t* lost_elem);
/* Any output pointer may be NULL. lost_flags01[i] is written 1 for lost particles and
 * 0 for live ones. */
int tp_particles_download(tp_particles* p, double* coords, int64_t* lost_flags01,
                          int32_t* lost_turn, int32_t* lost_elem);
/* Device-to-device copy of a whole particle set (same n, same device), enqueued on
 * dst's context stream, so it is ordered before later tp_track calls there; tp_sync waits
 * for it.  Restores a saved state for checkpoint/resume (the reference resumes by re-passing
 * coords/lost_flags/turn to ringpass!, src/lattice.jl:305-317). */
int tp_particles_copy(tp_particles* dst, const tp_particles* src);
/* Device pointers of the particle planes (for zero-copy interop); n is returned too. */
int tp_particles_device_ptrs(tp_particles* p, double** coords, int32_t** lost_turn,
                             int32_t** lost_elem, int64_t* n);

/* Track device-resident particles for nturns turns starting at turn number turn0.
 * nturns == 1 is linepass! (src/lattice.jl:260-297); nturns > 1 is ringpass!
 * (src/lattice.jl:305-317). Loss is checked at every element exit with check_lost
 * (src/tracking.jl:49-52) plus the table's aperture; a lost particle keeps its
 * element-entry coordinates and is skipped afterwards (src/lattice.jl:269-284).
 * When the turns take several launches (tp_set_launch turns_per_launch) and the call is
 * synchronous, the survivors are compacted into a particle list between launches, so lost
 * particles stop occupying lanes; results are the same bits either way. */
int tp_track(tp_ctx* ctx, const tp_table* t, tp_particles* p, int32_t turn0, int32_t nturns,
             uint32_t flags);

/* tp_track without any device allocation.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory during a tracking call (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same arguments, results
 * (bit for bit) and errors as tp_track; the survivor lists and the persistent launches'
 * queue counter, which tp_track allocates with the particle set on first use, live instead
 * in a library-owned static device array of 8 MiB per device (held until the process ends),
 * so particle sets larger than TP_NOALLOC_MAX_N particles are refused with TP_EINVAL.  Every
 * call on one device shares that array: calls that overlap on the same device (from
 * different contexts or threads, or TP_TRACK_ASYNC calls on different streams) race, and the
 * caller must serialise them. */
#define TP_NOALLOC_MAX_N 1048576
int tp_track_noalloc(tp_ctx* ctx, const tp_table* t, tp_particles* p, int32_t turn0, int32_t nturns,
                     uint32_t flags);

/* Wait for all work enqueued on the context's stream. */
int tp_sync(tp_ctx* ctx);

/* One-shot mirror of ringpass!/linepass! on host arrays (src/lattice.jl:260-317):
 * uploads coords and lost_flags, tracks, downloads. lost_flags follows the reference:
 * entries == 1 are skipped, newly lost particles are set to 1, other entries are left
 * as they were. lost_turn / lost_elem (may be NULL) receive the loss record. */
int tp_track_host(tp_ctx* ctx, const tp_table* t, double* coords, int64_t n, int64_t* lost_flags,
                  int32_t* lost_turn, int32_t* lost_elem, int32_t turn0, int32_t nturns,
                  uint32_t flags);

/* Reverse-mode gradient of  L = mean over surviving particles of (x^2 + y^2)  after nturns
 * turns (no reference counterpart: the reference's Enzyme extension is an empty stub,
 * ext/TrackPadEnzymeExt.jl:14-17; its only AD use is forward-mode Enzyme on a single
 * particle, test/verify_enzyme_compat.jl:13-63).  p holds the initial beam on entry and
 * the tracked beam on return (as tp_track).  The forward pass stores the beam every
 * ckpt_stride turns (0 = every turn); the reverse pass starts each segment from the
 * stored state at its end and runs every sub-step's inverse map with its vector-Jacobian
 * product.  Outputs (any may be NULL): *loss, *n_alive, dL_dk[n_seq] = dL/dk of each
 * lattice position's main strength (k1 of a Quadrupole, k2 of a Sextupole, k3 of an
 * Octupole; 0 for other positions), dL_dx0 (N x 6 column-major host array) = dL/d initial
 * coordinates.  Particles lost during tracking do not enter L and get zero gradient.
 * Supported: every element the tables hold (both Hamiltonians, t1/t2 and r1/r2
 * misalignments) except an ExactSBend in its drift limit (|angle/L| < 1e-6), which returns
 * TP_EUNSUPPORTED, as does a singular r1/r2 matrix.  A time-varying table runs each turn on its row (turns outside the
 * rows: TP_EINVAL); dL_dk of a position is then the derivative with respect to an offset
 * added to its strength at every turn.  flags: TP_TRACK_FAST / TP_TRACK_TIMED as for
 * tp_track. */
int tp_adjoint(tp_ctx* ctx, const tp_table* t, tp_particles* p, int32_t turn0, int32_t nturns,
               int32_t ckpt_stride, uint32_t flags, double* loss, int64_t* n_alive, double* dL_dk,
               double* dL_dx0);

/* ---- Batched lattice variants ----------------------------------------------------------
 * No reference counterpart: the reference tracks one lattice per linepass!/ringpass! call
 * (src/lattice.jl:260-317); a parameter sweep calls it once per lattice.  A batch binds
 * n_var (table, particle set) pairs -- lattice variants with their own beams, of uneven
 * sizes -- and runs all of them in one persistent launch: waves take (variant, 64-particle
 * tile) work items, ordered by decreasing cost n_seq x nturns, from a device queue.  The
 * launch sequence is captured once in a hipGraph and replayed while the turn counts (and,
 * for the adjoint, the checkpoint stride and flags) stay the same.
 *
 * The tables are copied into the batch (they may be destroyed afterwards); the particle
 * sets are referenced and stay alive until the batch is destroyed.  All variants must use
 * the same Hamiltonian (exact or expanded) and sit on ctx. */
int tp_batch_create(tp_ctx* ctx, int32_t n_var, tp_table* const* tables, tp_particles* const* parts,
                    tp_batch** out);
int tp_batch_destroy(tp_batch* b);

/* Work-group size of the persistent kernels (64..256, multiple of 64; 0 = 256) and their
 * register budgets, as for tp_set_waves.  The persistent grids hold as many blocks as stay
 * resident on every CU. */
int tp_batch_set_launch(tp_batch* b, int block, int fwd_waves, int back_waves);

/* Track variant v from turn turn0[v] (NULL = all 0) for nturns[v] turns, as tp_track on
 * each (table, particle set) pair would.  flags: TP_TRACK_FAST, TP_TRACK_ASYNC,
 * TP_TRACK_TIMED (tp_timing: kernel_ms of the graph launch). */
int tp_batch_track(tp_ctx* ctx, tp_batch* b, const int32_t* turn0, const int32_t* nturns, uint32_t flags);

/* tp_adjoint for every variant in one graph launch: a persistent forward pass that stores
 * each tile's state every ckpt_stride turns, a persistent reverse pass, and a per-variant
 * reduction in tile order.  Per variant v (outputs may be NULL): loss[v], n_alive[v];
 * dL_dk gets n_seq(v) entries per variant, concatenated in variant order; dL_dx0 gets the
 * N(v) x 6 column-major block of each variant, concatenated in variant order.  Supported
 * tables as for tp_adjoint.  flags: TP_TRACK_FAST, TP_TRACK_TIMED. */
int tp_batch_adjoint(tp_ctx* ctx, tp_batch* b, const int32_t* turn0, const int32_t* nturns, int32_t ckpt_stride,
                     uint32_t flags, double* loss, int64_t* n_alive, double* dL_dk, double* dL_dx0);

/* Shape of the last planned batch launch: work items, persistent grids (blocks) of the
 * forward and reverse kernels, checkpoint bytes. */
int tp_batch_info(const tp_batch* b, int32_t* n_items, int32_t* grid_fwd, int32_t* grid_back,
                  int64_t* ckpt_bytes);

/* ---- Autotuner ----------------------------------------------------------------------------
 * No reference counterpart.  Chooses launch shapes by timing candidates on the device. */
typedef struct tp_launch_cfg {
  int32_t block;             /* work-group size */
  int32_t ppl;               /* particles per lane (tp_autotune) */
  int32_t turns_per_launch;  /* tp_autotune: turns per launch (0 = all) */
  int32_t fwd_waves;         /* register budget of the tracking kernels (tp_set_waves) */
  int32_t back_waves;        /* register budget of the reverse kernels */
  int32_t candidates;        /* candidates timed (0: taken from the cache file) */
  double rate;               /* particle-turns/s of the choice on the probe (forward) */
} tp_launch_cfg;

/* tp_autotune flags */
#define TP_TUNE_ADJOINT 0x100u /* tp_batch_autotune: also time the reverse kernel budgets */

/* Time every candidate launch shape -- block {64, 128, 256} x particles per lane {1, 2} x
 * register budget {0, 5, 6} where compiled -- on a device copy of p for probe_turns turns
 * (p itself is not changed), apply the fastest to ctx (tp_set_launch, tp_set_waves), and
 * choose turns per launch so that one launch of N particles lasts about
 * target_launch_ms at the measured rate (0 = 1000 ms).  flags: TP_TRACK_FAST or 0, the
 * kernel flavour to tune.  cache_path (may be NULL): a text file of earlier choices keyed
 * by device, table contents, N, probe_turns and flags; a hit is applied without timing,
 * a miss is appended. */
int tp_autotune(tp_ctx* ctx, const tp_table* t, const tp_particles* p, int32_t probe_turns, uint32_t flags,
                double target_launch_ms, const char* cache_path, tp_launch_cfg* out);

/* The same for a batch: block and forward register budget timed on the batched forward
 * with turn counts probe_turns[v] (and, with TP_TUNE_ADJOINT, the reverse budget on the
 * batched adjoint with checkpoint stride ckpt_stride); the particle sets are restored
 * after every probe.  Applies the choice with tp_batch_set_launch. */
int tp_batch_autotune(tp_ctx* ctx, tp_batch* b, const int32_t* probe_turns, int32_t ckpt_stride, uint32_t flags,
                      const char* cache_path, tp_launch_cfg* out);

/* Jacobian of nturns turns, d(final coordinates) / d(initial coordinates), of every particle
 * of p: jac[36 q + 6 i + j] = d out_i / d in_j for particle q (row-major 6 x 6 per particle;
 * zero for particles lost during tracking).  The exact derivative of the tracked map, by six
 * reverse passes with the unit seeds e_i through the adjoint machinery of tp_adjoint -- the
 * GPU replacement of the finite-difference one_turn_map / findm66 (src/optics.jl:110-121,
 * 184-215), which the host-side optics functions (gettune, getchrom, find_closed_orbit_*)
 * build on.  p holds the initial beam on entry and the tracked beam on return.
 * ckpt_stride, flags and the supported tables as for tp_adjoint. */
int tp_jacobian(tp_ctx* ctx, const tp_table* t, tp_particles* p, int32_t turn0, int32_t nturns,
                int32_t ckpt_stride, uint32_t flags, double* jac);

#ifdef __cplusplus
}
#endif

#endif /* TRACKPAD_HIP_H */
