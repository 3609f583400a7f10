/* voce_batch.h -- many RVEs on one GPU, launch tuning, request serving.
 *
 * No counterpart in the reference, which advances one RVE per MPI job
 * (evp.c:7-32).  Every member of a batch is a full voce_ctx: the
 * single-RVE entry points of voce.h (voce_get_field, voce_write_fields,
 * voce_set_diagnostics) work on it unchanged.
 */
#ifndef VOCE_BATCH_H
#define VOCE_BATCH_H

#include <stddef.h>
#include <stdint.h>

#include "voce.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- FFT launch tuning ------------------------------------------------ */

/* Launch choices of the three FFT kernels for one grid size and precision:
 * plane-kernel variant (0: 64 threads / planes <= 32x32, 1: 256 / <= 64x64,
 * 2: 512 / <= 128x128, 3: 320 / <= 64x64, 4: line-parallel, square planes
 * of 16, 32, 48, 64 or 128 only, 5: split z and y passes, any plane up to
 * 512x512, 6: split passes with line-parallel transforms, 256x256 planes
 * only), x-pass variant (0: 256 threads,
 * 1: 128, 2: 64, 3: line-parallel, one wave per column (tile 2 or 4 waves),
 * Nx = 16, 32 or 64 only, 4: staged through LDS, a block per whole, half or
 * quarter 128-byte column block, Nx = 16, 32, 48, 64, 128 or 256 only, a
 * half or a quarter at 256),
 * x-pass tile (columns per block), and whether radix-8 stages are used in
 * the plane and x passes (variants 3 and 4 have fixed radices).  us_* are
 * the measured launch times (microseconds) of the chosen plane pair
 * (forward + inverse) and x pass. */
typedef struct voce_tune_choice {
    int plv, xv, tk, pl_r8, x_r8;
    double us_plane, us_x;
} voce_tune_choice;

/* Time every valid choice on this context's own buffers (state is saved
 * and restored; call between increments), keep the fastest, apply it to
 * the context and return it in *out (may be NULL).  Results are cached per
 * (nx, ny, nz, precision) for the process; with VOCE_TUNE_FILE set in the
 * environment they are also read from and appended to that text file. */
int voce_autotune(voce_ctx *ctx, voce_tune_choice *out);
/* Apply a known choice (e.g. from a file) without timing. */
int voce_apply_tuning(voce_ctx *ctx, const voce_tune_choice *choice);
/* The context's current launch choice (us_* are 0). */
int voce_get_tuning(const voce_ctx *ctx, voce_tune_choice *out);

/* ---- batches ---------------------------------------------------------- */

typedef struct voce_batch voce_batch;

/* n RVEs, one voce_setup each (grid sizes may differ; all must share one
 * precision).  tune = 1 times every FFT launch choice on each group of
 * equal-size members as a whole (all of its blocks in one launch, as the
 * iterations run) and keeps the fastest: throughput, where voce_autotune
 * measures one RVE's latency.  Returns 0 or the first member's error code. */
int voce_batch_create(const voce_setup *setups, int n, int tune, voce_batch **out);
/* Advance every member by one strain increment.  Iterations of all members
 * run in the same kernels (block maps over the members), replayed from a
 * captured HIP graph; converged members drop out of the work while the
 * rest iterate.  reports: n entries, or NULL. */
int voce_batch_increment(voce_batch *b, voce_incr_report *reports);
int voce_batch_size(const voce_batch *b);
/* Member i as a context (owned by the batch: do not destroy). */
voce_ctx *voce_batch_member(voce_batch *b, int i);
/* out[0..7]: members, groups (distinct grid sizes), total voxels, graph
 * replays, graph instantiations, FFT iterations launched, last increment's
 * largest iteration count, bytes of device memory held */
int voce_batch_info(const voce_batch *b, long out[8]);
void voce_batch_destroy(voce_batch *b);

/* ---- macroscopic velocity gradient per increment ------------------------ */

/* Velocity gradient L (row-major 3x3, 1/s) for the context's next
 * increment: that increment targets DisGradAvg_t + L dt (the predictor of
 * evolution.c:460 with the new L) and uses skew(L) as the applied spin. */
int voce_set_velgrad(voce_ctx *ctx, const double L[9]);

/* ---- request server (FE^2-style) ------------------------------------------ */

/* Strain-increment requests against resident RVEs.  Each registered id keeps
 * its state (stress, strain, plastic strain, CRSS, orientation, ...) on the
 * device between requests.  A request is admitted when its id is idle, an
 * in-flight slot is free and its per-increment workspace (spectra, partial
 * sums, pre-update stress) fits in the workspace pool; admitted increments
 * iterate together in the same kernels (block maps over the in-flight set),
 * new ones join between iteration chunks and finished ones leave.  An
 * audited request also runs, from a copy of the same state, in FP64 and
 * reports the macro-stress difference; the id's own state is not affected. */
typedef struct voce_server voce_server;

typedef struct voce_server_opts {
    int precision;            /* VOCE_MIXED or VOCE_FP64 for served increments */
    int max_rves;             /* ids that can be registered (table capacity) */
    int max_inflight;         /* increments iterating at once (0: memory-bound only) */
    size_t workspace_bytes;   /* workspace pool (0: an eighth of the free device memory, at most 16 GiB) */
    int tune;                 /* tune each grid size when first registered (timed for
                                 throughput on a temporary batch of copies of that RVE) */
    int chunk;                /* most FFT iterations launched between polls (0: 16); a round
                                 launches up to the nearest predicted end of an increment in
                                 flight (each id's previous iteration count), at least 2 */
} voce_server_opts;

typedef struct voce_result {
    long tag;                 /* as submitted */
    int id;
    int status;               /* 0 or an error code */
    int audited;              /* 1: the increment was also run in FP64 */
    voce_incr_report rep;     /* macro results (err_e/err_s/svm_it are NULL) */
    double audit_sig_rel;     /* audited: max|sig_avg - sig_avg_fp64| / max|sig_avg_fp64| */
    double t_submit, t_admit, t_done;   /* CLOCK_MONOTONIC seconds; t_done: result on the host */
} voce_result;

int voce_server_create(const voce_server_opts *opts, voce_server **out);
/* Register id (0 <= id < max_rves) with its RVE: builds its resident state. */
int voce_server_add_rve(voce_server *s, int id, const voce_setup *setup);
/* Queue one increment of id under velocity gradient L (row-major 3x3, 1/s).
 * VOCE_ERR_NOMEM: the id's per-increment workspace is larger than the
 * whole workspace pool, so it could never be admitted. */
int voce_server_submit(voce_server *s, int id, const double L[9], int audit, long tag);
/* One scheduling round: admit, iterate one chunk, retire finished
 * increments.  *inflight (may be NULL) receives the number of increments
 * still in flight.  Returns 0 or an error code. */
int voce_server_step(voce_server *s, int *inflight);
/* Move up to max finished results (completion order) into out; returns the count. */
int voce_server_poll(voce_server *s, voce_result *out, int max);
/* out[0..9]: registered ids, resident bytes, workspace pool bytes, pool bytes
 * in use, in flight, queued, completed, audits, steps, FFT iterations launched */
int voce_server_info(const voce_server *s, long out[10]);
/* Failure-handling test hook (no reference counterpart): with kind
 * VOCE_FAULT_SHADOW_ALLOC, the `after`-th next allocation of an audit's FP64
 * shadow reports VOCE_ERR_NOMEM (0 switches the hook off).  The failing
 * voce_server_step returns the error with the request back in the queue in
 * its place and nothing of it held; increments admitted before it in the
 * same round proceed. */
#define VOCE_FAULT_SHADOW_ALLOC 1
int voce_server_inject_fault(voce_server *s, int kind, int after);
/* The id's context (fields of its resident state; owned by the server). */
voce_ctx *voce_server_rve(voce_server *s, int id);
void voce_server_destroy(voce_server *s);

/* ---- synthetic microstructures ------------------------------------------ */

/* Periodic Voronoi grains on an nx*ny*nz grid (x slowest): seeds uniform in
 * [0,N)^3 from SplitMix64(seed), voxel centres at (i+0.5, j+0.5, k+0.5),
 * minimum-image distance, ties to the lowest seed id; then one Bunge triple
 * per grain, uniform on SO(3) (phi1 = 360u, Phi = acos(2u-1)*180/pi,
 * phi2 = 360u), rounded as "%.6f" prints it.  grain: V ints (1-based);
 * euler_by_grain: 3 * n_grains doubles (degrees). */
int voce_voronoi(const int dims[3], int n_grains, uint64_t seed, int *grain, double *euler_by_grain);
/* voce_load_case with the mesh generated instead of read: dims (overriding
 * the namelist's CellDim) and voce_voronoi(n_grains, seed), all phase 1. */
int voce_load_case_voronoi(const char *input_path, const char *base_dir, const int dims[3], int n_grains,
                           uint64_t seed, voce_case *out);

#ifdef __cplusplus
}
#endif
#endif
