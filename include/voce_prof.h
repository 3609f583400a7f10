/*
 * voce_prof.h -- kernel timing for measurement tools.
 *
 * No counterpart in the reference (it has no timers, see evolution.c:344-348
 * for its only progress output).  For benchmarks that must report a
 * kernel's own average duration: each kernel of every k-th FFT iteration is
 * launched with a pair of HIP events bound to its own dispatch
 * (hipExtLaunchKernelGGL start/stop events), so a sample is the kernel's
 * execution from its start to its end, without launch gaps (k chosen at
 * enable; a profiled dispatch costs a little queue time, so a sparse
 * sample keeps the timed run close to an untimed one).  The sampled
 * residue shifts by one per increment, so k consecutive increments sample
 * every iteration index once.  Only iterations that did work (before
 * convergence) are accumulated.
 */
#ifndef VOCE_PROF_H
#define VOCE_PROF_H

#include "voce.h"

#ifdef __cplusplus
extern "C" {
#endif

/* kernels of one FFT iteration, in launch order */
#define VOCE_PROF_PLANE_FWD 0   /* z r2c + y FFT of the stress, and (one extra block)
                                   the previous iteration's finaliser */
#define VOCE_PROF_XGREEN 1      /* x FFT, Green operator, inverse x FFT */
#define VOCE_PROF_PLANE_INV 2   /* inverse y FFT, z c2r, strain update */
#define VOCE_PROF_NR 3          /* per-voxel Newton-Raphson and block partial sums */
#define VOCE_PROF_FIN 4         /* reserved: 0 (the finaliser runs inside the next
                                   forward-plane launch) */
#define VOCE_PROF_NK 5

/* Timing off (0), or on for every on-th FFT iteration of each increment
 * (1: every iteration); either way the totals are reset. */
int voce_prof_enable(voce_ctx *ctx, int on);
/* Totals since the last enable: ms_total[k] milliseconds over count[k] launches. */
int voce_prof_read(voce_ctx *ctx, double ms_total[VOCE_PROF_NK], long count[VOCE_PROF_NK]);
/* out[0..5]: voxels, nx, ny, nz, precision, two-CRSS flag */
int voce_ctx_info(const voce_ctx *ctx, long out[6]);

#ifdef __cplusplus
}
#endif
#endif
