/*
 * stokesdt_amd_debug.h -- test and benchmark hooks of libstokesdt_amd.
 *
 * These have no counterpart in the reference: they expose single pipeline
 * stages for parity tests and per-kernel timing with HIP events on the
 * handle's own stream.
 */
#ifndef STOKESDT_AMD_DEBUG_H_
#define STOKESDT_AMD_DEBUG_H_

#include <stdint.h>
#include "stokesdt_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* G = A^T B (b x b, row-major) for row-major n x b host arrays, on the
 * device Gram kernel (FP64 MFMA). */
int sdt_debug_gram(int64_t n, int b, const double* A, const double* B, double* G);
/* Column-group Gram: for row-major n x btot host arrays split into groups of
 * bs columns (the last may be narrower), G[g] = A_g^T B_g stored as
 * [ngroup][bs][bs] (entries beyond a narrower group are zero). */
int sdt_debug_gram_groups(int64_t n, int btot, int bs, const double* A, const double* B,
                          double* G);
/* T^{1/2} E_1 (m x kb, row-major) of a symmetric band matrix T (m x m,
 * row-major, bandwidth kd): the host routine used by block Lanczos. */
int sdt_debug_band_sqrt(int m, int kd, const double* T, int kb, double* S);
/* Spread one force vector (3*npos, reference layout) with an updated
 * mobility and return the three real meshes [c][z][y][x] (3*dim^3). */
int sdt_debug_spme_spread(sdt_mob_t mob, const double* f, double* mesh);
/* Interpolate three given real meshes [c][z][y][x] (3*dim^3) at the
 * particles of an updated monodisperse mobility: v (3*npos, reference
 * layout) = alpha * sum of the B-spline weights times the mesh values, the
 * interpolation step of the reference's recip driver (src/mob/spme.cc:
 * 163-198). */
int sdt_debug_spme_interp(sdt_mob_t mob, const double* mesh, double alpha, double* v);
/* Mean device time (ms, HIP events on the handle's stream) of each stage of
 * one packed matvec with b right-hand sides over `reps` repetitions:
 * t[0] spread, t[1] forward x/y FFT, t[2] z FFT + influence + inverse z,
 * t[3] inverse y/x FFT, t[4] interp, t[5] real space, t[6] whole matvec. */
int sdt_bench_stage_times(sdt_mob_t mob, int b, int reps, double* t);
/* reps launches of a plain device copy of n doubles (reads and writes 8 n
 * bytes each), for scaling HBM byte counters; ms_avg = mean launch time. */
int sdt_bench_copy(int64_t n, int reps, double* ms_avg);
/* Mean device time (ms) of one full BD step sequence: runs `nsteps` steps
 * bracketed by events on the handle's stream. */
int sdt_bench_bd_steps(sdt_bd_t bd, int nsteps, double* ms_total);

/* Per-stage device time of the BD driver, recorded with HIP events on the
 * handle's stream around every launch group while profiling is on.
 * Stages: 0 spread, 1 forward x/y FFT, 2 z FFT + influence + inverse z,
 * 3 inverse y/x FFT, 4 interp, 5 real space, 6 Krylov block products,
 * 7 steric (pair list + force), 8 mobility update. sdt_bd_set_profiling(1)
 * clears the sums; sdt_bd_profile returns sums (ms) and counts. */
int sdt_bd_set_profiling(sdt_bd_t bd, int on);
int sdt_bd_profile(sdt_bd_t bd, double* sum_ms, int64_t* counts, int nstages);
/* right-hand sides processed per stage over the recorded launches (stages 0-5
 * record their launch width; the others 0) */
int sdt_bd_profile_columns(sdt_bd_t bd, int64_t* columns, int nstages);
/* Random sequential addition of n non-overlapping spheres of radius a in a
 * periodic box of side L (seeded); returns SDT_ECAP if max_attempts ran out
 * (then *placed < n). Synthetic-input generator for benchmarks. */
int sdt_util_rsa_positions(int n, double box, double a, uint64_t seed, int64_t max_attempts,
                           double* pos, int64_t* placed);

#ifdef __cplusplus
}
#endif

#endif /* STOKESDT_AMD_DEBUG_H_ */
