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
/* Launch forms of a mobility's kernels, for tests and measurements:
 * interp_form, meshes per LDS pass of the interp kernel: 0 the built-in
 * choice (1), else 1, 2 or 4 (4: FP32 only; an error at apply time
 * otherwise); z_form, threads and pencils per workgroup of the z transform:
 * 0 the built-in choice (FP32 (512, 16), FP64 (512, 8)), 1 .. 5 =
 * (1024, 16), (1024, 16), (512, 16), (512, 8), (256, 8), (256, 4);
 * spread_form: the point-owner spread's sums in wide launches (0 the
 * built-in choice: FP32 3 + 8, FP64 1 + 8; 1 forces staged in LDS, 2 forces
 * read by scalar loads (FP32), 3 forces staged in LDS and summed by the
 * matrix cores (FP32)), plus 4 to store the meshes as whole rows through
 * the stage's LDS, plus 8 for one workgroup per mesh block over every
 * column block (FP32 with 3); every form gives the same sums;
 * fft_form: 0 the fused x + y plane kernels where a plane fits LDS, 1 the
 * separate x and y passes. */
/* Right-hand sides per launch of an SPME mobility's mulvector (default 16;
 * [1, 1024]): a wider launch runs every column block of a product in one
 * spread and interpolation, as the block-Lanczos batches do. */
int sdt_debug_mob_launch_width(sdt_mob_t mob, int width);
/* Launch choices of an initialised SPME mobility for every launch width
 * (the autotuner's parameters; 0 = built-in): FFT radix order 0 .. 2,
 * interp strip width 1, 2 or 4 tiles, interp particle chunk 32 .. 128,
 * spread candidate stage 16 .. 64 KB. */
int sdt_debug_spme_tuning(sdt_mob_t mob, int fft_order, int interp_sxt, int interp_ch,
                          int spread_stage_kb);
int sdt_debug_spme_forms(sdt_mob_t mob, int interp_form, int z_form, int spread_form,
                         int fft_form);
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
 * 3 inverse y/x FFT, 4 interp, 5 real space (on a second stream beside
 * stages 0-3 in wide full products), 6 Krylov block products, 7 steric
 * (pair list + force), 8 mobility update, 9 the column-blocked input copy of
 * wide full products (elsewhere timed with the spread). sdt_bd_set_profiling(1)
 * clears the sums; sdt_bd_profile returns sums (ms) and counts. */
int sdt_bd_set_profiling(sdt_bd_t bd, int on);
int sdt_bd_profile(sdt_bd_t bd, double* sum_ms, int64_t* counts, int nstages);
/* right-hand sides processed per stage over the recorded launches (stages 0-5
 * and 9 record their launch width; the others 0) */
int sdt_bd_profile_columns(sdt_bd_t bd, int64_t* columns, int nstages);
/* Random sequential addition of n non-overlapping spheres of radius a in a
 * periodic box of side L (seeded); returns SDT_ECAP if max_attempts ran out
 * (then *placed < n). Synthetic-input generator for benchmarks. */
int sdt_util_rsa_positions(int n, double box, double a, uint64_t seed, int64_t max_attempts,
                           double* pos, int64_t* placed);

/* Plain FP64 GPU path: the reference algorithm on the device as it is
 * structured on the CPU, the baseline of the FP32-mixed path. MobSpme
 * (src/mob/mob_spme.cc:53-262) with a per-RHS reciprocal loop
 * (src/mob/spme.cc:714-839), stored N x p^3 weight and index tables
 * (:366-509, :625-633), spread by FP64 global atomics, three R2C and three
 * C2R library FFTs per RHS, a separate influence kernel (:201-310) and
 * interpolation kernel (:163-198), and stored FP64 3x3 real-space blocks
 * (src/mob/mob_spme.cc:135-202) applied by SpMV3x3 (src/util/sparse.cc:
 * 79-153); BlockLanczos (src/brwn/brwn_lanczos.cc:85-201) group after group
 * with Householder QR and Y = V S R every iteration; the BD step of
 * src/simbox/simbox_bd.cc:296-366. FP64 only; host pointers. create also
 * runs the reference's Init checks. mode: 0 full, 1 real, 2 recip. */
typedef struct sdt_plain_mob_s* sdt_plain_mob_t;
typedef struct sdt_plain_bd_s* sdt_plain_bd_t;
int sdt_plain_mob_create(int npos, const double* rdi, double box_size, double xi, double rmax,
                         int dim, int porder, sdt_plain_mob_t* out);
int sdt_plain_mob_update(sdt_plain_mob_t mob, const double* pos);
int sdt_plain_mob_mulvector(sdt_plain_mob_t mob, int mode, int num_rhs, double alpha, int ldf,
                            const double* f, double beta, int ldv, double* v);
/* y ~= M^{1/2} z by BlockLanczos in groups of max_nrhs on the plain
 * mobility; iters[g] = iteration of convergence of group g (-1 if none). */
int sdt_plain_lanczos(sdt_plain_mob_t mob, int max_iters, int max_nrhs, double tol, int num_rhs,
                      int ldz, const double* z, int ldy, double* y, int* iters, int cap,
                      int* ngroups);
int sdt_plain_mob_destroy(sdt_plain_mob_t mob);
int sdt_plain_bd_create(int npos, const double* pos, const double* rdi, double box_size, double xi,
                        double rmax, int dim, int porder, int max_iters, int max_nrhs, double tol,
                        double steric_r0, double steric_k0, double dt, int mob_interval,
                        uint64_t seed, sdt_plain_bd_t* out);
int sdt_plain_bd_init(sdt_plain_bd_t bd);
int sdt_plain_bd_advance(sdt_plain_bd_t bd, int nsteps);
int sdt_plain_bd_get_positions(sdt_plain_bd_t bd, double* pos);
int sdt_plain_bd_last_iters(sdt_plain_bd_t bd, int* iters, int cap, int* ngroups);
/* stages as sdt_bd_profile */
int sdt_plain_bd_set_profiling(sdt_plain_bd_t bd, int on);
int sdt_plain_bd_profile(sdt_plain_bd_t bd, double* sum_ms, int64_t* counts, int64_t* columns,
                         int nstages);
int sdt_plain_bd_destroy(sdt_plain_bd_t bd);

#ifdef __cplusplus
}
#endif

#endif /* STOKESDT_AMD_DEBUG_H_ */
