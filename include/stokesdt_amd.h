/*
 * stokesdt_amd.h -- C ABI of the MI355X (gfx950) Brownian-dynamics step.
 *
 * Drop-in boundary for the xing-liu/stokesdt hot path. Every entry point
 * returns an int status (SDT_OK = 0, negative on error) and leaves a message
 * for sdt_last_error(). Vectors use the reference conventions: positions are
 * AoS [x0 y0 z0 x1 ...] (3 x npos doubles), radii npos doubles, blocks of
 * vectors are column-major 3*npos x nrhs with an explicit leading dimension.
 * Functions without the _dev suffix take host pointers; _dev variants take
 * device pointers (hipMalloc'd) and keep the data on the GPU.
 *
 * Citations are file:line in the reference source tree.
 */
#ifndef STOKESDT_AMD_H_
#define STOKESDT_AMD_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SDT_OK 0
#define SDT_EINVAL (-1)     /* invalid argument */
#define SDT_ESTATE (-2)     /* call out of order */
#define SDT_ENOMEM (-3)     /* device allocation failed */
#define SDT_ECONVERGE (-4)  /* Lanczos did not converge (result still written) */
#define SDT_EHIP (-5)       /* HIP runtime error */
#define SDT_ECAP (-6)       /* capacity exceeded */

#define SDT_FP64 0          /* all arithmetic in FP64 */
#define SDT_FP32 1          /* spread, FFT, influence, interp and real space in
                               FP32 with FP32 accumulation; vectors, Krylov
                               basis and small dense algebra in FP64 */

typedef struct sdt_mob_s* sdt_mob_t;
typedef struct sdt_brwn_s* sdt_brwn_t;
typedef struct sdt_pairlist_s* sdt_pairlist_t;
typedef struct sdt_force_s* sdt_force_t;
typedef struct sdt_bd_s* sdt_bd_t;
typedef struct sdt_simbox_s* sdt_simbox_t;

/* Thread-local message of the last failing call on this thread. */
const char* sdt_last_error(void);
/* Library version string. */
const char* sdt_version(void);

/* ------------------------------------------------------------------ MobSpme
 * Replaces MobSpme (src/mob/mob_spme.h:21-125): ctor
 * MobSpme(npos, rdi, box_size, xi, rmax, dim, porder) src/mob/mob_spme.cc:19-39,
 * Init :53-132, Update :205-216, MulVector :219-236, RealMulVector :239-249,
 * RecipMulVector :252-262, dim() src/mob/mob_spme.h:94.
 * precision: SDT_FP64 or SDT_FP32. The handle owns device state and one HIP
 * stream; a handle is not re-entrant, distinct handles may run concurrently.
 */
int sdt_mob_spme_create(int npos, const double* rdi, double box_size, double xi, double rmax,
                        int dim, int porder, int precision, sdt_mob_t* out);
int sdt_mob_spme_init(sdt_mob_t mob);
int sdt_mob_spme_update(sdt_mob_t mob, const double* pos, const double* rdi);
int sdt_mob_spme_update_dev(sdt_mob_t mob, const double* d_pos, const double* d_rdi);
/* v := alpha * M * f + beta * v (src/mob/mob_base.h:52-69) */
int sdt_mob_spme_mulvector(sdt_mob_t mob, int num_rhs, double alpha, int ldf, const double* f,
                           double beta, int ldv, double* v);
int sdt_mob_spme_mulvector_real(sdt_mob_t mob, int num_rhs, double alpha, int ldf,
                                const double* f, double beta, int ldv, double* v);
int sdt_mob_spme_mulvector_recip(sdt_mob_t mob, int num_rhs, double alpha, int ldf,
                                 const double* f, double beta, int ldv, double* v);
/* mode: 0 full, 1 real, 2 recip; device pointers */
int sdt_mob_spme_mulvector_dev(sdt_mob_t mob, int mode, int num_rhs, double alpha, int ldf,
                               const double* d_f, double beta, int ldv, double* d_v);
int sdt_mob_spme_dim(sdt_mob_t mob, int* dim);
/* number of real-space pairs (3x3 blocks, self included) of the last update */
int sdt_mob_spme_real_nnz(sdt_mob_t mob, int64_t* nnz);
int sdt_mob_spme_destroy(sdt_mob_t mob);

/* ----------------------------------------------------------------- MobEwald
 * Replaces MobEwald(npos, rdi, box_size, tol[, xi]) (src/mob/mob_ewald.h,
 * ctor src/mob/mob_ewald.cc:19-50, Init :53-110, Update :112-130,
 * MulVector through MobMatrix :132-152, GetMatrix): the dense 3npos x 3npos
 * Ewald-summed RPY matrix (EwaldKernel, src/mob/ewald.cc:86-330, truncated
 * at tol) stored on the device and applied by a GEMM; FP64, npos <= 30000.
 * xi <= 0 selects the reference's 10^(1/6) sqrt(pi) / L (mob_ewald.cc:51).
 * The handle is a mobility like an SPME handle: sdt_mob_init / _update /
 * _mulvector / _dim / _destroy below (or the sdt_mob_spme_ names of the same
 * MobBase methods) and the Brownian solvers accept either kind; the
 * real/recip split and the real-pair count exist only for SPME handles. */
int sdt_mob_ewald_create(int npos, const double* rdi, double box_size, double tol, double xi,
                         sdt_mob_t* out);
/* MobMatrix::GetMatrix: the symmetric matrix, column-major, ldm >= 3npos */
int sdt_mob_ewald_get_matrix(sdt_mob_t mob, int ldm, double* mat);
/* MobBase (src/mob/mob_base.h:29-69) on any mobility handle */
int sdt_mob_init(sdt_mob_t mob);
int sdt_mob_update(sdt_mob_t mob, const double* pos, const double* rdi);
int sdt_mob_mulvector(sdt_mob_t mob, int num_rhs, double alpha, int ldf, const double* f,
                      double beta, int ldv, double* v);
int sdt_mob_dim(sdt_mob_t mob, int* dim);
int sdt_mob_destroy(sdt_mob_t mob);

/* One-shot SPME matvec with the semantics of the MATLAB entry point
 * v = rpy_spme_mex(pos, rdi, L, f, mode, xi, rmax, dim, porder)
 * (mex/rpy_spme_mex.cc:8-75): alpha = 1, beta = 0, mode "full", "real" or
 * "recip"; FP64. sdt_rpy_spme_ex adds the precision choice. */
int sdt_rpy_spme(const double* pos, const double* rdi, int npos, double L, const double* f,
                 int nrhs, int ldf, const char* mode, double xi, double rmax, int dim, int porder,
                 double* v, int ldv);
int sdt_rpy_spme_ex(const double* pos, const double* rdi, int npos, double L, const double* f,
                    int nrhs, int ldf, const char* mode, double xi, double rmax, int dim,
                    int porder, int precision, double* v, int ldv);

/* Dense Ewald-summed RPY matrix with the semantics of the MATLAB entry point
 * mat = rpy_ewald_matrix_mex(pos, rdi, L, tol, mode, xi)
 * (mex/rpy_ewald_matrix_mex.cc:10-64 over MobDebug::GetMatrix,
 * src/mob/mob_debug.cc, and EwaldKernel, src/mob/ewald.cc:86-330): pos AoS
 * 3 x npos, mode "full", "real" or "recip", tol the truncation tolerance of
 * both sums; xi <= 0 selects the reference's default 10^(1/6) sqrt(pi) / L.
 * mat receives the symmetric 3npos x 3npos matrix (leading dimension
 * ldm >= 3npos); FP64 on the device, npos <= 30000. */
int sdt_rpy_ewald_matrix(const double* pos, const double* rdi, int npos, double L, double tol,
                         const char* mode, double xi, double* mat, int ldm);

/* EwaldVectorKernel (src/mob/ewald.cc:333-536), v = M f, on a sample of
 * rows: the Ewald-summed RPY mobility to tolerance tol (the tables and
 * truncation of sdt_rpy_ewald_matrix) applied to f (3 npos values) and
 * returned for the particles rows[0 .. nrows-1] (v: 3 nrows values), with
 * no limit on npos: the matrix is never formed and the reciprocal sum runs
 * over structure factors of f. xi <= 0 picks the splitting that makes the
 * sample cheapest (the sum depends on xi only within tol); *xi_used (may be
 * NULL) receives the one used. FP64 on the device; host arrays. */
int sdt_rpy_ewald_rows(const double* pos, const double* rdi, int npos, double L, double tol,
                       double xi, int nrows, const int* rows, const double* f, double* v,
                       double* xi_used);

/* Diffusion coefficients of tools/calc_diff (calc_diff.cc:100-140): pos
 * holds nframes > 2 frames of npos AoS positions, frame after frame, dt the
 * time between frames. For lag i = 1 .. nframes/2 - 1,
 * D[i-1] = sum over origins j < nframes/2 and particles of
 * |r(j+i) - r(j)|^2 / (6 npos (nframes/2) i dt), the value calc_diff prints
 * (every particle counts with radius 1, as there). *nlags receives
 * nframes/2 - 1; D must hold that many values. FP64 on the device. */
int sdt_diffusion(const double* pos, int nframes, int npos, double dt, double* D, int* nlags);
/* No counterpart in the reference: sdt_diffusion for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops). d_pos and d_D are device pointers on the current device; the work
 * is enqueued on stream (a hipStream_t, NULL for the default stream) and the
 * call returns without waiting. No hipMalloc / hipFree: the partial sums live
 * in a static device array of 2^20 doubles (8 MiB, held on every GPU the
 * library runs on until the process ends), so ceil(npos / 256) *
 * (nframes / 2 - 1) must not exceed 2^20 (else SDT_ECAP). Calls that
 * overlap on one GPU share that array and race. */
int sdt_diffusion_dev_noalloc(const double* d_pos, int nframes, int npos, double dt, double* d_D,
                              int* nlags, void* stream);

/* ------------------------------------------------------------- BrwnRandom
 * Replaces BrwnRandom(dim, num_vecs, max_nrhs) (src/brwn/brwn_random.h:16-86),
 * Init (src/brwn/brwn_random.cc:42-116) and Compute(mob, nrhs, ldz, z, ldy,
 * y) (:118-263): with a mobility handle, num_vecs Gaussian vectors (the
 * Philox stream of seed, one stream index per sampling; the reference
 * draws from MKL with seed 5678) span V = orth(M^3 Omega); y = V B^{1/2}
 * V^T z + alpha (I - V V^T) z with B = V^T M V and alpha making
 * |y|^2 = z^T M z, per column. mob = NULL reuses the last sampling.
 * num_vecs <= min(dim, 128). Column-major z, y (ldz, ldy >= dim). */
typedef struct sdt_brwn_random_s* sdt_brwn_random_t;
int sdt_brwn_random_create(int dim, int num_vecs, int max_nrhs, uint64_t seed,
                           sdt_brwn_random_t* out);
int sdt_brwn_random_init(sdt_brwn_random_t br);
int sdt_brwn_random_compute(sdt_brwn_random_t br, sdt_mob_t mob, int num_rhs, int ldz,
                            const double* z, int ldy, double* y);
int sdt_brwn_random_compute_dev(sdt_brwn_random_t br, sdt_mob_t mob, int num_rhs, int ldz,
                                const double* d_z, int ldy, double* d_y);
int sdt_brwn_random_destroy(sdt_brwn_random_t br);

/* --------------------------------------------------------------- PairList
 * Replaces PairListR / PairListS (src/pairlist/pair_base.h:30-69, pair_r.h,
 * pair_s.h): mode 'R' (min-image r <= cutoff) or 'S' (2r/(a_i+a_j) <
 * cutoff); self pairs included. Init returns the estimated nnz
 * (src/pairlist/pair_base.cc:34-71), Build the exact nnz (:118-123),
 * GetPairs fills caller buffers of npos+1 and nnz ints, rows sorted
 * (:126-140). Sets are bit-identical to the reference's. */
int sdt_pairlist_create(int npos, const double* rdi, double box_size, double cutoff, char mode,
                        sdt_pairlist_t* out);
int sdt_pairlist_init(sdt_pairlist_t pl, int64_t* est_nnz);
int sdt_pairlist_build(sdt_pairlist_t pl, const double* pos, int64_t* nnz);
int sdt_pairlist_get_pairs(sdt_pairlist_t pl, int32_t* rowptr, int32_t* colidx);
int sdt_pairlist_destroy(sdt_pairlist_t pl);
/* One-shot [ia, ja] = pairlist_mex(pos, rdi, L, cutoff, mode)
 * (mex/pairlist_mex.cc:9-80): 0-based CSR. If ja_cap < nnz, or ia / ja
 * are null (a size query), only *nnz is set and SDT_ECAP is returned (call
 * again with buffers of npos + 1 and nnz). */
int sdt_pairlist(const double* pos, const double* rdi, int npos, double L, double cutoff,
                 char mode, int32_t* ia, int32_t* ja, int64_t ja_cap, int64_t* nnz);

/* ------------------------------------------------------------ BrwnLanczos
 * Replaces BrwnLanczos(dim, max_iters, max_nrhs, tol)
 * (src/brwn/brwn_lanczos.h:45-46), Init (src/brwn/brwn_lanczos.cc:224-275)
 * and Compute(mob, nrhs, ldz, z, ldy, y) (:278-304): y ~= M^{1/2} z by block
 * Lanczos in groups of max_nrhs (max 16) columns. Returns SDT_ECONVERGE if
 * a group did not converge within max_iters (y is still written, as the
 * reference does). ldy is honoured. */
int sdt_brwn_lanczos_create(int dim, int max_iters, int max_nrhs, double tol, sdt_brwn_t* out);
int sdt_brwn_lanczos_init(sdt_brwn_t br);
int sdt_brwn_lanczos_compute(sdt_brwn_t br, sdt_mob_t mob, int num_rhs, int ldz, const double* z,
                             int ldy, double* y);
int sdt_brwn_lanczos_compute_dev(sdt_brwn_t br, sdt_mob_t mob, int num_rhs, int ldz,
                                 const double* d_z, int ldy, double* d_y);
/* iteration index at convergence of each group of the last compute (-1 if
 * not converged), like the reference's return value */
int sdt_brwn_lanczos_iters(sdt_brwn_t br, int* iters, int cap, int* ngroups);
/* 0: no reorthogonalisation (reference); 1: full FP64 reorthogonalisation */
int sdt_brwn_lanczos_set_reorth(sdt_brwn_t br, int reorth);
int sdt_brwn_lanczos_destroy(sdt_brwn_t br);

/* --------------------------------------------------------------- BrwnChol
 * Replaces BrwnChol(dim) (src/brwn/brwn_chol.h), Init
 * (src/brwn/brwn_chol.cc:29-49) and Compute(mob, nrhs, ldz, z, ldy, y)
 * (:52-83): y = L z with M = L L^T the lower Cholesky factor of a matrix
 * mobility (a dense Ewald handle; other handles give SDT_EINVAL), factored
 * on every call as the reference does. SDT_ECONVERGE if M is not positive
 * definite. FP64 on the device, dim <= 90000. */
typedef struct sdt_brwn_chol_s* sdt_brwn_chol_t;
int sdt_brwn_chol_create(int dim, sdt_brwn_chol_t* out);
int sdt_brwn_chol_init(sdt_brwn_chol_t br);
int sdt_brwn_chol_compute(sdt_brwn_chol_t br, sdt_mob_t mob, int num_rhs, int ldz,
                          const double* z, int ldy, double* y);
int sdt_brwn_chol_compute_dev(sdt_brwn_chol_t br, sdt_mob_t mob, int num_rhs, int ldz,
                              const double* d_z, int ldy, double* d_y);
int sdt_brwn_chol_destroy(sdt_brwn_chol_t br);

/* ------------------------------------------------------------ StericForce
 * Replaces StericForce(npos, rdi, box_size, r0, k0)
 * (src/force/force_steric.cc:56-69), Init (:77-96), ForceBase::Compute
 * (src/force/force_base.cc:24-28: zero then accumulate) and Accumulate
 * (src/force/force_steric.cc:99-112). */
int sdt_steric_create(int npos, const double* rdi, double box_size, double r0, double k0,
                      sdt_force_t* out);
int sdt_steric_init(sdt_force_t fo);
int sdt_steric_compute(sdt_force_t fo, const double* pos, const double* rdi, double* f);
int sdt_steric_accumulate(sdt_force_t fo, const double* pos, const double* rdi, double* f);
int sdt_steric_destroy(sdt_force_t fo);
/* One-shot f = force_steric_mex(pos, rdi, L, r0, k0)
 * (mex/force_steric_mex.cc:8-50): pos AoS 3 x npos, f 3 x npos (zeroed,
 * then the steric forces); L, r0 and k0 must be positive. */
int sdt_force_steric(const double* pos, const double* rdi, int npos, double L, double r0,
                     double k0, double* f);

/* --------------------------------------------------------------- BD step
 * BDSimBox::Advance (src/simbox/simbox_bd.cc:296-366) with SPME mobility,
 * block Lanczos and steric forces, positions resident on the device. Every
 * mob_interval steps a Gaussian block is drawn (Philox-4x32-10, stream =
 * step / mob_interval), the mobility is updated and D = M^{1/2} Z computed;
 * every step disp = dt M f + sqrt(2 dt) D[:, step % mob_interval] and
 * pos += disp (kT = 1). */
int sdt_bd_create(int npos, const double* pos, const double* rdi, double box_size, double xi,
                  double rmax, int dim, int porder, int precision, int max_iters, int max_nrhs,
                  double tol, double steric_r0, double steric_k0, double dt, int mob_interval,
                  uint64_t seed, sdt_bd_t* out);
int sdt_bd_init(sdt_bd_t bd);
int sdt_bd_advance(sdt_bd_t bd, int nsteps);
int sdt_bd_get_positions(sdt_bd_t bd, double* pos);
int sdt_bd_last_iters(sdt_bd_t bd, int* iters, int cap, int* ngroups);
int sdt_bd_destroy(sdt_bd_t bd);

/* Tails of the BD step as one replayed HIP graph per mobility interval
 * (default on): the forces, M f and position update of a step are captured
 * once per interval and replayed; results are bit-identical to eager launches
 * (off). No counterpart in the reference. */
int sdt_bd_set_graphs(sdt_bd_t bd, int on);
/* The project autotuner on a BD handle (after init; no counterpart in the
 * reference): picks the SPME launch choices (FFT radix order, interp strip
 * and chunk, spread stage) for single-column and Lanczos-batch launches by
 * timing them on the current positions; tune_cache (may be NULL) as for
 * sdt_ensemble_init. choices[8] and ms[4] as sdt_ensemble_tuning. */
int sdt_bd_autotune(sdt_bd_t bd, const char* tune_cache, int* choices, double* ms);

/* -------------------------------------------------------------- Ensemble
 * An ensemble of independent BD systems of uneven size, box and mesh
 * advanced together (no counterpart in the reference, whose BDSimBox
 * advances one system, src/simbox/simbox_bd.cc:296-366; each member is that
 * step). System s has npos[s] particles (positions and radii of all systems
 * one after another in pos / rdi), box box_size[s], Ewald xi[s], real-space
 * cutoff rmax[s] and mesh dim[s]; porder, precision, Lanczos, steric, dt and
 * mob_interval are shared; system s draws its Gaussian blocks from seed + s.
 * Each mobility interval's updates and Lanczos solves run side by side (one
 * HIP stream per system); every other step is four HIP graph launches
 * (lanes of systems) holding every system's step tail. init with autotune != 0 picks each system's SPME
 * launch choices (FFT radix order, interp strip and chunk, spread stage) by
 * timing them on the system, for single-column and Lanczos-batch launches;
 * tune_cache (may be NULL) names a text file of earlier choices.
 * advance writes each step's host wall time in ms to step_ms (may be NULL).
 * sdt_ensemble_tuning: choices[8] = {fft_order, interp_sxt, interp_ch,
 * spread_stage_kb} narrow then wide (0 = heuristic); ms[4] = default and
 * tuned matvec times, narrow then wide. */
typedef struct sdt_ensemble_s* sdt_ensemble_t;
int sdt_ensemble_create(int nsys, const int* npos, const double* pos, const double* rdi,
                        const double* box_size, const double* xi, const double* rmax,
                        const int* dim, int porder, int precision, int max_iters, int max_nrhs,
                        double tol, double steric_r0, double steric_k0, double dt,
                        int mob_interval, uint64_t seed, sdt_ensemble_t* out);
int sdt_ensemble_init(sdt_ensemble_t e, int autotune, const char* tune_cache);
int sdt_ensemble_set_graphs(sdt_ensemble_t e, int on);
int sdt_ensemble_advance(sdt_ensemble_t e, int nsteps, double* step_ms);
int sdt_ensemble_get_positions(sdt_ensemble_t e, int sys, double* pos);
int sdt_ensemble_tuning(sdt_ensemble_t e, int sys, int* choices, double* ms);
int sdt_ensemble_last_iters(sdt_ensemble_t e, int sys, int* iters, int cap, int* ngroups);
int sdt_ensemble_destroy(sdt_ensemble_t e);

/* --------------------------------------------------------------- Serving
 * A request loop for BD jobs (no counterpart in the reference, whose only
 * run loop is bin/main/main.cc:129-158): continuous batching on one GPU.
 * kind: 0 optimised FP64, 1 FP32-mixed, 2 plain FP64 path. A job is a
 * system (npos particles, box, xi, rmax, mesh dim) asking for nsteps BD
 * steps (the step of src/simbox/simbox_bd.cc:296-366 with the shared
 * porder, Lanczos, steric, dt and mob_interval), its Gaussian stream from
 * seed, arriving arrival_s seconds after sdt_serve_run starts.
 * sdt_serve_run admits arrived jobs first come first served while a slot
 * (max_active) is free and the job's predicted device footprint
 * (sdt_serve_footprint) fits what is left of capacity_bytes; a job whose
 * footprint exceeds the whole capacity is refused (state 3). Every admitted
 * job steps on its own stream and host thread with its state resident on
 * the device, so jobs admitted while others run overlap with them; its
 * reservation is released when it ends. Returns when every job has ended, or
 * at deadline_s (> 0), when running jobs stop after their current step
 * (state 4) and queued ones stay queued (state 0).
 * sdt_serve_job: state (0 queued, 1 running, 2 done, 3 failed, 4 stopped),
 * times[3] = {arrival, admission, end} in seconds (-1: not yet), the steps
 * done and the reserved footprint; a failed job's message goes to
 * sdt_last_error. sdt_serve_job_positions: the positions after the last
 * step (3 npos). sdt_serve_job_step_ms: host wall time of every step.
 * sdt_serve_stats: stats[3] = {wall seconds of the run, peak reserved bytes,
 * peak jobs on the device}. */
typedef struct sdt_serve_s* sdt_serve_t;
int sdt_serve_create(int kind, double capacity_bytes, int max_active, int porder, int max_iters,
                     int max_nrhs, double tol, double steric_r0, double steric_k0, double dt,
                     int mob_interval, sdt_serve_t* out);
int sdt_serve_footprint(sdt_serve_t sv, int npos, int dim, double box_size, double rmax,
                        double* bytes);
int sdt_serve_submit(sdt_serve_t sv, int npos, const double* pos, const double* rdi,
                     double box_size, double xi, double rmax, int dim, int nsteps, uint64_t seed,
                     double arrival_s, int* job);
int sdt_serve_run(sdt_serve_t sv, double deadline_s);
int sdt_serve_job(sdt_serve_t sv, int job, int* state, double* times, int* steps_done,
                  double* footprint_bytes);
int sdt_serve_job_positions(sdt_serve_t sv, int job, double* pos);
int sdt_serve_job_step_ms(sdt_serve_t sv, int job, double* ms, int cap, int* n);
int sdt_serve_stats(sdt_serve_t sv, double* stats);
int sdt_serve_destroy(sdt_serve_t sv);

/* ---------------------------------------------------------------- SimBox
 * Replaces BDSimBox(config_file, model_file, xyz_file)
 * (src/simbox/simbox_bd.cc:31-294) with Init and Advance (:296-366), and the
 * file formats of MoleculeIO (src/util/molecule_io.cc:25-414):
 *  - config "key value" lines with the reference keys (mob-method spme, or
 *    the dense Ewald sum for any other value, as the reference; ewald-tol;
 *    brownian-method lanczos, random (SPME only) or Cholesky for any other
 *    value (dense mobility only, as the reference); spme-xi, spme-rmax,
 *    spme-nfft-grids, spme-porder, mob-interval,
 *    lanczos-maxiters/-maxnrhs/-tol, random-numvecs/-maxnrhs, delta-t,
 *    traj-interval, traj-startstep, traj-output) plus "precision fp64|fp32"
 *    (SPME) and "seed <uint64>";
 *  - model lines "par type radius charge" and "steric r0 k0" (bond lines
 *    are rejected);
 *  - XYZ: npos, "step time Lx Ly Lz [seed]", "type x y z"; the box must be
 *    cubic.
 * create parses the files on the host (no device use); init allocates the
 * device state. Trajectory frames and sdt_simbox_write_xyz keep particle
 * types and full precision and record the seed, so any frame restarts the
 * run (a start inside a mobility interval recomputes that interval's
 * displacements from the restart positions).
 * sdt_simbox_params: dparams[7] = {spme-xi, spme-rmax, lanczos-tol,
 * steric r0, steric k0, delta-t, ewald-tol}; iparams[11] = {spme-nfft-grids,
 * spme-porder, mob-interval, lanczos-maxiters, lanczos-maxnrhs,
 * traj-interval, traj-startstep, precision, seed, mobility (0 SPME, 1 dense
 * Ewald), Brownian method (0 Lanczos, 1 random, 2 Cholesky)}. */
int sdt_simbox_create(const char* config_file, const char* model_file, const char* xyz_file,
                      sdt_simbox_t* out);
int sdt_simbox_init(sdt_simbox_t sb);
int sdt_simbox_advance(sdt_simbox_t sb, int nsteps);
int sdt_simbox_npos(sdt_simbox_t sb, int* npos);
int sdt_simbox_state(sdt_simbox_t sb, int64_t* step, double* time);
int sdt_simbox_params(sdt_simbox_t sb, double* dparams, int64_t* iparams);
int sdt_simbox_get_positions(sdt_simbox_t sb, double* pos);
int sdt_simbox_get_radii(sdt_simbox_t sb, double* rdi);
int sdt_simbox_write_xyz(sdt_simbox_t sb, const char* path);
int sdt_simbox_destroy(sdt_simbox_t sb);

/* ----------------------------------------------------------------- misc */
/* Gaussian block of the BD driver (Philox-4x32-10 + Box-Muller, FP64):
 * element e of (seed, stream); replaces RndStream::Gaussian
 * (src/util/rnd_stream.cc:44-65), whose MKL stream cannot be reproduced. */
int sdt_gaussian(uint64_t seed, uint32_t stream, int64_t n, double* out);
/* Free and total memory of the current device, in bytes. */
int sdt_device_memory(double* free_bytes, double* total_bytes);
/* Block until all work of every handle finished. */
int sdt_device_synchronize(void);

#ifdef __cplusplus
}
#endif

#endif /* STOKESDT_AMD_H_ */
