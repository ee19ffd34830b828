// Matrix-free real-space RPY-Ewald matvec.
#pragma once
#include "sdt_internal.hpp"

namespace sdt {

// Y = beta Y + alpha * M_real X over a CSR pair list (self included) whose
// rows and columns index the same particle order as pos/rdi and the packed
// vectors X, Y (row-major 3N x b).
// Overlapping pairs (r < a_i + a_j) of a CSR pair list, built once per
// update: per-row CSR over the rows that have any (rows[]).
struct RealOverlap {
    DevBuf<int> cnt, has, rowptr, hrp, col, rows, scratch;
    long long npairs = 0;
    int nrows = 0;
    void build(int npos, const int* d_rowptr, const int* d_colidx, const double* d_pos,
               const double* d_rdi, double box, hipStream_t s);
};

// Y += alpha * (overlapping-pair blocks) X, in FP64
void real_overlap(const RealOverlap& ov, const double* d_pos, const double* d_rdi, double box,
                  double xi, int b, double alpha, const double* X, double* Y, hipStream_t s);

// without the overlapping pairs (those are added by real_overlap), in
// arithmetic T. For b >= 8 the kernel gathers a column-blocked copy of X in
// T, Xb[cb][j][c][r] with real_block_cols<T>(b) columns per block
// (zero-padded), held in scratch (real_scratch_bytes<T>(npos, b) bytes, 0
// for b < 8): written by real_block_copy, or by real_matvec itself unless
// copied is set.
template <typename T>
int real_block_cols(int b);
template <typename T>
size_t real_scratch_bytes(int npos, int b);
template <typename T>
void real_block_copy(int npos, int b, const double* X, void* scratch, hipStream_t s);
template <typename T>
void real_matvec(int npos, const int* d_rowptr, const int* d_colidx, const double* d_pos,
                 const double* d_rdi, double box, double xi, int b, double alpha, const double* X,
                 double beta, double* Y, void* scratch, bool copied, hipStream_t s);

// Dense Ewald-summed RPY matrix on the device: setup builds the image and
// k-vector tables of (L, xi, tol) for mode 0 full, 1 real or 2 recip
// (XitoRmax / XitoKmax truncation, src/mob/ewald.cc:19-83, 643-797); build
// writes the symmetric 3N x 3N matrix of device positions and radii into
// d_mat (column-major, leading dimension ld), npos <= 30000.
struct EwaldMatrixGpu {
    double box = 0, xi = 0, rmax = 0;
    int parts = 0, nr = 0, nk = 0;
    DevBuf<double> d_img, d_k;
    void setup(double box, double xi, double tol, int mode, hipStream_t s);
    void build(int npos, const double* d_pos, const double* d_rdi, long long ld, double* d_mat,
               hipStream_t s) const;
    // v_r = sum_j M(rows[r], j) f_j for nrows rows (3 values each), the
    // matrix never formed: the real sum per pair over the image table, the
    // reciprocal sum through structure factors of f; d_sk holds 12 nk doubles
    void rows_mul(int npos, const double* d_pos, const double* d_rdi, int nrows,
                  const int* d_rows, const double* d_f, double* d_sk, double* d_v,
                  hipStream_t s) const;
};

// Dense Ewald-summed RPY matrix (MobDebug::GetMatrix): mode 0 full, 1 real
// part, 2 reciprocal part; host pos (AoS) and rdi, host output 3N x 3N with
// leading dimension ldm (symmetric). Synchronises.
void ewald_matrix(int npos, const double* pos, const double* rdi, double box, double xi,
                  double tol, int mode, int ldm, double* mat, hipStream_t s);

// EwaldVectorKernel (src/mob/ewald.cc:333-536) on a sample of rows: host
// pos (AoS), rdi, rows (nrows indices) and f (3 npos); v receives 3 nrows
// values of M f. xi <= 0 picks the splitting that balances the real-space
// pairs of the sampled rows against the structure factors over all npos
// particles (the sum does not depend on xi beyond tol). Synchronises.
void ewald_rows(int npos, const double* pos, const double* rdi, double box, double xi, double tol,
                int nrows, const int* rows, const double* f, double* v, hipStream_t s);
double ewald_rows_xi(int npos, int nrows, double box, double tol);

}  // namespace sdt
