// Tall-skinny FP64 block kernels and vector layout conversion.
#pragma once
#include "sdt_internal.hpp"

namespace sdt {

// G (16 x 16 tile holding b x b, row-major, device) = A^T B for packed
// n x b blocks A, B (b <= 16)
void gram(long long n, int b, const double* A, const double* B, double* d_G, DevBuf<double>& part,
          hipStream_t s);
// Column groups of a packed n x btot block (leading dimension btot): group g
// holds columns [g bs, min((g + 1) bs, btot)). d_G[g][16][16] = A_g^T B_g.
void gram_groups(long long n, int btot, int bs, const double* A, const double* B, double* d_G,
                 DevBuf<double>& part, hipStream_t s);
// C = beta C + alpha [A_0 .. A_{nblk-1}] S, blocks n x b at stride blk_stride,
// S row-major (nblk*b) x b on the device
void tsmm(long long n, int b, int nblk, const double* A, long long blk_stride, const double* d_S,
          double alpha, double beta, double* C, hipStream_t s);
// per group: C_g = beta C_g + alpha sum_k A_{k,g} S[g][k] with S laid out
// [ngroup][nblk][bs][bs] (rows/columns beyond a narrower last group unused)
void tsmm_groups(long long n, int btot, int bs, int nblk, const double* A, long long blk_stride,
                 const double* d_S, double alpha, double beta, double* C, hipStream_t s);
// the same with one coefficient block per group (nblk = 1); C may alias A
void tsmm_groups1(long long n, int btot, int bs, const double* A, const double* d_S, double alpha,
                  double beta, double* C, hipStream_t s);
// sum of squares (synchronises)
double sumsq(long long n, const double* x, DevBuf<double>& part, hipStream_t s);
// reference layout (column-major 3N x b, leading dimension ld, original
// particle order) <-> packed layout (row-major 3N x b, sorted order)
void pack(int npos, int b, const int* perm, int ldf, const double* f, double* X, hipStream_t s);
void unpack(int npos, int b, const int* perm, int ldv, const double* X, double* v, hipStream_t s);
// columns cols[0 .. nc) of a packed block (nrow rows, leading dimension ld)
// to a compact nrow x nc block (to_compact), or back
void column_subset(long long nrow, int ld, int nc, const int* d_cols, const double* src,
                   double* dst, bool to_compact, hipStream_t s);
// Device-side steps of the lockstep block Lanczos iteration, one column
// group (bs <= 16) per workgroup. The per-group record of an iteration holds
// kLzRec doubles: [0, 256) H_i (raw 16 x 16 Gram tile), [256, 512) B_i
// (upper, 16 x 16), [512] 1 if the CholQR succeeded. Groups with mask 0 get
// zero coefficient blocks.
constexpr int kLzRec = 520;
// mode 0: S = H from the Gram tiles d_G (also stored into d_rec); mode 1:
// S = B^T from the record d_prev of the previous iteration
void lz_coef(int ng, int bs, int btot, const int* d_mask, const double* d_G, const double* d_prev,
             double* d_rec, double* d_S, int mode, hipStream_t s);
// one CholQR pass on the Gram tiles d_G: S = R^{-1}; pass 0 keeps R in
// d_R1 (ng x 256), pass 1 stores B = R R1 and the success flag in d_rec
void lz_chol(int ng, int bs, int btot, const int* d_mask, const double* d_G, double* d_R1,
             double* d_rec, double* d_S, int pass, hipStream_t s);
void axpby(long long n, double a, const double* x, double b, double* y, hipStream_t s);
void fill(long long n, double v, double* p, hipStream_t s);

}  // namespace sdt
