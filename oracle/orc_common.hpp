// Oracle for the StokesDT Brownian-dynamics step: shared declarations.
//
// TEST INFRASTRUCTURE ONLY. This directory is a CPU FP64 (C++/OpenMP)
// restatement of the reference algorithm. It is used by tests/ as the
// checker, by __graft_entry__.smoke() as the checker, and by bench.py as the
// timed CPU baseline. The product library (stokesdt_amd/) never links,
// loads or calls it.
//
// Each function cites the reference file:line it restates (paths relative
// to the xing-liu/stokesdt source tree).
#pragma once

#include <complex>
#include <cstdint>
#include <cstddef>
#include <vector>

namespace orc {

typedef std::complex<double> cplx;

// ---------------------------------------------------------------- FFT ------
// Mixed-radix (any factorisation; radices 2,3,4,5 have dedicated
// butterflies) self-sorting Stockham transform of one complex sequence.
struct FFTPlan1D {
    int n = 0;
    std::vector<int> radix;          // stage radices, product == n
    std::vector<cplx> tw;            // per-stage twiddles, concatenated
    std::vector<int> tw_off;         // offset of each stage in tw
    void init(int n_);
    // sign = -1 forward (e^{-i...}), +1 backward (unnormalised)
    void exec(cplx* x, cplx* work, int sign) const;
};

// 3D real transforms on a compact in-place mesh of K^3 reals stored as
// [z][y][x] with x padded to ldx = 2*(K/2+1) doubles (in-place R2C layout,
// as the reference's DFTI descriptors at src/mob/spme.cc:539-618 use, with
// FFTPadLen padding dropped: padding does not change any value).
struct FFT3D {
    int K = 0;
    int ldx = 0;           // doubles per x-row (2*(K/2+1))
    FFTPlan1D plan;
    void init(int K_);
    void forward(double* mesh) const;    // R2C in place, unscaled
    void backward(double* mesh) const;   // C2R in place, scaled by 1/K^3
};

// ------------------------------------------------------------- linalg ------
// Symmetric eigendecomposition A = V diag(w) V^T (row-major n x n A, V
// columns are eigenvectors, ascending w). Householder tridiagonalisation
// followed by implicit QL with Wilkinson-type shifts.
void sym_eig(int n, const double* A, int lda, double* w, double* V, int ldv);
// Tridiagonal eigendecomposition (d: n diagonal, e: n-1 off-diagonal).
void tridiag_eig(int n, const double* d, const double* e, double* w,
                 double* V, int ldv);
// Householder QR of a column-major m x n matrix (m >= n): on exit A holds
// the explicit Q (m x n) and R (n x n, column-major, upper) is returned.
void householder_qr(int m, int n, double* A, int lda, double* R, int ldr);
// Cholesky A = L L^T (row-major, lower). Returns 0 on success, k>0 if the
// leading minor k is not positive definite.
int cholesky_lower(int n, double* A, int lda);

// ------------------------------------------------------------ helpers ------
inline size_t pad_len(size_t len, size_t size) {
    // PadLen from src/util/common.h:61-65 (64-byte alignment)
    size_t len0 = 64 / size;
    return (len + len0 - 1) / len0 * len0;
}
double frob_rel_err(int nrows, int ncols, int ld1, const double* a,
                    int ld2, const double* b);

// --------------------------------------------------------- mobility -------
class Mob {
  public:
    virtual ~Mob() {}
    virtual int dim() const = 0;
    // v := alpha * M * f + beta * v, nrhs column-major columns
    virtual void mul_vector(int nrhs, double alpha, int ldf, const double* f,
                            double beta, int ldv, double* v) = 0;
};

// Philox-4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) and Box-Muller.
void philox4x32_10(const uint32_t ctr[4], const uint32_t key[2],
                   uint32_t out[4]);
void gaussian_fill(uint64_t seed, uint32_t stream, int64_t n, double* out);

}  // namespace orc
