// Small dense FP64 host linear algebra for the Krylov solver.
#pragma once
#include <vector>

namespace sdt {

// G = R^T R for symmetric positive definite b x b G (row-major); R upper
// (row-major). Returns false if G is not numerically positive definite.
bool chol_upper(int b, const double* G, double* R);
// inverse of an upper-triangular b x b matrix (row-major)
void inv_upper(int b, const double* R, double* Rinv);
// C = A * B for b x b row-major upper-triangular matrices
void mul_upper(int b, const double* A, const double* B, double* C);

// S = T^{1/2} E_1 for a symmetric band matrix T (m x m, row-major, lower
// bandwidth kd) with E_1 the first kb columns of the identity. Band to
// tridiagonal reduction by Givens bulge chasing, implicit QL on the
// tridiagonal, with every rotation applied to the thin m x kb matrix only,
// so the cost is O(m^2 (kd + kb)) instead of O(m^3). Eigenvalues below zero
// (rounding) are clamped to zero.
// The lower triangle of T's band is read (T is taken to be symmetric).
class BandSqrt {
  public:
    void compute(int m, int kd, const double* T, int kb, double* S);
    // the same with T in band storage: T(i, j), |i - j| <= kd, at
    // Tb[i * ldt + j - i + kd] (rows 0 .. m-1; entries outside T unused)
    void compute_band(int m, int kd, const double* Tb, int ldt, int kb, double* S);

    // a Givens rotation on rows (p, p+1)
    struct Rot { int p; double c, s; };

  private:
    std::vector<double> A_;
    std::vector<Rot> band_rot_, ql_rot_;
    std::vector<double> d_, e_, X_;
};

}  // namespace sdt
