// Oracle Krylov Brownian displacements (test infrastructure only; see
// orc_common.hpp).
#pragma once
#include <vector>
#include "orc_common.hpp"

namespace orc {

class BrwnLanczos {
  public:
    BrwnLanczos(int dim, int max_iters, int max_nrhs, double tol);
    bool init();
    void compute(Mob* mob, int num_rhs, int ldz, const double* z, int ldy,
                 double* y);
    const std::vector<int>& iters() const { return iters_; }
    int max_nrhs() const { return max_nrhs_; }

  private:
    int lanczos(Mob* mob, const double* z, double* y);
    int block_lanczos(Mob* mob, int b, int ldz, const double* z, int ldy,
                      double* y);
    int dim_, max_iters_, max_nrhs_;
    double tol_;
    int ldv_ = 0;
    int last_iters_ = 0;
    std::vector<double> v_, y_old_;
    std::vector<int> iters_;
};

int chol_displacements(int n, const double* mat, int ldm, int nrhs, int ldz,
                       const double* z, int ldy, double* y);

// dense operator wrapper (row-major symmetric matrix)
class DenseMob : public Mob {
  public:
    DenseMob(int n, const double* mat, int ldm) : n_(n), m_(mat), ld_(ldm) {}
    int dim() const override { return n_; }
    void mul_vector(int nrhs, double alpha, int ldf, const double* f,
                    double beta, int ldv, double* v) override {
        #pragma omp parallel for schedule(static)
        for (int i = 0; i < n_; i++) {
            for (int r = 0; r < nrhs; r++) {
                double acc = 0.0;
                const double* fr = f + (size_t)r * ldf;
                for (int j = 0; j < n_; j++) acc += m_[(size_t)i * ld_ + j] * fr[j];
                double* vr = v + (size_t)r * ldv;
                vr[i] = (beta == 0.0 ? 0.0 : beta * vr[i]) + alpha * acc;
            }
        }
    }

  private:
    int n_;
    const double* m_;
    int ld_;
};

}  // namespace orc
