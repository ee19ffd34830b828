// Dense Ewald-summed mobility (MobEwald) and Cholesky Brownian displacements
// (BrwnChol) on the GPU.
#pragma once
#include "mob_gpu.hpp"
#include "real_gpu.hpp"

typedef struct _rocblas_handle* rocblas_handle;

namespace sdt {

// MobEwald (src/mob/mob_ewald.h, src/mob/mob_ewald.cc:19-152): the 3N x 3N
// Ewald-summed RPY matrix stored on the device (EwaldKernel,
// src/mob/ewald.cc:86-330, to tolerance tol), applied by a GEMM. The
// particle order is the caller's (perm() is the identity).
class MobEwaldGpu : public MobGpu {
  public:
    // xi <= 0 selects the reference's 10^(1/6) sqrt(pi) / L (mob_ewald.cc:51)
    MobEwaldGpu(int npos, const double* rdi, double box, double tol, double xi);
    ~MobEwaldGpu() override;
    void init() override;
    void update(const double* pos, const double* rdi, bool dev) override;
    // mode must be 0 (full): the reference's MobEwald has no split
    void mul_vector(int mode, int nrhs, double alpha, int ldf, const double* f, double beta,
                    int ldv, double* v, bool dev) override;
    void enqueue_mul_vector(int mode, int nrhs, double alpha, int ldf, const double* d_f,
                            double beta, int ldv, double* d_v) override;
    void apply_packed(int mode, int b, double alpha, const double* X, double beta,
                      double* Y) override;
    int npos() const override { return npos_; }
    hipStream_t stream() const override { return stream_; }
    const int* perm() const override { return d_perm_.get(); }
    bool updated() const override { return updated_; }
    // MobMatrix::GetMatrix: the matrix (column-major, symmetric) into a host
    // or device buffer with leading dimension ld
    void get_matrix(int ld, double* mat, bool dev);
    const double* d_matrix() const { return d_mat_.get(); }
    double xi() const { return xi_; }

  private:
    int npos_;
    double box_, tol_, xi_;
    std::vector<double> rdi_;
    bool inited_ = false, updated_ = false;
    hipStream_t stream_ = nullptr;
    rocblas_handle blas_ = nullptr;
    EwaldMatrixGpu ew_;
    DevBuf<double> d_pos_, d_rdi_, d_mat_, d_io_f_, d_io_v_;
    DevBuf<int> d_perm_;
};

// BrwnChol (src/brwn/brwn_chol.cc:52-83): y = L z with M = L L^T, the lower
// Cholesky factor of the matrix mobility, recomputed on every call.
class BrwnCholGpu {
  public:
    explicit BrwnCholGpu(int dim);
    ~BrwnCholGpu();
    void init();
    void compute(MobGpu* mob, int num_rhs, int ldz, const double* z, int ldy, double* y, bool dev);

  private:
    int dim_;
    bool inited_ = false;
    rocblas_handle blas_ = nullptr;
    DevBuf<double> d_L_, d_B_, d_io_z_, d_io_y_;
    DevBuf<int> d_info_;
};

}  // namespace sdt
