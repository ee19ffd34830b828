// Brownian displacements by randomised sampling (BrwnRandom,
// src/brwn/brwn_random.h:16-86) on the GPU.
#pragma once
#include <rocblas/rocblas.h>
#include "mob_gpu.hpp"

namespace sdt {

class BrwnRandomGpu {
  public:
    // num_vecs random vectors (1 .. min(dim, 128)); max_nrhs as the
    // reference's; the sample vectors come from the Gaussian stream of seed
    // (stream index = number of samplings so far)
    BrwnRandomGpu(int dim, int num_vecs, int max_nrhs, unsigned long long seed);
    ~BrwnRandomGpu();
    void init();
    // BrwnRandom::Compute (src/brwn/brwn_random.cc:118-263): a non-null mob
    // draws a new sample and rebuilds the subspace; y = first + alpha
    // (I - V V^T) z per column with |y|^2 = z^T M z. z, y column-major, host
    // or device (dev).
    void compute(MobGpu* mob, int num_rhs, int ldz, const double* z, int ldy, double* y,
                 bool dev);
    static constexpr int kMaxVecs = 128;

  private:
    void sample(MobGpu* mob);
    void cholqr(hipStream_t s);
    int dim_, nv_, max_nrhs_;
    unsigned long long seed_;
    unsigned draws_ = 0;
    bool inited_ = false;
    rocblas_handle h_ = nullptr;
    MobGpu* mob_ = nullptr;
    DevBuf<double> d_V_, d_W_, d_X_, d_G_, d_T_, d_F_, d_S_, d_io_z_, d_io_y_, d_alpha_;
};

// cyclic Jacobi eigen-decomposition of a symmetric n x n matrix (row-major):
// w = eigenvalues, Q = eigenvectors as columns (row-major)
void sym_eig(int n, const double* A, double* w, double* Q);

}  // namespace sdt
