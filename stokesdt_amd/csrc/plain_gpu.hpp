// The plain FP64 GPU form of the reference's BD step: the reference
// algorithm moved to the device as it is structured on the CPU, with no
// fusion. It is the baseline the FP32-mixed path is measured against, and a
// second, independent device implementation of the same operator.
//  - MobSpme recip (src/mob/spme.cc:714-839): a loop over right-hand sides;
//    per RHS the mesh is zeroed, spread with FP64 global atomics from stored
//    N x p^3 weight and index tables (:366-509, :625-633), transformed by
//    three R2C and three C2R FFTs (hipFFT, one per component), multiplied by
//    the influence tensor in its own kernel (:201-310) and interpolated in
//    its own kernel (:163-198);
//  - MobSpme real (src/mob/mob_spme.cc:135-202): stored FP64 3x3 blocks over
//    the pair list, applied by a block-row SpMV (src/util/sparse.cc:79-153);
//  - BlockLanczos (src/brwn/brwn_lanczos.cc:85-201): one group of max_nrhs
//    columns after another, Householder QR (rocSOLVER geqrf / orgqr), the
//    block products on rocBLAS, T^{1/2} E_1 on the host each iteration and
//    Y = V S R recomputed each iteration for the convergence test.
#pragma once
#include <memory>
#include <hipfft/hipfft.h>
#include "mob_gpu.hpp"

typedef struct _rocblas_handle* rocblas_handle;

namespace sdt {

class PlainMobGpu {
  public:
    PlainMobGpu(int npos, const double* rdi, double box, double xi, double rmax, int dim,
                int porder);
    ~PlainMobGpu();
    void init();
    // device positions (AoS, original order)
    void update(const double* d_pos);
    // v := alpha M f + beta v on device vectors in the reference layout;
    // mode 0 full, 1 real, 2 recip
    void mul_vector(int mode, int nrhs, double alpha, int ldf, const double* d_f, double beta,
                    int ldv, double* d_v);
    int dim() const { return 3 * npos_; }
    hipStream_t stream() const { return stream_; }
    long long nnz() const { return nnz_; }
    StageTimer* timer = nullptr;

  private:
    void recip_one(double alpha, const double* f, double beta, double* v);
    void round_trip(const double* f, const double* scale_spread, int flag);
    int npos_, K_, porder_, ldp_;
    double box_, xi_, rmax_;
    std::vector<double> rdi_;
    bool mono_ = true, inited_ = false, updated_ = false;
    hipStream_t stream_ = nullptr;
    hipfftHandle fwd_ = 0, bwd_ = 0;
    bool plans_ = false;
    PairListGpu pairs_;
    DevBuf<double> d_rdi_, d_a2_, d_P_, d_grid_, d_lm2_, d_bsp_, d_val_;
    DevBuf<int> d_ind_;
    DevBuf<double2> d_spec_;
    long long nnz_ = 0;
};

class PlainLanczosGpu {
  public:
    PlainLanczosGpu(int dim, int max_iters, int max_nrhs, double tol);
    ~PlainLanczosGpu();
    void init();
    // device z / y in the reference layout; one BlockLanczos per group
    void compute(PlainMobGpu* mob, int num_rhs, int ldz, const double* d_z, int ldy, double* d_y);
    const std::vector<int>& iters() const { return iters_; }
    StageTimer* timer = nullptr;
    // when set, a solve stops between iterations with SDT_ESTATE "cancelled"
    const std::atomic<bool>* cancel = nullptr;

  private:
    int block(PlainMobGpu* mob, int b, int ldz, const double* z, int ldy, double* y);
    int dim_, max_iters_, max_nrhs_;
    double tol_;
    bool inited_ = false;
    rocblas_handle blas_ = nullptr;
    std::vector<int> iters_;
    DevBuf<double> d_V_, d_tau_, d_C_, d_H_, d_Yold_;
    BandSqrt bsq_;
};

// BDSimBox::Advance (src/simbox/simbox_bd.cc:296-366) on the plain path.
class PlainBDGpu {
  public:
    PlainBDGpu(int npos, const double* pos, const double* rdi, double box, double xi, double rmax,
               int dim, int porder, int max_iters, int max_nrhs, double tol, double r0, double k0,
               double dt, int mob_interval, unsigned long long seed);
    void init();
    void advance(int nsteps);
    void get_positions(double* pos);
    const std::vector<int>& last_iters() const { return brwn_->iters(); }
    StageTimer& timer() { return timer_; }
    long long istep() const { return istep_; }
    void set_cancel(const std::atomic<bool>* c) { brwn_->cancel = c; }

  private:
    int npos_;
    double dt_;
    int mob_interval_;
    unsigned long long seed_;
    std::vector<double> h_pos0_;
    std::unique_ptr<PlainMobGpu> mob_;
    std::unique_ptr<PlainLanczosGpu> brwn_;
    std::unique_ptr<StericGpu> steric_;
    DevBuf<double> d_pos_, d_Z_, d_D_, d_f_;
    long long istep_ = 0;
    bool steric_on_ = true;
    int ldm_ = 0;
    StageTimer timer_;
};

}  // namespace sdt
