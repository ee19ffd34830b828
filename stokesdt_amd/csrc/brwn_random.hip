// Brownian displacements by randomised sampling (BrwnRandom,
// src/brwn/brwn_random.cc:118-263) on the GPU.
//
// Sampling: V = M^3 Omega for num_vecs Gaussian vectors, orthonormalised
// (CholQR2: two Cholesky-QR passes; the reference uses Householder QR, the
// same subspace), B = V^T M V and W = V B^{1/2}. B is symmetric positive
// semi-definite, so the reference's U diag(sqrt(s)) U^T from the SVD of B is
// B^{1/2} from its eigen-decomposition (|eigenvalues|, as singular values).
// Application: first = W V^T z, second = V V^T z - z, and alpha solving
// |first - alpha second|^2 = z^T M z, y = first - alpha second. The GEMMs are
// rocBLAS calls on column-major blocks; M is the SPME mobility.
#include <algorithm>
#include <cmath>
#include <vector>
#include "block_gpu.hpp"
#include "brwn_random.hpp"
#include "host_linalg.hpp"

namespace sdt {

#define SDT_RB_CHECK(expr)                                                              \
    do {                                                                                \
        rocblas_status st_ = (expr);                                                    \
        if (st_ != rocblas_status_success)                                              \
            throw Error{SDT_EHIP, std::string("rocBLAS call failed: ") +                \
                                      rocblas_status_to_string(st_)};                   \
    } while (0)

void sym_eig(int n, const double* A0, double* w, double* Q) {
    std::vector<double> A(A0, A0 + (size_t)n * n);
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) Q[i * n + j] = (i == j) ? 1.0 : 0.0;
    auto at = [&](int i, int j) -> double& { return A[(size_t)i * n + j]; };
    for (int sweep = 0; sweep < 64; sweep++) {
        double off = 0.0, tot = 0.0;
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) {
                tot += at(i, j) * at(i, j);
                if (i != j) off += at(i, j) * at(i, j);
            }
        if (off <= 1e-32 * tot) break;
        for (int p = 0; p < n - 1; p++)
            for (int q = p + 1; q < n; q++) {
                const double apq = at(p, q);
                if (apq == 0.0) continue;
                // rotation in the (p, q) plane that zeroes A(p, q)
                const double th = (at(q, q) - at(p, p)) / (2.0 * apq);
                const double t = (th >= 0.0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < n; k++) {
                    const double kp = at(k, p), kq = at(k, q);
                    at(k, p) = c * kp - s * kq;
                    at(k, q) = s * kp + c * kq;
                }
                for (int k = 0; k < n; k++) {
                    const double pk = at(p, k), qk = at(q, k);
                    at(p, k) = c * pk - s * qk;
                    at(q, k) = s * pk + c * qk;
                }
                for (int k = 0; k < n; k++) {
                    const double kp = Q[k * n + p], kq = Q[k * n + q];
                    Q[k * n + p] = c * kp - s * kq;
                    Q[k * n + q] = s * kp + c * kq;
                }
            }
    }
    for (int i = 0; i < n; i++) w[i] = at(i, i);
}

// y_k = F_k - alpha_k S_k (column-major, y with leading dimension ldy)
__global__ void rnd_combine_kernel(long long dim, int nrhs, const double* __restrict__ F,
                                   const double* __restrict__ S, const double* __restrict__ alpha,
                                   int ldy, double* __restrict__ y) {
    const long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    if (r >= dim || k >= nrhs) return;
    y[(size_t)k * ldy + r] = F[(size_t)k * dim + r] - alpha[k] * S[(size_t)k * dim + r];
}

BrwnRandomGpu::BrwnRandomGpu(int dim, int num_vecs, int max_nrhs, unsigned long long seed)
    : dim_(dim), nv_(num_vecs), max_nrhs_(max_nrhs), seed_(seed) {}

BrwnRandomGpu::~BrwnRandomGpu() {
    {
        CaptureLock lk(capture_mutex());
        if (h_) (void)rocblas_destroy_handle(h_);
    }
}

void BrwnRandomGpu::init() {
    // BrwnRandom::Init (src/brwn/brwn_random.cc:42-53)
    SDT_REQUIRE(dim_ > 0, SDT_EINVAL, "the dimension must be positive");
    SDT_REQUIRE(nv_ > 0 && nv_ <= dim_, SDT_EINVAL,
                "the number of random vectors must be in [1, dim]");
    SDT_REQUIRE(nv_ <= kMaxVecs, SDT_ECAP, "at most 128 random vectors");
    SDT_REQUIRE(max_nrhs_ > 0, SDT_EINVAL, "max_nrhs must be positive");
    if (!h_) {
        CaptureLock lk(capture_mutex());
        SDT_RB_CHECK(rocblas_create_handle(&h_));
    }
    const size_t blk = (size_t)dim_ * nv_;
    d_V_.reserve(blk);
    d_W_.reserve(blk);
    d_X_.reserve(blk);
    d_G_.reserve((size_t)nv_ * nv_);
    inited_ = true;
}

void BrwnRandomGpu::cholqr(hipStream_t s) {
    const double one = 1.0, zero = 0.0;
    std::vector<double> G((size_t)nv_ * nv_), R((size_t)nv_ * nv_), Ri((size_t)nv_ * nv_);
    for (int pass = 0; pass < 2; pass++) {
        // G = V^T V
        SDT_RB_CHECK(rocblas_dgemm(h_, rocblas_operation_transpose, rocblas_operation_none, nv_,
                                   nv_, dim_, &one, d_V_.get(), dim_, d_V_.get(), dim_, &zero,
                                   d_G_.get(), nv_));
        SDT_HIP_CHECK(hipMemcpyAsync(G.data(), d_G_.get(), sizeof(double) * G.size(),
                                     hipMemcpyDeviceToHost, s));
        SDT_HIP_CHECK(hipStreamSynchronize(s));
        for (int i = 0; i < nv_; i++)
            for (int j = 0; j < i; j++) G[i * nv_ + j] = G[j * nv_ + i] =
                0.5 * (G[i * nv_ + j] + G[j * nv_ + i]);
        SDT_REQUIRE(chol_upper(nv_, G.data(), R.data()), SDT_ECONVERGE,
                    "random sampling: the sampled vectors are rank deficient");
        inv_upper(nv_, R.data(), Ri.data());
        SDT_HIP_CHECK(hipMemcpyAsync(d_G_.get(), Ri.data(), sizeof(double) * Ri.size(),
                                     hipMemcpyHostToDevice, s));
        // V = V R^{-1}: the row-major R^{-1} read column-major is its
        // transpose, hence op = transpose
        SDT_RB_CHECK(rocblas_dgemm(h_, rocblas_operation_none, rocblas_operation_transpose, dim_,
                                   nv_, nv_, &one, d_V_.get(), dim_, d_G_.get(), nv_, &zero,
                                   d_X_.get(), dim_));
        std::swap(d_V_.ptr, d_X_.ptr);
        std::swap(d_V_.n, d_X_.n);
    }
}

void BrwnRandomGpu::sample(MobGpu* mob) {
    hipStream_t s = mob->stream();
    SDT_RB_CHECK(rocblas_set_stream(h_, s));
    const double one = 1.0, zero = 0.0;
    // omega, then V = M^3 omega (src/brwn/brwn_random.cc:123-128)
    gaussian_fill_device(seed_, draws_++, (long long)dim_ * nv_, d_X_.get(), s);
    mob->mul_vector(0, nv_, 1.0, dim_, d_X_.get(), 0.0, dim_, d_V_.get(), true);
    mob->mul_vector(0, nv_, 1.0, dim_, d_V_.get(), 0.0, dim_, d_X_.get(), true);
    mob->mul_vector(0, nv_, 1.0, dim_, d_X_.get(), 0.0, dim_, d_V_.get(), true);
    cholqr(s);
    // B = V^T M V (:134-137)
    mob->mul_vector(0, nv_, 1.0, dim_, d_V_.get(), 0.0, dim_, d_X_.get(), true);
    SDT_RB_CHECK(rocblas_dgemm(h_, rocblas_operation_transpose, rocblas_operation_none, nv_, nv_,
                               dim_, &one, d_V_.get(), dim_, d_X_.get(), dim_, &zero, d_G_.get(),
                               nv_));
    std::vector<double> B((size_t)nv_ * nv_), Q((size_t)nv_ * nv_), w(nv_), H((size_t)nv_ * nv_);
    SDT_HIP_CHECK(hipMemcpyAsync(B.data(), d_G_.get(), sizeof(double) * B.size(),
                                 hipMemcpyDeviceToHost, s));
    SDT_HIP_CHECK(hipStreamSynchronize(s));
    for (int i = 0; i < nv_; i++)
        for (int j = 0; j < i; j++) B[i * nv_ + j] = B[j * nv_ + i] =
            0.5 * (B[i * nv_ + j] + B[j * nv_ + i]);
    // B^{1/2} = Q diag(sqrt|w|) Q^T (:139-146)
    sym_eig(nv_, B.data(), w.data(), Q.data());
    for (int i = 0; i < nv_; i++)
        for (int j = 0; j < nv_; j++) {
            double acc = 0.0;
            for (int k = 0; k < nv_; k++)
                acc += Q[i * nv_ + k] * std::sqrt(std::fabs(w[k])) * Q[j * nv_ + k];
            H[i * nv_ + j] = acc;
        }
    SDT_HIP_CHECK(hipMemcpyAsync(d_G_.get(), H.data(), sizeof(double) * H.size(),
                                 hipMemcpyHostToDevice, s));
    // W = V B^{1/2} (B^{1/2} symmetric: either storage order)
    SDT_RB_CHECK(rocblas_dgemm(h_, rocblas_operation_none, rocblas_operation_none, dim_, nv_, nv_,
                               &one, d_V_.get(), dim_, d_G_.get(), nv_, &zero, d_W_.get(), dim_));
    SDT_HIP_CHECK(hipStreamSynchronize(s));
    mob_ = mob;
}

void BrwnRandomGpu::compute(MobGpu* mob, int num_rhs, int ldz, const double* z, int ldy,
                            double* y, bool dev) {
    SDT_REQUIRE(inited_, SDT_ESTATE, "compute called before init");
    SDT_REQUIRE(num_rhs >= 1, SDT_EINVAL, "num_rhs must be positive");
    SDT_REQUIRE(ldz >= dim_ && ldy >= dim_, SDT_EINVAL, "leading dimensions must be >= dim");
    SDT_REQUIRE(z && y, SDT_EINVAL, "null vector");
    if (mob) {
        SDT_REQUIRE(mob->dim() == dim_, SDT_EINVAL, "mobility dimension differs from dim");
        sample(mob);
    }
    SDT_REQUIRE(mob_ != nullptr, SDT_ESTATE, "no sampled subspace: call with a mobility first");
    MobGpu* m = mob_;
    hipStream_t s = m->stream();
    SDT_RB_CHECK(rocblas_set_stream(h_, s));
    const double one = 1.0, zero = 0.0, mone = -1.0;
    const size_t span_z = (size_t)ldz * (num_rhs - 1) + dim_;
    const size_t span_y = (size_t)ldy * (num_rhs - 1) + dim_;
    const double* dz = z;
    double* dy = y;
    if (!dev) {
        d_io_z_.reserve(span_z);
        d_io_y_.reserve(span_y);
        SDT_HIP_CHECK(hipMemcpyAsync(d_io_z_.get(), z, sizeof(double) * span_z,
                                     hipMemcpyHostToDevice, s));
        dz = d_io_z_.get();
        dy = d_io_y_.get();
    }
    const size_t cols = (size_t)dim_ * num_rhs;
    d_T_.reserve((size_t)nv_ * num_rhs);
    d_F_.reserve(cols);
    d_S_.reserve(cols);
    d_alpha_.reserve(num_rhs);
    // T = V^T z; first = W T; second = V T - z (:188-213)
    SDT_RB_CHECK(rocblas_dgemm(h_, rocblas_operation_transpose, rocblas_operation_none, nv_,
                               num_rhs, dim_, &one, d_V_.get(), dim_, dz, ldz, &zero, d_T_.get(),
                               nv_));
    SDT_RB_CHECK(rocblas_dgemm(h_, rocblas_operation_none, rocblas_operation_none, dim_, num_rhs,
                               nv_, &one, d_W_.get(), dim_, d_T_.get(), nv_, &zero, d_F_.get(),
                               dim_));
    SDT_HIP_CHECK(hipMemcpy2DAsync(d_S_.get(), sizeof(double) * dim_, dz, sizeof(double) * ldz,
                                   sizeof(double) * dim_, num_rhs, hipMemcpyDeviceToDevice, s));
    SDT_RB_CHECK(rocblas_dgemm(h_, rocblas_operation_none, rocblas_operation_none, dim_, num_rhs,
                               nv_, &one, d_V_.get(), dim_, d_T_.get(), nv_, &mone, d_S_.get(),
                               dim_));
    // M z into y (:214-215), then the per-column quadratic for alpha
    m->mul_vector(0, num_rhs, 1.0, ldz, dz, 0.0, ldy, dy, true);
    std::vector<double> zy(num_rhs), ss(num_rhs), fs(num_rhs), ff(num_rhs), al(num_rhs);
    SDT_RB_CHECK(rocblas_set_pointer_mode(h_, rocblas_pointer_mode_host));
    SDT_RB_CHECK(rocblas_ddot_strided_batched(h_, dim_, dz, 1, ldz, dy, 1, ldy, num_rhs, zy.data()));
    SDT_RB_CHECK(rocblas_ddot_strided_batched(h_, dim_, d_S_.get(), 1, dim_, d_S_.get(), 1, dim_,
                                              num_rhs, ss.data()));
    SDT_RB_CHECK(rocblas_ddot_strided_batched(h_, dim_, d_F_.get(), 1, dim_, d_S_.get(), 1, dim_,
                                              num_rhs, fs.data()));
    SDT_RB_CHECK(rocblas_ddot_strided_batched(h_, dim_, d_F_.get(), 1, dim_, d_F_.get(), 1, dim_,
                                              num_rhs, ff.data()));
    for (int k = 0; k < num_rhs; k++) {
        // a alpha^2 + b alpha + c = 0 (:216-237)
        const double a = ss[k], b = -2.0 * fs[k], c = ff[k] - zy[k];
        al[k] = (-b + std::sqrt(b * b - 4.0 * a * c)) / (2.0 * a);
    }
    SDT_HIP_CHECK(hipMemcpyAsync(d_alpha_.get(), al.data(), sizeof(double) * num_rhs,
                                 hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(rnd_combine_kernel, dim3((unsigned)ceil_div((long long)dim_, 256), num_rhs),
                       dim3(256), 0, s, (long long)dim_, num_rhs, d_F_.get(), d_S_.get(),
                       d_alpha_.get(), ldy, dy);
    SDT_HIP_CHECK(hipGetLastError());
    if (!dev)
        SDT_HIP_CHECK(hipMemcpy2DAsync(y, sizeof(double) * ldy, dy, sizeof(double) * ldy,
                                       sizeof(double) * dim_, num_rhs, hipMemcpyDeviceToHost, s));
    SDT_HIP_CHECK(hipStreamSynchronize(s));
}

}  // namespace sdt
