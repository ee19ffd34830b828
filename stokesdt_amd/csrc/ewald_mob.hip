// Dense Ewald-summed mobility and Cholesky Brownian displacements on the
// GPU: the matrix is built by ewald_block_kernel (one thread per 3x3 block,
// FP64) and applied by rocBLAS GEMMs; the Cholesky factor comes from
// rocSOLVER potrf and y = L z from rocBLAS trmm.
#include <algorithm>
#include <cmath>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>
#include "ewald_gpu.hpp"

namespace sdt {

#define SDT_EW_BLAS(expr)                                                               \
    do {                                                                                \
        rocblas_status st_ = (expr);                                                    \
        if (st_ != rocblas_status_success)                                              \
            throw Error{SDT_EHIP, std::string(#expr) + " failed: " +                   \
                                      rocblas_status_to_string(st_)};                   \
    } while (0)

__global__ void iota_kernel(int n, int* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = i;
}

static size_t span(int rows, int ncols, int ld) { return (size_t)ld * (ncols - 1) + rows; }

// ------------------------------------------------------------ MobEwald -----
MobEwaldGpu::MobEwaldGpu(int npos, const double* rdi, double box, double tol, double xi)
    : npos_(npos), box_(box), tol_(tol), xi_(xi) {
    SDT_REQUIRE(npos > 0 && rdi != nullptr, SDT_EINVAL, "npos must be positive and rdi non-null");
    rdi_.assign(rdi, rdi + npos);
    // the reference's default splitting (src/mob/mob_ewald.cc:51)
    if (!(xi_ > 0.0) && box > 0.0) xi_ = pow(10.0, 1.0 / 6.0) * sqrt(M_PI) / box;
    SDT_HIP_CHECK(stream_create(&stream_));
}

MobEwaldGpu::~MobEwaldGpu() {
    {
        CaptureLock lk(capture_mutex());
        if (blas_) (void)rocblas_destroy_handle(blas_);
    }
    if (stream_) (void)stream_destroy(stream_);
}

void MobEwaldGpu::init() {
    // src/mob/mob_ewald.cc:55-110
    SDT_REQUIRE(box_ > 0.0, SDT_EINVAL, "box size must be positive");
    SDT_REQUIRE(tol_ > 0.0 && tol_ < 1.0, SDT_EINVAL, "the Ewald tolerance must be in (0, 1)");
    SDT_REQUIRE(npos_ <= 30000, SDT_ECAP, "the dense Ewald mobility holds at most 30000 particles");
    SDT_REQUIRE(*std::min_element(rdi_.begin(), rdi_.end()) > 0.0, SDT_EINVAL,
                "particle radii must be positive");
    const size_t n = (size_t)3 * npos_;
    d_mat_.reserve(n * n);
    d_pos_.reserve(n);
    d_rdi_.reserve(npos_);
    d_perm_.reserve(npos_);
    SDT_HIP_CHECK(hipMemcpyAsync(d_rdi_.get(), rdi_.data(), sizeof(double) * npos_,
                                 hipMemcpyHostToDevice, stream_));
    hipLaunchKernelGGL(iota_kernel, dim3(ceil_div(npos_, 256)), dim3(256), 0, stream_, npos_,
                       d_perm_.get());
    SDT_HIP_CHECK(hipGetLastError());
    ew_.setup(box_, xi_, tol_, 0, stream_);
    {
        CaptureLock lk(capture_mutex());
        SDT_EW_BLAS(rocblas_create_handle(&blas_));
    }
    SDT_EW_BLAS(rocblas_set_stream(blas_, stream_));
    inited_ = true;
}

void MobEwaldGpu::update(const double* pos, const double* rdi, bool dev) {
    SDT_REQUIRE(inited_, SDT_ESTATE, "update called before init");
    SDT_REQUIRE(pos != nullptr, SDT_EINVAL, "pos is null");
    hipStream_t s = stream_;
    const size_t n = (size_t)3 * npos_;
    SDT_HIP_CHECK(hipMemcpyAsync(d_pos_.get(), pos, sizeof(double) * n,
                                 dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s));
    if (rdi)
        SDT_HIP_CHECK(hipMemcpyAsync(d_rdi_.get(), rdi, sizeof(double) * npos_,
                                     dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s));
    ew_.build(npos_, d_pos_.get(), d_rdi_.get(), (long long)n, d_mat_.get(), s);
    SDT_HIP_CHECK(hipStreamSynchronize(s));
    updated_ = true;
}

void MobEwaldGpu::apply_packed(int mode, int b, double alpha, const double* X, double beta,
                               double* Y) {
    SDT_REQUIRE(updated_, SDT_ESTATE, "mulvector called before update");
    SDT_REQUIRE(mode == 0, SDT_EINVAL, "the dense Ewald mobility has no real/recip split");
    const int n = 3 * npos_;
    if (fill_y) fill_y(stream());
    // row-major n x b blocks are column-major b x n: Y^T = alpha X^T M + beta Y^T
    SDT_EW_BLAS(rocblas_set_pointer_mode(blas_, rocblas_pointer_mode_host));
    SDT_EW_BLAS(rocblas_dgemm(blas_, rocblas_operation_none, rocblas_operation_none, b, n, n,
                              &alpha, X, b, d_mat_.get(), n, &beta, Y, b));
}

void MobEwaldGpu::mul_vector(int mode, int nrhs, double alpha, int ldf, const double* f,
                             double beta, int ldv, double* v, bool dev) {
    SDT_REQUIRE(updated_, SDT_ESTATE, "mulvector called before update");
    SDT_REQUIRE(mode == 0, SDT_EINVAL, "the dense Ewald mobility has no real/recip split");
    SDT_REQUIRE(nrhs >= 1 && f && v, SDT_EINVAL, "null vector or no right-hand side");
    const int n = 3 * npos_;
    SDT_REQUIRE(ldf >= n && ldv >= n, SDT_EINVAL, "leading dimensions must be at least 3*npos");
    hipStream_t s = stream_;
    const double* df = f;
    double* dv = v;
    if (!dev) {
        d_io_f_.reserve(span(n, nrhs, ldf));
        d_io_v_.reserve(span(n, nrhs, ldv));
        SDT_HIP_CHECK(hipMemcpyAsync(d_io_f_.get(), f, sizeof(double) * span(n, nrhs, ldf),
                                     hipMemcpyHostToDevice, s));
        if (beta != 0.0)
            SDT_HIP_CHECK(hipMemcpyAsync(d_io_v_.get(), v, sizeof(double) * span(n, nrhs, ldv),
                                         hipMemcpyHostToDevice, s));
        df = d_io_f_.get();
        dv = d_io_v_.get();
    }
    // MobMatrix::MulVector: v = alpha M f + beta v (dsymm in the reference)
    SDT_EW_BLAS(rocblas_set_pointer_mode(blas_, rocblas_pointer_mode_host));
    SDT_EW_BLAS(rocblas_dgemm(blas_, rocblas_operation_none, rocblas_operation_none, n, nrhs, n,
                              &alpha, d_mat_.get(), n, df, ldf, &beta, dv, ldv));
    if (!dev)
        SDT_HIP_CHECK(hipMemcpy2DAsync(v, sizeof(double) * ldv, dv, sizeof(double) * ldv,
                                       sizeof(double) * n, nrhs, hipMemcpyDeviceToHost, s));
    SDT_HIP_CHECK(hipStreamSynchronize(s));
}

void MobEwaldGpu::enqueue_mul_vector(int mode, int nrhs, double alpha, int ldf, const double* d_f,
                                     double beta, int ldv, double* d_v) {
    SDT_REQUIRE(updated_, SDT_ESTATE, "mulvector called before update");
    SDT_REQUIRE(mode == 0, SDT_EINVAL, "the dense Ewald mobility has no real/recip split");
    const int n = 3 * npos_;
    SDT_REQUIRE(nrhs >= 1 && ldf >= n && ldv >= n, SDT_EINVAL, "bad shapes");
    SDT_EW_BLAS(rocblas_set_pointer_mode(blas_, rocblas_pointer_mode_host));
    SDT_EW_BLAS(rocblas_dgemm(blas_, rocblas_operation_none, rocblas_operation_none, n, nrhs, n,
                              &alpha, d_mat_.get(), n, d_f, ldf, &beta, d_v, ldv));
}

void MobEwaldGpu::get_matrix(int ld, double* mat, bool dev) {
    SDT_REQUIRE(updated_, SDT_ESTATE, "get_matrix called before update");
    const int n = 3 * npos_;
    SDT_REQUIRE(mat != nullptr && ld >= n, SDT_EINVAL, "null matrix or ld < 3*npos");
    SDT_HIP_CHECK(hipMemcpy2DAsync(mat, sizeof(double) * ld, d_mat_.get(), sizeof(double) * n,
                                   sizeof(double) * n, n,
                                   dev ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, stream_));
    SDT_HIP_CHECK(hipStreamSynchronize(stream_));
}

// ------------------------------------------------------------ BrwnChol -----
BrwnCholGpu::BrwnCholGpu(int dim) : dim_(dim) {}

BrwnCholGpu::~BrwnCholGpu() {
    {
        CaptureLock lk(capture_mutex());
        if (blas_) (void)rocblas_destroy_handle(blas_);
    }
}

void BrwnCholGpu::init() {
    // src/brwn/brwn_chol.cc:29-49
    SDT_REQUIRE(dim_ > 0, SDT_EINVAL, "dimension must be positive");
    SDT_REQUIRE(dim_ <= 90000, SDT_ECAP, "Cholesky displacements: at most 30000 particles");
    d_L_.reserve((size_t)dim_ * dim_);
    d_info_.reserve(1);
    {
        CaptureLock lk(capture_mutex());
        SDT_EW_BLAS(rocblas_create_handle(&blas_));
    }
    inited_ = true;
}

void BrwnCholGpu::compute(MobGpu* mob, int num_rhs, int ldz, const double* z, int ldy, double* y,
                          bool dev) {
    SDT_REQUIRE(inited_, SDT_ESTATE, "compute called before init");
    MobEwaldGpu* me = dynamic_cast<MobEwaldGpu*>(mob);
    SDT_REQUIRE(me != nullptr, SDT_EINVAL,
                "Cholesky displacements need a matrix mobility (the dense Ewald mobility)");
    SDT_REQUIRE(me->updated(), SDT_ESTATE, "mobility not updated");
    SDT_REQUIRE(me->dim() == dim_, SDT_EINVAL, "mobility dimension mismatch");
    SDT_REQUIRE(num_rhs >= 1 && ldz >= dim_ && ldy >= dim_ && z && y, SDT_EINVAL, "bad shapes");
    const int n = dim_;
    hipStream_t s = me->stream();
    SDT_EW_BLAS(rocblas_set_stream(blas_, s));
    SDT_EW_BLAS(rocblas_set_pointer_mode(blas_, rocblas_pointer_mode_host));
    // L L^T = M (dpotrf, lower)
    me->get_matrix(n, d_L_.get(), true);
    SDT_EW_BLAS(rocsolver_dpotrf(blas_, rocblas_fill_lower, n, d_L_.get(), n, d_info_.get()));
    int info = 0;
    SDT_HIP_CHECK(hipMemcpyAsync(&info, d_info_.get(), sizeof(int), hipMemcpyDeviceToHost, s));
    SDT_HIP_CHECK(hipStreamSynchronize(s));
    SDT_REQUIRE(info == 0, SDT_ECONVERGE,
                "Cholesky displacements: the mobility matrix is not positive definite (leading "
                "minor " + std::to_string(info) + ")");
    // y = L z (dtrmm, lower, no transpose)
    const double* dz = z;
    double* dy = y;
    if (!dev) {
        d_io_z_.reserve(span(n, num_rhs, ldz));
        d_io_y_.reserve(span(n, num_rhs, ldy));
        SDT_HIP_CHECK(hipMemcpyAsync(d_io_z_.get(), z, sizeof(double) * span(n, num_rhs, ldz),
                                     hipMemcpyHostToDevice, s));
        dz = d_io_z_.get();
        dy = d_io_y_.get();
    }
    d_B_.reserve((size_t)n * num_rhs);
    SDT_HIP_CHECK(hipMemcpy2DAsync(d_B_.get(), sizeof(double) * n, dz, sizeof(double) * ldz,
                                   sizeof(double) * n, num_rhs, hipMemcpyDeviceToDevice, s));
    const double one = 1.0;
    SDT_EW_BLAS(rocblas_dtrmm(blas_, rocblas_side_left, rocblas_fill_lower, rocblas_operation_none,
                              rocblas_diagonal_non_unit, n, num_rhs, &one, d_L_.get(), n,
                              d_B_.get(), n, dy, ldy));
    if (!dev)
        SDT_HIP_CHECK(hipMemcpy2DAsync(y, sizeof(double) * ldy, dy, sizeof(double) * ldy,
                                       sizeof(double) * n, num_rhs, hipMemcpyDeviceToHost, s));
    SDT_HIP_CHECK(hipStreamSynchronize(s));
}

}  // namespace sdt
