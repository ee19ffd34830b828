// The plain FP64 GPU form of the reference's BD step (see plain_gpu.hpp):
// stored weight/index tables, per-RHS mesh round trips with global FP64
// atomics, library FFTs, a separate influence kernel, stored 3x3 blocks with
// a block-row SpMV, and BlockLanczos one group at a time with Householder
// QR and the per-iteration Y recomputation. Nothing here is fused or
// batched across right-hand sides: it is the baseline of the optimised path.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>
#include "ewald_scalars.hpp"
#include "plain_gpu.hpp"

namespace sdt {

#define SDT_PL_BLAS(expr)                                                               \
    do {                                                                                \
        rocblas_status st_ = (expr);                                                    \
        if (st_ != rocblas_status_success)                                              \
            throw Error{SDT_EHIP, std::string(#expr) + " failed: " +                   \
                                      rocblas_status_to_string(st_)};                   \
    } while (0)

#define SDT_PL_FFT(expr)                                                                \
    do {                                                                                \
        hipfftResult r_ = (expr);                                                       \
        if (r_ != HIPFFT_SUCCESS)                                                       \
            throw Error{SDT_EHIP, std::string(#expr) + " failed, hipfftResult " +      \
                                      std::to_string((int)r_)};                         \
    } while (0)

// ------------------------------------------------------------- tables ------
// w[x] = M_P(x + 1 - frac), x = 0..P-1 (BSpline(P, floor(c) - c + x),
// src/mob/spme.cc:38-68,469-473), by Cox-de Boor recursion in the order.
template <int P>
__device__ __forceinline__ void plain_bspline(double frac, double* w) {
    const double u = 1.0 - frac;
    w[0] = 1.0;
#pragma unroll
    for (int k = 1; k < P; k++) w[k] = 0.0;
#pragma unroll
    for (int n = 2; n <= P; n++) {
        const double inv = 1.0 / (n - 1);
#pragma unroll
        for (int k = n - 1; k >= 1; k--) w[k] = ((u + k) * w[k] + (n - u - k) * w[k - 1]) * inv;
        w[0] = u * w[0] * inv;
    }
}

// UpdateSpmeEngine_ (src/mob/spme.cc:366-509): per particle the p^3 weights
// P[i][j] = qx qy qz and mesh indices ind[i][j] = (zz K + yy) K + xx, with
// j = (z p + y) p + x (positions wrapped into [0, L) first).
template <int P>
__global__ void plain_tables_kernel(int npos, const double* __restrict__ pos, double box, int K,
                                    int ldp, double* __restrict__ Pt, int* __restrict__ ind) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npos) return;
    double q[3][P];
    int base[3];
    for (int d = 0; d < 3; d++) {
        double x = fmod(pos[3 * i + d], box);
        x = (x >= 0.0 ? x : box + x);
        const double c = x / box * K;
        const double g = floor(c);
        const int b0 = (P % 2 == 0) ? (int)(g - P / 2 + 1) : (int)(g - (P - 1) / 2);
        base[d] = ((b0 % K) + K) % K;
        plain_bspline<P>(c - g, q[d]);
    }
    double* Pi = Pt + (size_t)i * ldp;
    int* Ii = ind + (size_t)i * ldp;
    for (int z = 0; z < P; z++) {
        const int zz = (z + base[2]) % K;
        for (int y = 0; y < P; y++) {
            const int yy = (y + base[1]) % K;
            for (int x = 0; x < P; x++) {
                const int xx = (x + base[0]) % K;
                Pi[(z * P + y) * P + x] = q[0][x] * q[1][y] * q[2][z];
                Ii[(z * P + y) * P + x] = (zz * K + yy) * K + xx;
            }
        }
    }
}

// ComputeSpline (src/mob/spme.cc:71-160): lm2 at every half-spectrum point
// [z][y][kx], 0 at k = 0
__global__ void plain_lm2_kernel(int K, double box, double xi, const double* __restrict__ bsp,
                                 double* __restrict__ lm2) {
    const int Kh = K / 2 + 1;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)K * K * Kh) return;
    const int kx = (int)(t % Kh);
    const int y = (int)((t / Kh) % K);
    const int z = (int)(t / ((long long)Kh * K));
    const int my = (y <= K / 2) ? y : y - K;
    const int mz = (z <= K / 2) ? z : z - K;
    const double w = 2.0 * M_PI / box;
    const double kk = w * w * ((double)kx * kx + (double)my * my + (double)mz * mz);
    double v = 0.0;
    if (kk > 0.0) {
        const double kx2 = kk / (xi * xi);
        v = 6.0 * M_PI * exp(-0.25 * kx2) / kk * (1.0 + kx2 * (0.25 + 0.125 * kx2)) * bsp[kx] *
            bsp[y] * bsp[z];
    }
    lm2[t] = v;
}

// ------------------------------------------------------ mesh round trip ----
// Spread (src/mob/spme.cc:313-363): grid[c][ind] += P f_c (x s_i), one thread
// per (particle, stencil point), FP64 global atomics.
__global__ void plain_spread_kernel(int npos, int p3, int ldp, const double* __restrict__ Pt,
                                    const int* __restrict__ ind, const double* __restrict__ f,
                                    const double* __restrict__ scale, long long K3,
                                    double* __restrict__ grid) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)npos * p3) return;
    const int i = (int)(t / p3), j = (int)(t % p3);
    const double s = scale ? scale[i] : 1.0;
    const double pv = Pt[(size_t)i * ldp + j] * s;
    const int id = ind[(size_t)i * ldp + j];
    atomicAdd(&grid[id], pv * f[3 * i]);
    atomicAdd(&grid[K3 + id], pv * f[3 * i + 1]);
    atomicAdd(&grid[2 * K3 + id], pv * f[3 * i + 2]);
}

// ApplyInfluence (src/mob/spme.cc:201-310): g := lm2 s (I - k k^T / k^2) g
// with s = 1 - k^2/3 (flag 0), 1 (flag 1) or -k^2/6 (flag 2); k = 0 is 0.
__global__ void plain_influence_kernel(int K, double box, const double* __restrict__ lm2,
                                       int flag, long long nspec, double2* __restrict__ g) {
    const int Kh = K / 2 + 1;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nspec) return;
    const int x = (int)(t % Kh);
    const int y = (int)((t / Kh) % K);
    const int z = (int)(t / ((long long)Kh * K));
    const double w = 2.0 * M_PI / box;
    const double kx = w * x, ky = w * ((y <= K / 2) ? y : y - K), kz = w * ((z <= K / 2) ? z : z - K);
    const double kk = kx * kx + ky * ky + kz * kz;
    double2 v0 = g[t], v1 = g[nspec + t], v2 = g[2 * nspec + t];
    if (kk == 0.0) {
        g[t] = g[nspec + t] = g[2 * nspec + t] = double2{0.0, 0.0};
        return;
    }
    double m2 = lm2[t];
    if (flag == 0) m2 *= 1.0 - kk / 3.0;
    else if (flag == 2) m2 *= -kk / 6.0;
    const double ik = 1.0 / kk;
    const double B0 = m2 - m2 * kx * kx * ik, B1 = -m2 * kx * ky * ik, B2 = -m2 * kx * kz * ik;
    const double B3 = m2 - m2 * ky * ky * ik, B4 = -m2 * ky * kz * ik, B5 = m2 - m2 * kz * kz * ik;
    g[t] = double2{B0 * v0.x + B1 * v1.x + B2 * v2.x, B0 * v0.y + B1 * v1.y + B2 * v2.y};
    g[nspec + t] = double2{B1 * v0.x + B3 * v1.x + B4 * v2.x, B1 * v0.y + B3 * v1.y + B4 * v2.y};
    g[2 * nspec + t] = double2{B2 * v0.x + B4 * v1.x + B5 * v2.x, B2 * v0.y + B4 * v1.y + B5 * v2.y};
}

// Interpolate (src/mob/spme.cc:163-198): v_i = beta v_i + a0 s_i sum_j P grid[ind],
// one wave per particle: the lanes read the particle's table rows together
// (contiguous) and the partial sums meet by cross-lane shuffles
__global__ void plain_interp_kernel(int npos, int p3, int ldp, const double* __restrict__ Pt,
                                    const int* __restrict__ ind, const double* __restrict__ grid,
                                    long long K3, const double* __restrict__ scale, double a0,
                                    double beta, double* __restrict__ v) {
    const int i = (int)((blockIdx.x * (size_t)blockDim.x + threadIdx.x) >> 6);
    const int lane = threadIdx.x & 63;
    if (i >= npos) return;
    const double* Pi = Pt + (size_t)i * ldp;
    const int* Ii = ind + (size_t)i * ldp;
    double t0 = 0.0, t1 = 0.0, t2 = 0.0;
    for (int j = lane; j < p3; j += 64) {
        const int id = Ii[j];
        const double pv = Pi[j];
        t0 += grid[id] * pv;
        t1 += grid[K3 + id] * pv;
        t2 += grid[2 * K3 + id] * pv;
    }
    for (int off = 32; off > 0; off >>= 1) {
        t0 += __shfl_xor(t0, off, 64);
        t1 += __shfl_xor(t1, off, 64);
        t2 += __shfl_xor(t2, off, 64);
    }
    if (lane != 0) return;
    const double a = scale ? a0 * scale[i] : a0;
    const double b0 = beta == 0.0 ? 0.0 : beta * v[3 * i];
    const double b1 = beta == 0.0 ? 0.0 : beta * v[3 * i + 1];
    const double b2 = beta == 0.0 ? 0.0 : beta * v[3 * i + 2];
    v[3 * i] = b0 + a * t0;
    v[3 * i + 1] = b1 + a * t1;
    v[3 * i + 2] = b2 + a * t2;
}

// ----------------------------------------------------------- real space ----
// BuildSparseReal (src/mob/mob_spme.cc:135-202): the 3x3 block of every
// listed pair (row-major), self blocks self_a I, overlap correction for
// r < a_i + a_j.
__global__ void plain_blocks_kernel(int npos, const int* __restrict__ rowptr,
                                    const int* __restrict__ colidx, const double* __restrict__ pos,
                                    const double* __restrict__ rdi, double box, double xi,
                                    double* __restrict__ val) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npos) return;
    const double aa = rdi[i];
    const double self_a = 1.0 / aa - xi / sqrt(M_PI) * (6.0 - 40.0 / 3.0 * xi * xi * aa * aa);
    for (int k = rowptr[i]; k < rowptr[i + 1]; k++) {
        const int j = colidx[k];
        double* b = val + (size_t)k * 9;
        if (j == i) {
            for (int e = 0; e < 9; e++) b[e] = (e % 4 == 0) ? self_a : 0.0;
            continue;
        }
        const double dx = remainder(pos[3 * j] - pos[3 * i], box);
        const double dy = remainder(pos[3 * j + 1] - pos[3 * i + 1], box);
        const double dz = remainder(pos[3 * j + 2] - pos[3 * i + 2], box);
        const double ab = rdi[j];
        const double rr = dx * dx + dy * dy + dz * dz;
        const double r = sqrt(rr);
        double xa, ya;
        ewald_scalars<double>(xi, r, aa, ab, &xa, &ya);
        if (pair_rr(dx, dy, dz) < (aa + ab) * (aa + ab)) {
            const double s2 = aa * aa + ab * ab;
            xa += -(1.5 - s2 * 0.5 / rr) / r + 2.0 / (aa + ab) - 3.0 * r / 8.0 / s2;
            ya += -(0.75 + s2 * 0.25 / rr) / r + 2.0 / (aa + ab) - 9.0 * r / 16.0 / s2;
        }
        const double e[3] = {dx / r, dy / r, dz / r};
        for (int a = 0; a < 3; a++)
            for (int c = 0; c < 3; c++) b[3 * a + c] = (xa - ya) * e[a] * e[c] + (a == c ? ya : 0.0);
    }
}

// SpMV3x3 (src/util/sparse.cc:79-153): y_i = beta y_i + alpha sum_k B_k x_j,
// one thread per (block row, right-hand side)
__global__ void plain_spmv_kernel(int npos, int nrhs, const int* __restrict__ rowptr,
                                  const int* __restrict__ colidx, const double* __restrict__ val,
                                  double alpha, int ldx, const double* __restrict__ x, double beta,
                                  int ldy, double* __restrict__ y) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)npos * nrhs) return;
    const int i = (int)(t % npos), r = (int)(t / npos);
    const double* xr = x + (size_t)r * ldx;
    double* yr = y + (size_t)r * ldy;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    for (int k = rowptr[i]; k < rowptr[i + 1]; k++) {
        const int j = colidx[k];
        const double* b = val + (size_t)k * 9;
        const double v0 = xr[3 * j], v1 = xr[3 * j + 1], v2 = xr[3 * j + 2];
        a0 += b[0] * v0 + b[1] * v1 + b[2] * v2;
        a1 += b[3] * v0 + b[4] * v1 + b[5] * v2;
        a2 += b[6] * v0 + b[7] * v1 + b[8] * v2;
    }
    const double c0 = beta == 0.0 ? 0.0 : beta * yr[3 * i];
    const double c1 = beta == 0.0 ? 0.0 : beta * yr[3 * i + 1];
    const double c2 = beta == 0.0 ? 0.0 : beta * yr[3 * i + 2];
    yr[3 * i] = c0 + alpha * a0;
    yr[3 * i + 1] = c1 + alpha * a1;
    yr[3 * i + 2] = c2 + alpha * a2;
}

__global__ void plain_square_kernel(int n, const double* __restrict__ a, double* __restrict__ a2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a2[i] = a[i] * a[i];
}

// ------------------------------------------------------------ PlainMob -----
PlainMobGpu::PlainMobGpu(int npos, const double* rdi, double box, double xi, double rmax, int dim,
                         int porder)
    : npos_(npos), K_(dim), porder_(porder), ldp_(0), box_(box), xi_(xi), rmax_(rmax) {
    SDT_REQUIRE(npos > 0 && rdi != nullptr, SDT_EINVAL, "npos must be positive and rdi non-null");
    rdi_.assign(rdi, rdi + npos);
    SDT_HIP_CHECK(stream_create(&stream_));
}

PlainMobGpu::~PlainMobGpu() {
    if (plans_) {
        {
            CaptureLock lk(capture_mutex());
            (void)hipfftDestroy(fwd_);
        }
        {
            CaptureLock lk(capture_mutex());
            (void)hipfftDestroy(bwd_);
        }
    }
    if (stream_) (void)stream_destroy(stream_);
}

void PlainMobGpu::init() {
    // the checks of src/mob/mob_spme.cc:55-93
    SDT_REQUIRE(box_ > 0.0 && rmax_ > 0.0 && xi_ > 0.0 && K_ > 0, SDT_EINVAL,
                "box, rmax, xi and dim must be positive");
    SDT_REQUIRE(porder_ >= 2 && porder_ <= 12, SDT_EINVAL, "porder must be in [2, 12]");
    SDT_REQUIRE(2.0 * rmax_ < box_, SDT_EINVAL, "rmax must be less than box_size/2");
    SDT_REQUIRE(4 * porder_ <= K_, SDT_EINVAL, "porder must be smaller than dim/4");
    SDT_REQUIRE(K_ % 2 == 0, SDT_EINVAL, "dim must be even");
    const double mx = *std::max_element(rdi_.begin(), rdi_.end());
    const double mn = *std::min_element(rdi_.begin(), rdi_.end());
    SDT_REQUIRE(mn > 0.0, SDT_EINVAL, "particle radii must be positive");
    mono_ = (mx == 1.0 && mn == 1.0);
    const int p3 = porder_ * porder_ * porder_;
    ldp_ = (p3 + 7) / 8 * 8;   // PadLen(p^3, 8), src/mob/spme.cc:625-633
    const size_t K3 = (size_t)K_ * K_ * K_, nspec = (size_t)K_ * K_ * (K_ / 2 + 1);
    d_P_.reserve((size_t)npos_ * ldp_);
    d_ind_.reserve((size_t)npos_ * ldp_);
    d_grid_.reserve(3 * K3);
    d_spec_.reserve(3 * nspec);
    d_lm2_.reserve(nspec);
    d_rdi_.reserve(npos_);
    d_a2_.reserve(npos_);
    SDT_HIP_CHECK(sync_memcpy(d_rdi_.get(), rdi_.data(), sizeof(double) * npos_,
                            hipMemcpyHostToDevice));
    hipLaunchKernelGGL(plain_square_kernel, dim3(ceil_div(npos_, 256)), dim3(256), 0, stream_,
                       npos_, d_rdi_.get(), d_a2_.get());
    std::vector<double> bsp;
    bspline_moduli(K_, porder_, bsp);
    d_bsp_.reserve(K_);
    SDT_HIP_CHECK(sync_memcpy(d_bsp_.get(), bsp.data(), sizeof(double) * K_, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(plain_lm2_kernel, dim3(ceil_div(nspec, 256)), dim3(256), 0, stream_, K_,
                       box_, xi_, d_bsp_.get(), d_lm2_.get());
    SDT_HIP_CHECK(hipGetLastError());
    {
        CaptureLock lk(capture_mutex());
        SDT_PL_FFT(hipfftPlan3d(&fwd_, K_, K_, K_, HIPFFT_D2Z));
    }
    {
        CaptureLock lk(capture_mutex());
        SDT_PL_FFT(hipfftPlan3d(&bwd_, K_, K_, K_, HIPFFT_Z2D));
    }
    plans_ = true;
    SDT_PL_FFT(hipfftSetStream(fwd_, stream_));
    SDT_PL_FFT(hipfftSetStream(bwd_, stream_));
    SDT_REQUIRE(pairs_.init(npos_, rdi_.data(), box_, rmax_, 'R') > 0, SDT_EINVAL,
                "failed to initialise the real-space pair list");
    SDT_HIP_CHECK(hipStreamSynchronize(stream_));
    inited_ = true;
}

void PlainMobGpu::update(const double* d_pos) {
    SDT_REQUIRE(inited_, SDT_ESTATE, "update called before init");
    hipStream_t s = stream_;
    StageScope scope(timer, 8, s);
#define SDT_PLAIN_TABLES(PP)                                                                \
    case PP:                                                                                \
        hipLaunchKernelGGL(plain_tables_kernel<PP>, dim3(ceil_div(npos_, 128)), dim3(128), 0, \
                           s, npos_, d_pos, box_, K_, ldp_, d_P_.get(), d_ind_.get());      \
        break;
    switch (porder_) {
        SDT_PLAIN_TABLES(2) SDT_PLAIN_TABLES(3) SDT_PLAIN_TABLES(4) SDT_PLAIN_TABLES(5)
        SDT_PLAIN_TABLES(6) SDT_PLAIN_TABLES(7) SDT_PLAIN_TABLES(8) SDT_PLAIN_TABLES(9)
        SDT_PLAIN_TABLES(10) SDT_PLAIN_TABLES(11) SDT_PLAIN_TABLES(12)
    }
#undef SDT_PLAIN_TABLES
    SDT_HIP_CHECK(hipGetLastError());
    // pair list in the original order with sorted rows (PairListR), then
    // the stored blocks
    nnz_ = pairs_.build(d_pos, nullptr, nullptr, true, s);
    d_val_.reserve((size_t)nnz_ * 9);
    hipLaunchKernelGGL(plain_blocks_kernel, dim3(ceil_div(npos_, 128)), dim3(128), 0, s, npos_,
                       pairs_.d_rowptr.get(), pairs_.d_colidx.get(), d_pos, d_rdi_.get(), box_,
                       xi_, d_val_.get());
    SDT_HIP_CHECK(hipGetLastError());
    SDT_HIP_CHECK(hipStreamSynchronize(s));
    updated_ = true;
}

// one mesh round trip: zero, spread (scaled by scale_spread), 3 R2C,
// influence(flag), 3 C2R
void PlainMobGpu::round_trip(const double* f, const double* scale_spread, int flag) {
    hipStream_t s = stream_;
    const long long K3 = (long long)K_ * K_ * K_;
    const long long nspec = (long long)K_ * K_ * (K_ / 2 + 1);
    const int p3 = porder_ * porder_ * porder_;
    {
        StageScope sc(timer, 0, s, 1);
        SDT_HIP_CHECK(hipMemsetAsync(d_grid_.get(), 0, sizeof(double) * 3 * K3, s));
        hipLaunchKernelGGL(plain_spread_kernel, dim3(ceil_div((size_t)npos_ * p3, 256)), dim3(256),
                           0, s, npos_, p3, ldp_, d_P_.get(), d_ind_.get(), f, scale_spread, K3,
                           d_grid_.get());
        SDT_HIP_CHECK(hipGetLastError());
    }
    {
        StageScope sc(timer, 1, s, 1);
        for (int c = 0; c < 3; c++)
            SDT_PL_FFT(hipfftExecD2Z(fwd_, d_grid_.get() + c * K3,
                                     reinterpret_cast<hipfftDoubleComplex*>(d_spec_.get() + c * nspec)));
    }
    {
        StageScope sc(timer, 2, s, 1);
        hipLaunchKernelGGL(plain_influence_kernel, dim3(ceil_div(nspec, 256)), dim3(256), 0, s, K_,
                           box_, d_lm2_.get(), flag, nspec, d_spec_.get());
        SDT_HIP_CHECK(hipGetLastError());
    }
    {
        StageScope sc(timer, 3, s, 1);
        for (int c = 0; c < 3; c++)
            SDT_PL_FFT(hipfftExecZ2D(bwd_,
                                     reinterpret_cast<hipfftDoubleComplex*>(d_spec_.get() + c * nspec),
                                     d_grid_.get() + c * K3));
    }
}

// ComputeSpmeRecip_ for one right-hand side (src/mob/spme.cc:743-838)
void PlainMobGpu::recip_one(double alpha, const double* f, double beta, double* v) {
    hipStream_t s = stream_;
    const long long K3 = (long long)K_ * K_ * K_;
    const int p3 = porder_ * porder_ * porder_;
    // alpha K^3 / L^3 with the unnormalised inverse transform: alpha / L^3
    const double a0 = alpha / (box_ * box_ * box_);
    auto interp = [&](const double* scale, double b) {
        StageScope sc(timer, 4, s, 1);
        hipLaunchKernelGGL(plain_interp_kernel, dim3(ceil_div(npos_, 4)), dim3(256), 0, s, npos_,
                           p3, ldp_, d_P_.get(), d_ind_.get(), d_grid_.get(), K3, scale, a0, b, v);
        SDT_HIP_CHECK(hipGetLastError());
    };
    if (mono_) {
        round_trip(f, nullptr, 0);
        interp(nullptr, beta);
    } else {
        round_trip(f, nullptr, 1);
        interp(nullptr, beta);
        round_trip(f, d_a2_.get(), 2);
        interp(nullptr, 1.0);
        round_trip(f, nullptr, 2);
        interp(d_a2_.get(), 1.0);
    }
}

void PlainMobGpu::mul_vector(int mode, int nrhs, double alpha, int ldf, const double* d_f,
                             double beta, int ldv, double* d_v) {
    SDT_REQUIRE(updated_, SDT_ESTATE, "mulvector called before update");
    SDT_REQUIRE(mode >= 0 && mode <= 2 && nrhs >= 1, SDT_EINVAL, "bad mode or num_rhs");
    SDT_REQUIRE(ldf >= 3 * npos_ && ldv >= 3 * npos_, SDT_EINVAL,
                "leading dimensions must be at least 3*npos");
    hipStream_t s = stream_;
    // MobSpme::MulVector (src/mob/mob_spme.cc:219-236): recip with (alpha,
    // beta) one right-hand side after another, then real with (alpha, 1)
    if (mode == 0 || mode == 2)
        for (int r = 0; r < nrhs; r++)
            recip_one(alpha, d_f + (size_t)r * ldf, beta, d_v + (size_t)r * ldv);
    if (mode == 0 || mode == 1) {
        StageScope sc(timer, 5, s, nrhs);
        const double rb = (mode == 0) ? 1.0 : beta;
        hipLaunchKernelGGL(plain_spmv_kernel, dim3(ceil_div((size_t)npos_ * nrhs, 128)), dim3(128),
                           0, s, npos_, nrhs, pairs_.d_rowptr.get(), pairs_.d_colidx.get(),
                           d_val_.get(), alpha, ldf, d_f, rb, ldv, d_v);
        SDT_HIP_CHECK(hipGetLastError());
    }
}

// ------------------------------------------------------- PlainLanczos ------
PlainLanczosGpu::PlainLanczosGpu(int dim, int max_iters, int max_nrhs, double tol)
    : dim_(dim), max_iters_(max_iters), max_nrhs_(max_nrhs), tol_(tol) {}

PlainLanczosGpu::~PlainLanczosGpu() {
    {
        CaptureLock lk(capture_mutex());
        if (blas_) (void)rocblas_destroy_handle(blas_);
    }
}

void PlainLanczosGpu::init() {
    // src/brwn/brwn_lanczos.cc:224-275
    SDT_REQUIRE(dim_ > 0 && max_iters_ > 0 && max_nrhs_ > 0 && tol_ > 0.0, SDT_EINVAL,
                "dimension, max_iters, max_nrhs and tol must be positive");
    if (max_nrhs_ > dim_ / 3) max_nrhs_ = dim_ / 3;
    {
        CaptureLock lk(capture_mutex());
        SDT_PL_BLAS(rocblas_create_handle(&blas_));
    }
    d_V_.reserve((size_t)dim_ * max_nrhs_ * (max_iters_ + 1));
    d_Yold_.reserve((size_t)dim_ * max_nrhs_ * 2);
    d_tau_.reserve(max_nrhs_);
    d_C_.reserve((size_t)(max_iters_ + 1) * max_nrhs_ * max_nrhs_);
    d_H_.reserve((size_t)max_nrhs_ * max_nrhs_);
    inited_ = true;
}

// BlockLanczos (src/brwn/brwn_lanczos.cc:85-201) for one group of b columns;
// returns the iteration of convergence or -1
int PlainLanczosGpu::block(PlainMobGpu* mob, int b, int ldz, const double* z, int ldy, double* y) {
    const int n = dim_;
    hipStream_t s = mob->stream();
    SDT_PL_BLAS(rocblas_set_stream(blas_, s));
    SDT_PL_BLAS(rocblas_set_pointer_mode(blas_, rocblas_pointer_mode_host));
    const rocblas_operation N_ = rocblas_operation_none, T_ = rocblas_operation_transpose;
    const double one = 1.0, zero = 0.0, mone = -1.0;
    double* V = d_V_.get();
    double* Ya = d_Yold_.get();
    double* Yb = Ya + (size_t)n * b;
    const size_t blk = (size_t)n * b;
    auto download = [&](const double* dsrc, int rows, int cols, int ld, std::vector<double>& h) {
        h.assign((size_t)rows * cols, 0.0);
        SDT_HIP_CHECK(hipMemcpy2DAsync(h.data(), sizeof(double) * rows, dsrc, sizeof(double) * ld,
                                       sizeof(double) * rows, cols, hipMemcpyDeviceToHost, s));
        SDT_HIP_CHECK(hipStreamSynchronize(s));
    };
    // [V_0, R] = qr(Z)
    SDT_HIP_CHECK(hipMemcpy2DAsync(V, sizeof(double) * n, z, sizeof(double) * ldz,
                                   sizeof(double) * n, b, hipMemcpyDeviceToDevice, s));
    double normz = 0.0;
    SDT_PL_BLAS(rocblas_dnrm2(blas_, (rocblas_int)blk, V, 1, &normz));
    SDT_PL_BLAS(rocsolver_dgeqrf(blas_, n, b, V, n, d_tau_.get()));
    std::vector<double> R, H, Bq, Bprev((size_t)b * b, 0.0);
    download(V, b, b, n, R);   // column-major, upper triangle is R
    for (int j = 0; j < b; j++)
        for (int k = j + 1; k < b; k++) R[(size_t)j * b + k] = 0.0;
    SDT_PL_BLAS(rocsolver_dorgqr(blas_, n, b, b, V, n, d_tau_.get()));
    const int lt = 2 * b + 1;
    std::vector<double> Tb((size_t)(max_iters_ + 1) * b * lt, 0.0), S, C;
    auto T = [&](int r, int c) -> double& { return Tb[(size_t)r * lt + (c - r + b)]; };
    double* Ycur = Ya;
    double* Yprev = Yb;
    int iter;
    bool conv = false;
    for (iter = 0; iter < max_iters_; iter++) {
        if (cancel && cancel->load()) {
            SDT_HIP_CHECK(hipStreamSynchronize(s));
            throw Error{SDT_ESTATE, "cancelled"};
        }
        double* Vi = V + (size_t)iter * blk;
        double* W = V + (size_t)(iter + 1) * blk;
        double beta = 0.0;
        if (iter > 0) {
            // W = V_{i-1} B_{i-1}^T
            SDT_HIP_CHECK(hipMemcpyAsync(d_H_.get(), Bprev.data(), sizeof(double) * b * b,
                                         hipMemcpyHostToDevice, s));
            StageScope sc(timer, 6, s);
            SDT_PL_BLAS(rocblas_dgemm(blas_, N_, T_, n, b, b, &one, V + (size_t)(iter - 1) * blk, n,
                                      d_H_.get(), b, &zero, W, n));
            beta = -1.0;
        }
        // W = M V_i - W
        mob->mul_vector(0, b, 1.0, n, Vi, beta, n, W);
        {
            StageScope sc(timer, 6, s);
            // H = V_i^T W ; W -= V_i H
            SDT_PL_BLAS(rocblas_dgemm(blas_, T_, N_, b, b, n, &one, Vi, n, W, n, &zero, d_H_.get(),
                                      b));
            SDT_PL_BLAS(rocblas_dgemm(blas_, N_, N_, n, b, b, &mone, Vi, n, d_H_.get(), b, &one, W,
                                      n));
            // [V_{i+1}, B_i] = qr(W)
            SDT_PL_BLAS(rocsolver_dgeqrf(blas_, n, b, W, n, d_tau_.get()));
        }
        download(d_H_.get(), b, b, b, H);
        download(W, b, b, n, Bq);
        // band of T: H_i on the diagonal (upper triangle of H, as the
        // reference reads it), B_i below it for the next iteration
        const int o = iter * b;
        for (int j = 0; j < b; j++)
            for (int k = 0; k <= j; k++) {
                T(o + j, o + k) = H[(size_t)j * b + k];
                T(o + k, o + j) = H[(size_t)j * b + k];
            }
        for (int j = 0; j < b; j++)
            for (int k = 0; k < b; k++) Bprev[(size_t)j * b + k] = (k <= j) ? Bq[(size_t)j * b + k] : 0.0;
        for (int j = 0; j < b; j++)
            for (int k = 0; k <= j; k++) {
                T(o + b + k, o + j) = Bq[(size_t)j * b + k];
                T(o + j, o + b + k) = Bq[(size_t)j * b + k];
            }
        {
            StageScope sc(timer, 6, s);
            SDT_PL_BLAS(rocsolver_dorgqr(blas_, n, b, b, W, n, d_tau_.get()));
        }
        // S = T^{1/2} E_1 (m x b, row-major), C = S R (column-major, ld m)
        const int m = (iter + 1) * b;
        S.resize((size_t)m * b);
        bsq_.compute_band(m, b, Tb.data(), lt, b, S.data());
        C.assign((size_t)m * b, 0.0);
        for (int c = 0; c < b; c++)
            for (int i = 0; i < m; i++) {
                double acc = 0.0;
                for (int k = 0; k <= c; k++) acc += S[(size_t)i * b + k] * R[(size_t)c * b + k];
                C[(size_t)c * m + i] = acc;
            }
        SDT_HIP_CHECK(hipMemcpyAsync(d_C_.get(), C.data(), sizeof(double) * C.size(),
                                     hipMemcpyHostToDevice, s));
        {
            // Y = [V_0 .. V_i] C
            StageScope sc(timer, 6, s);
            SDT_PL_BLAS(rocblas_dgemm(blas_, N_, N_, n, b, m, &one, V, n, d_C_.get(), m, &zero, Ycur,
                                      n));
        }
        if (iter > 0) {
            // ||Y_i - Y_{i-1}||_F / ||Z||_F < tol
            double normy = 0.0;
            SDT_PL_BLAS(rocblas_daxpy(blas_, (rocblas_int)blk, &mone, Ycur, 1, Yprev, 1));
            SDT_PL_BLAS(rocblas_dnrm2(blas_, (rocblas_int)blk, Yprev, 1, &normy));
            if (normy / normz < tol_) {
                conv = true;
                break;
            }
        }
        std::swap(Ycur, Yprev);
    }
    const double* Yout = conv ? Ycur : Yprev;
    SDT_HIP_CHECK(hipMemcpy2DAsync(y, sizeof(double) * ldy, Yout, sizeof(double) * n,
                                   sizeof(double) * n, b, hipMemcpyDeviceToDevice, s));
    return conv ? iter : -1;
}

void PlainLanczosGpu::compute(PlainMobGpu* mob, int num_rhs, int ldz, const double* d_z, int ldy,
                              double* d_y) {
    SDT_REQUIRE(inited_, SDT_ESTATE, "compute called before init");
    SDT_REQUIRE(mob->dim() == dim_, SDT_EINVAL, "mobility dimension mismatch");
    SDT_REQUIRE(num_rhs >= 1 && ldz >= dim_ && ldy >= dim_, SDT_EINVAL, "bad shapes");
    iters_.clear();
    // BrwnLanczos::Compute (src/brwn/brwn_lanczos.cc:279-304): groups of
    // max_nrhs columns, one after another
    for (int c0 = 0; c0 < num_rhs; c0 += max_nrhs_) {
        const int b = std::min(max_nrhs_, num_rhs - c0);
        iters_.push_back(block(mob, b, ldz, d_z + (size_t)c0 * ldz, ldy, d_y + (size_t)c0 * ldy));
    }
    SDT_HIP_CHECK(hipStreamSynchronize(mob->stream()));
}

// -------------------------------------------------------------- PlainBD ----
PlainBDGpu::PlainBDGpu(int npos, const double* pos, const double* rdi, double box, double xi,
                       double rmax, int dim, int porder, int max_iters, int max_nrhs, double tol,
                       double r0, double k0, double dt, int mob_interval, unsigned long long seed)
    : npos_(npos), dt_(dt), mob_interval_(mob_interval), seed_(seed) {
    SDT_REQUIRE(npos > 0 && pos && rdi, SDT_EINVAL, "bad particle arrays");
    SDT_REQUIRE(mob_interval >= 1 && dt > 0.0, SDT_EINVAL, "mob_interval and dt must be positive");
    h_pos0_.assign(pos, pos + 3 * (size_t)npos);
    steric_on_ = r0 > 0.0 && k0 != 0.0;
    mob_.reset(new PlainMobGpu(npos, rdi, box, xi, rmax, dim, porder));
    brwn_.reset(new PlainLanczosGpu(3 * npos, max_iters, max_nrhs, tol));
    steric_.reset(new StericGpu(npos, rdi, box, r0, k0));
}

void PlainBDGpu::init() {
    mob_->init();
    brwn_->init();
    if (steric_on_) steric_->init();
    mob_->timer = &timer_;
    brwn_->timer = &timer_;
    steric_->timer = &timer_;
    ldm_ = 3 * npos_;
    d_pos_.reserve((size_t)ldm_);
    d_Z_.reserve((size_t)ldm_ * mob_interval_);
    d_D_.reserve((size_t)ldm_ * mob_interval_);
    d_f_.reserve((size_t)ldm_);
    SDT_HIP_CHECK(sync_memcpy(d_pos_.get(), h_pos0_.data(), sizeof(double) * ldm_,
                            hipMemcpyHostToDevice));
}

// the step of src/simbox/simbox_bd.cc:296-366 (positions stay on the device)
void PlainBDGpu::advance(int nsteps) {
    hipStream_t s = mob_->stream();
    for (int k = 0; k < nsteps; k++) {
        const int irhs = (int)(istep_ % mob_interval_);
        if (irhs == 0) {
            const long long blk = istep_ / mob_interval_;
            gaussian_fill_device(seed_, (unsigned)blk, (long long)mob_interval_ * ldm_, d_Z_.get(), s);
            mob_->update(d_pos_.get());
            brwn_->compute(mob_.get(), mob_interval_, ldm_, d_Z_.get(), ldm_, d_D_.get());
        }
        double* disp = d_D_.get() + (size_t)irhs * ldm_;
        fill(ldm_, 0.0, d_f_.get(), s);
        if (steric_on_) steric_->accumulate(d_pos_.get(), d_f_.get(), true, s);
        mob_->mul_vector(0, 1, dt_, ldm_, d_f_.get(), sqrt(2.0 * dt_), ldm_, disp);
        axpby(ldm_, 1.0, disp, 1.0, d_pos_.get(), s);
        istep_++;
    }
    SDT_HIP_CHECK(hipStreamSynchronize(s));
    timer_.collect();
}

void PlainBDGpu::get_positions(double* pos) {
    SDT_HIP_CHECK(sync_memcpy(pos, d_pos_.get(), sizeof(double) * ldm_, hipMemcpyDeviceToHost));
}

}  // namespace sdt
