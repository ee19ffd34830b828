// Matrix-free real-space RPY-Ewald matvec for gfx950 (FP32 and FP64).
//
// Reference: the real-space part of MobSpme is a stored 3x3-block CSR built
// once per update (MobSpme::BuildSparseReal, src/mob/mob_spme.cc:135-202,
// with scalars_ewald_F from src/mob/ewald.h:55-85) and applied by SpMV3x3
// (src/util/sparse.cc:79-153). Here the blocks are recomputed inside the
// matvec from FP64 position differences, for a chunk of right-hand sides per
// thread, so a matvec streams only the column indices and the input vector.
#include <algorithm>
#include <cmath>
#include "real_gpu.hpp"
#include "sort_gpu.hpp"

namespace sdt {

// squared min-image distance with explicit fused operations, so that every
// kernel that classifies a pair as overlapping rounds identically
__device__ __forceinline__ double pair_rr(double dx, double dy, double dz) {
    return fma(dx, dx, fma(dy, dy, dz * dz));
}

template <typename T>
__device__ __forceinline__ void ewald_scalars(T xi, T r, T aa, T ab, T* xa, T* ya) {
    const T xir = xi * r;
    const T xir2 = xir * xir;
    const T r2 = r * r;
    const T r3 = r2 * r;
    const T aa2 = aa * aa, ab2 = ab * ab;
    const T erfcxir = erfc(xir);
    const T expxir2 = xi * T(0.56418958354775628694807945156077) * exp(-xir2);  // xi/sqrt(pi)
    const T a0 = T(2) * (T(-3) + xir2 * T(2)) * expxir2 + erfcxir / r;
    const T b0 = T(2) * (T(1) + xir2 * T(-2)) * expxir2 + erfcxir / r;
    const T a2 = T(4) / r2 * (T(1) + xir2 * (T(14) + xir2 * (T(-20) + xir2 * T(4)))) * expxir2 +
                 T(2) / r3 * erfcxir;
    const T b2 = T(-4) / r2 * (T(3) + xir2 * (T(2) + xir2 * (T(-16) + xir2 * T(4)))) * expxir2 -
                 T(6) / r3 * erfcxir;
    const T s6 = (aa2 + ab2) / T(6);
    *ya = T(0.75) * (a0 + a2 * s6);
    *xa = T(0.75) * (b0 + b2 * s6) + *ya;
}

// Y[(3 i + c) b + r] = beta Y + alpha sum_j B_ij X_j for r in [r0, r0 + RC),
// over the self block and the non-overlapping pairs (r >= a_i + a_j);
// overlapping pairs are added by real_overlap_kernel. LPR consecutive lanes
// share a row, each taking every LPR-th pair, so a row's gathers are in
// flight together; partial sums are combined with cross-lane shuffles.
// BLK: X is the column-blocked copy Xb[cb][j][c][r] (RC columns per block,
// zero-padded), so the 3 RC values a pair gathers are one contiguous run of
// 12 RC or 24 RC bytes instead of three pieces of rows b values apart.
template <typename T, int RC, int LPR, typename XT, bool BLK>
__global__ void __launch_bounds__(256)
real_matvec_kernel(int npos, const int* __restrict__ rowptr, const int* __restrict__ colidx,
                   const double* __restrict__ pos, const double* __restrict__ rdi, double box,
                   double xi, int b, double alpha, const XT* __restrict__ X, double beta,
                   double* __restrict__ Y) {
    // rows are in spatially sorted order: XCD-contiguous block ranges keep
    // the gathered neighbour rows of X within one L2
    const unsigned lid = xcd_logical_id(blockIdx.x + blockIdx.y * gridDim.x, gridDim.x * gridDim.y);
    const int bxl = (int)(lid % gridDim.x), byl = (int)(lid / gridDim.x);
    const int tid = bxl * blockDim.x + threadIdx.x;
    const int i = tid / LPR, sub = tid % LPR;
    const bool valid = i < npos;
    const int r0 = byl * RC;
    const int rc = min(RC, b - r0);
    T acc[3][RC];
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
        for (int r = 0; r < RC; r++) acc[c][r] = T(0);
    if (valid) {
        const double x1 = pos[3 * i], y1 = pos[3 * i + 1], z1 = pos[3 * i + 2];
        const double aa_d = rdi[i];
        const T aa = (T)aa_d;
        const T txi = (T)xi;
        // self block: (1/a - xi/sqrt(pi) (6 - 40/3 xi^2 a^2)) I
        const T self_a =
            (T)(1.0 / aa_d - xi / sqrt(M_PI) * (6.0 - 40.0 / 3.0 * xi * xi * aa_d * aa_d));
        const int kend = rowptr[i + 1];
        for (int k = rowptr[i] + sub; k < kend; k += LPR) {
            const int j = colidx[k];
            XT xv[3][RC];
            if constexpr (BLK) {
                const XT* xj = X + ((size_t)byl * npos + j) * (3 * RC);
#pragma unroll
                for (int r = 0; r < RC; r++) {
                    xv[0][r] = xj[r];
                    xv[1][r] = xj[RC + r];
                    xv[2][r] = xj[2 * RC + r];
                }
            } else {
                const XT* xj = X + (size_t)(3 * j) * b + r0;
#pragma unroll
                for (int r = 0; r < RC; r++) {
                    xv[0][r] = xv[1][r] = xv[2][r] = XT(0);
                    if (r < rc) {
                        xv[0][r] = xj[r];
                        xv[1][r] = xj[b + r];
                        xv[2][r] = xj[2 * b + r];
                    }
                }
            }
            T bxx, bxy, bxz, byy, byz, bzz;
            if (j == i) {
                bxx = byy = bzz = self_a;
                bxy = bxz = byz = T(0);
            } else {
                const double dx = remainder(pos[3 * j] - x1, box);
                const double dy = remainder(pos[3 * j + 1] - y1, box);
                const double dz = remainder(pos[3 * j + 2] - z1, box);
                const double ab_d = rdi[j];
                const double sab = aa_d + ab_d;
                if (pair_rr(dx, dy, dz) < sab * sab) continue;   // overlap pass
                const T rx = (T)dx, ry = (T)dy, rz = (T)dz;
                const T rr = rx * rx + ry * ry + rz * rz;
                const T r = sqrt(rr);
                const T ab = (T)ab_d;
                T xa, ya;
                ewald_scalars<T>(txi, r, aa, ab, &xa, &ya);
                const T ir = T(1) / r;
                const T ex = rx * ir, ey = ry * ir, ez = rz * ir;
                const T d = xa - ya;
                bxx = d * ex * ex + ya; bxy = d * ex * ey; bxz = d * ex * ez;
                byy = d * ey * ey + ya; byz = d * ey * ez; bzz = d * ez * ez + ya;
            }
#pragma unroll
            for (int r = 0; r < RC; r++) {
                const T v0 = (T)xv[0][r], v1 = (T)xv[1][r], v2 = (T)xv[2][r];
                acc[0][r] += bxx * v0 + bxy * v1 + bxz * v2;
                acc[1][r] += bxy * v0 + byy * v1 + byz * v2;
                acc[2][r] += bxz * v0 + byz * v1 + bzz * v2;
            }
        }
    }
#pragma unroll
    for (int off = LPR / 2; off > 0; off >>= 1)
#pragma unroll
        for (int c = 0; c < 3; c++)
#pragma unroll
            for (int r = 0; r < RC; r++) acc[c][r] += __shfl_xor(acc[c][r], off, 64);
    if (valid && sub == 0) {
#pragma unroll
        for (int c = 0; c < 3; c++) {
            double* yi = Y + (size_t)(3 * i + c) * b + r0;
#pragma unroll
            for (int r = 0; r < RC; r++) {
                if (r < rc) yi[r] = (beta == 0.0 ? 0.0 : beta * yi[r]) + alpha * (double)acc[c][r];
            }
        }
    }
}

// Xb[cb][j][c][r] = X[(3 j + c) b + cb RC + r], 0 beyond column b: the
// column-blocked copy the gather kernel reads (OT = float: FP32 arithmetic,
// half the bytes per gathered neighbour)
template <typename OT, int RC>
__global__ void block_copy_kernel(int npos, int b, int nblk, const double* __restrict__ x,
                                  OT* __restrict__ y) {
    const long long n = (long long)nblk * npos * 3 * RC;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += stride) {
        const int r = (int)(e % RC);
        const long long q = e / RC;            // (cb npos + j) 3 + c
        const long long row = q % (3LL * npos);
        const int cb = (int)(q / (3LL * npos));
        const int col = cb * RC + r;
        y[e] = col < b ? (OT)x[row * b + col] : OT(0);
    }
}

template <typename T, int RC, typename XT, bool BLK>
static void launch_real(int npos, const int* d_rowptr, const int* d_colidx, const double* d_pos,
                        const double* d_rdi, double box, double xi, int b, double alpha,
                        const XT* X, double beta, double* Y, hipStream_t s) {
    // lanes per row: 4 for the blocked gathers (measured at b = 100), 8 for
    // the in-place ones (b < 8)
    constexpr int LPR = BLK ? 4 : 8;
    dim3 grid(ceil_div((size_t)npos * LPR, 256), ceil_div(b, RC));
    hipLaunchKernelGGL((real_matvec_kernel<T, RC, LPR, XT, BLK>), grid, dim3(256), 0, s, npos,
                       d_rowptr, d_colidx, d_pos, d_rdi, box, xi, b, alpha, X, beta, Y);
    SDT_HIP_CHECK(hipGetLastError());
}

// columns per block of the blocked copy (0: X is gathered in place)
template <typename T>
int real_block_cols(int b) {
    if (b < 8) return 0;
    return (sizeof(T) == 4 && b >= 16) ? 16 : 8;
}

template <typename T>
size_t real_scratch_bytes(int npos, int b) {
    const int rc = real_block_cols<T>(b);
    if (rc == 0) return 0;
    return (size_t)ceil_div(b, rc) * rc * 3 * npos * sizeof(T);
}

template <typename T>
void real_block_copy(int npos, int b, const double* X, void* scratch, hipStream_t s) {
    const int rc = real_block_cols<T>(b);
    SDT_REQUIRE(rc > 0 && scratch != nullptr, SDT_EINVAL, "real space: no blocked copy for b");
    const int nblk = ceil_div(b, rc);
    const long long n = (long long)nblk * npos * 3 * rc;
    const dim3 grid(std::min<long long>(ceil_div(n, 256), 8192));
    if (rc == 16)
        hipLaunchKernelGGL((block_copy_kernel<T, 16>), grid, dim3(256), 0, s, npos, b, nblk, X,
                           static_cast<T*>(scratch));
    else
        hipLaunchKernelGGL((block_copy_kernel<T, 8>), grid, dim3(256), 0, s, npos, b, nblk, X,
                           static_cast<T*>(scratch));
    SDT_HIP_CHECK(hipGetLastError());
}

template <typename T>
void real_matvec(int npos, const int* d_rowptr, const int* d_colidx, const double* d_pos,
                 const double* d_rdi, double box, double xi, int b, double alpha, const double* X,
                 double beta, double* Y, void* scratch, bool copied, hipStream_t s) {
    const int rc = real_block_cols<T>(b);
    SDT_REQUIRE(rc == 0 || scratch != nullptr, SDT_EINVAL, "real space: scratch buffer missing");
    if (rc > 0 && !copied) real_block_copy<T>(npos, b, X, scratch, s);
    if (rc == 16) {
        if constexpr (sizeof(T) == 4)
            launch_real<T, 16, T, true>(npos, d_rowptr, d_colidx, d_pos, d_rdi, box, xi, b, alpha,
                                        static_cast<const T*>(scratch), beta, Y, s);
    } else if (rc == 8) {
        launch_real<T, 8, T, true>(npos, d_rowptr, d_colidx, d_pos, d_rdi, box, xi, b, alpha,
                                   static_cast<const T*>(scratch), beta, Y, s);
    } else if (b == 1) {
        launch_real<T, 1, double, false>(npos, d_rowptr, d_colidx, d_pos, d_rdi, box, xi, b, alpha,
                                         X, beta, Y, s);
    } else {
        launch_real<T, 4, double, false>(npos, d_rowptr, d_colidx, d_pos, d_rdi, box, xi, b, alpha,
                                         X, beta, Y, s);
    }
}

// ------------------------------------------------------------ overlaps -----
// Pairs closer than a_i + a_j carry the overlap correction of
// src/mob/mob_spme.cc:170-175, which cancels the 1/r^3 Ewald terms: they are
// listed once per update and always evaluated in FP64.
__device__ __forceinline__ bool overlapping(int i, int j, const double* pos, const double* rdi,
                                            double box) {
    if (i == j) return false;
    const double dx = remainder(pos[3 * j] - pos[3 * i], box);
    const double dy = remainder(pos[3 * j + 1] - pos[3 * i + 1], box);
    const double dz = remainder(pos[3 * j + 2] - pos[3 * i + 2], box);
    const double sab = rdi[i] + rdi[j];
    return pair_rr(dx, dy, dz) < sab * sab;
}

__global__ void overlap_count_kernel(int npos, const int* __restrict__ rowptr,
                                     const int* __restrict__ colidx, const double* __restrict__ pos,
                                     const double* __restrict__ rdi, double box,
                                     int* __restrict__ cnt, int* __restrict__ has) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npos) return;
    int c = 0;
    for (int k = rowptr[i]; k < rowptr[i + 1]; k++) c += overlapping(i, colidx[k], pos, rdi, box);
    cnt[i] = c;
    has[i] = c > 0;
}

__global__ void overlap_fill_kernel(int npos, const int* __restrict__ rowptr,
                                    const int* __restrict__ colidx, const double* __restrict__ pos,
                                    const double* __restrict__ rdi, double box,
                                    const int* __restrict__ orp, const int* __restrict__ hrp,
                                    int* __restrict__ ocol, int* __restrict__ orows) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npos) return;
    int o = orp[i];
    if (o == orp[i + 1]) return;
    orows[hrp[i]] = i;
    for (int k = rowptr[i]; k < rowptr[i + 1]; k++) {
        const int j = colidx[k];
        if (overlapping(i, j, pos, rdi, box)) ocol[o++] = j;
    }
}

void RealOverlap::build(int npos, const int* d_rowptr, const int* d_colidx, const double* d_pos,
                        const double* d_rdi, double box, hipStream_t s) {
    cnt.reserve((size_t)npos);
    has.reserve((size_t)npos);
    rowptr.reserve((size_t)npos + 1);
    hrp.reserve((size_t)npos + 1);
    const unsigned g = ceil_div((size_t)npos, 256);
    hipLaunchKernelGGL(overlap_count_kernel, dim3(g), dim3(256), 0, s, npos, d_rowptr, d_colidx,
                       d_pos, d_rdi, box, cnt.get(), has.get());
    SDT_HIP_CHECK(hipGetLastError());
    exclusive_scan(cnt.get(), npos, rowptr.get(), scratch, s);
    exclusive_scan(has.get(), npos, hrp.get(), scratch, s);
    int tot[2] = {0, 0};
    SDT_HIP_CHECK(hipMemcpyAsync(&tot[0], rowptr.get() + npos, sizeof(int), hipMemcpyDeviceToHost, s));
    SDT_HIP_CHECK(hipMemcpyAsync(&tot[1], hrp.get() + npos, sizeof(int), hipMemcpyDeviceToHost, s));
    SDT_HIP_CHECK(hipStreamSynchronize(s));
    npairs = tot[0];
    nrows = tot[1];
    if (nrows == 0) return;
    col.reserve((size_t)npairs);
    rows.reserve((size_t)nrows);
    hipLaunchKernelGGL(overlap_fill_kernel, dim3(g), dim3(256), 0, s, npos, d_rowptr, d_colidx,
                       d_pos, d_rdi, box, rowptr.get(), hrp.get(), col.get(), rows.get());
    SDT_HIP_CHECK(hipGetLastError());
}

// Y[(3 i + c) b + r] += alpha sum over the overlapping pairs of row i, one
// thread per (listed row, block of kOvRC columns): each pair's FP64 block is
// evaluated once per thread and applied to its columns; runs after
// real_matvec_kernel.
constexpr int kOvRC = 4;

__global__ void __launch_bounds__(256)
real_overlap_kernel(int nrows, const int* __restrict__ rows,
                                    const int* __restrict__ orp, const int* __restrict__ ocol,
                                    const double* __restrict__ pos, const double* __restrict__ rdi,
                                    double box, double xi, int b, double alpha,
                                    const double* __restrict__ X, double* __restrict__ Y) {
    const int nblk = (b + kOvRC - 1) / kOvRC;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)nrows * nblk) return;
    const int i = rows[t / nblk];
    const int r0 = (int)(t % nblk) * kOvRC;
    const int rc = min(kOvRC, b - r0);
    const double aa = rdi[i];
    double acc[3][kOvRC];
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
        for (int r = 0; r < kOvRC; r++) acc[c][r] = 0.0;
    for (int k = orp[i]; k < orp[i + 1]; k++) {
        const int j = ocol[k];
        const double dx = remainder(pos[3 * j] - pos[3 * i], box);
        const double dy = remainder(pos[3 * j + 1] - pos[3 * i + 1], box);
        const double dz = remainder(pos[3 * j + 2] - pos[3 * i + 2], box);
        const double ab = rdi[j];
        const double rr = dx * dx + dy * dy + dz * dz;
        const double rv = sqrt(rr);
        double xa, ya;
        ewald_scalars<double>(xi, rv, aa, ab, &xa, &ya);
        const double s2 = aa * aa + ab * ab;
        xa += -(1.5 - s2 * 0.5 / rr) / rv + 2.0 / (aa + ab) - 3.0 * rv / 8.0 / s2;
        ya += -(0.75 + s2 * 0.25 / rr) / rv + 2.0 / (aa + ab) - 9.0 * rv / 16.0 / s2;
        const double ex = dx / rv, ey = dy / rv, ez = dz / rv;
        const double d = xa - ya;
        const double bxx = d * ex * ex + ya, bxy = d * ex * ey, bxz = d * ex * ez;
        const double byy = d * ey * ey + ya, byz = d * ey * ez, bzz = d * ez * ez + ya;
        const double* xj = X + (size_t)(3 * j) * b + r0;
#pragma unroll
        for (int r = 0; r < kOvRC; r++) {
            if (r < rc) {
                const double v0 = xj[r], v1 = xj[b + r], v2 = xj[2 * b + r];
                acc[0][r] += bxx * v0 + bxy * v1 + bxz * v2;
                acc[1][r] += bxy * v0 + byy * v1 + byz * v2;
                acc[2][r] += bxz * v0 + byz * v1 + bzz * v2;
            }
        }
    }
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
        for (int r = 0; r < kOvRC; r++)
            if (r < rc) Y[(size_t)(3 * i + c) * b + r0 + r] += alpha * acc[c][r];
}

void real_overlap(const RealOverlap& ov, const double* d_pos, const double* d_rdi, double box,
                  double xi, int b, double alpha, const double* X, double* Y, hipStream_t s) {
    if (ov.nrows == 0) return;
    const size_t nt = (size_t)ov.nrows * ((b + kOvRC - 1) / kOvRC);
    hipLaunchKernelGGL(real_overlap_kernel, dim3(ceil_div(nt, 256)), dim3(256), 0, s, ov.nrows,
                       ov.rows.get(), ov.rowptr.get(), ov.col.get(), d_pos, d_rdi, box, xi, b,
                       alpha, X, Y);
    SDT_HIP_CHECK(hipGetLastError());
}

template int real_block_cols<float>(int);
template int real_block_cols<double>(int);
template size_t real_scratch_bytes<float>(int, int);
template size_t real_scratch_bytes<double>(int, int);
template void real_block_copy<float>(int, int, const double*, void*, hipStream_t);
template void real_block_copy<double>(int, int, const double*, void*, hipStream_t);
template void real_matvec<float>(int, const int*, const int*, const double*, const double*, double,
                                 double, int, double, const double*, double, double*, void*, bool,
                                 hipStream_t);
template void real_matvec<double>(int, const int*, const int*, const double*, const double*, double,
                                  double, int, double, const double*, double, double*, void*, bool,
                                  hipStream_t);

}  // namespace sdt
