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
#include <vector>
#include "ewald_scalars.hpp"
#include "real_gpu.hpp"
#include "sort_gpu.hpp"

namespace sdt {

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
        const double ibox = 1.0 / box;
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
                // listed pairs lie within the cutoff, so the min image equals
                // remainder(., box) bit for bit
                const double dx = min_image(pos[3 * j] - x1, box, ibox);
                const double dy = min_image(pos[3 * j + 1] - y1, box, ibox);
                const double dz = min_image(pos[3 * j + 2] - z1, box, ibox);
                const double ab_d = rdi[j];
                const double sab = aa_d + ab_d;
                if (pair_rr(dx, dy, dz) < sab * sab) continue;   // overlap pass
                const T rx = (T)dx, ry = (T)dy, rz = (T)dz;
                const T rr = rx * rx + ry * ry + rz * rz;
                const T r = sqrt(rr);
                const T ab = (T)ab_d;
                const T ir = T(1) / r;
                T xa, ya;
                ewald_scalars_ir<T>(txi, r, ir, aa, ab, &xa, &ya);
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
    // listed pairs: the same min image as the matvec's (remainder's bits)
    const double ibox = 1.0 / box;
    const double dx = min_image(pos[3 * j] - pos[3 * i], box, ibox);
    const double dy = min_image(pos[3 * j + 1] - pos[3 * i + 1], box, ibox);
    const double dz = min_image(pos[3 * j + 2] - pos[3 * i + 2], box, ibox);
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
    const double ibox = 1.0 / box;
    for (int k = orp[i]; k < orp[i + 1]; k++) {
        const int j = ocol[k];
        const double dx = min_image(pos[3 * j] - pos[3 * i], box, ibox);
        const double dy = min_image(pos[3 * j + 1] - pos[3 * i + 1], box, ibox);
        const double dz = min_image(pos[3 * j + 2] - pos[3 * i + 2], box, ibox);
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

// ------------------------------------------------------ dense Ewald matrix --
// The Ewald-summed RPY matrix of MobDebug::GetMatrix (src/mob/mob_debug.cc),
// i.e. EwaldKernel (src/mob/ewald.cc:86-330) over the tables of
// InitRealTable / InitRecipTable (:643-797) with the truncation of
// XitoRmax / XitoKmax (:19-83). One thread per 3x3 block (i, j), j fastest;
// the image and k-vector tables are read at wave-uniform addresses.

namespace {

double ewald_real_tail(double r, double xi) {
    const double xir = xi * r, xir2 = xir * xir, r2 = r * r;
    return erfc(xir) * (0.75 + 1.5 / r2) / r +
           exp(-xir2) * xi / sqrt(M_PI) *
               (4.5 + 3.0 * xir2 + (3.0 + xir2 * (14.0 + xir2 * (4.0 + xir2))) / r2);
}

double ewald_recip_tail(double k, double xi) {
    const double kx = k / xi, kx2 = kx * kx, k2 = k * k;
    return 6.0 * M_PI * exp(-kx2 / 4.0) * (1.0 + k2 / 3.0) / k2 *
           (1.0 + kx2 * (1.0 / 4.0 + kx2 / 8.0));
}

// shrink the cutoff by 1 % while the bound (one term at the cutoff plus six
// one period further) stays below tol; xi is detuned by the safety factor
double ewald_rmax(double xi, double tol, double box) {
    const double x0 = xi / 1.15;
    double rmax = 10.0 * sqrt(-log(tol)) / xi;
    while (ewald_real_tail(rmax * 0.99, x0) + 6.0 * ewald_real_tail(rmax * 0.99 + box, x0) < tol)
        rmax *= 0.99;
    return rmax;
}

double ewald_kmax(double xi, double tol, double box) {
    const double x0 = xi * 1.15, dk = 2.0 * M_PI / box;
    double kmax = 20.0 * sqrt(-log(tol)) * xi;
    while (ewald_recip_tail(kmax * 0.99, x0) + 6.0 * ewald_recip_tail(kmax * 0.99 + dk, x0) < tol)
        kmax *= 0.99;
    return kmax;
}

}  // namespace

constexpr int kEwaldK = 8;   // k1, k2, k3, ex, ey, ez, |k|^2, exp factor

__global__ void __launch_bounds__(256)
ewald_block_kernel(int npos, const double* __restrict__ pos, const double* __restrict__ rdi,
                   double box, double xi, double rmax2, int nr, const double* __restrict__ img,
                   int nk, const double* __restrict__ ktab, int parts, long long ld,
                   double* __restrict__ mat) {
    // blocks with j >= i, each stored at (i, j) and (j, i), as EwaldKernel
    // does: the matrix is symmetric bit for bit
    const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
    if (j >= npos || j < i) return;
    const double aa = rdi[i], ab = rdi[j], aa2 = aa * aa, ab2 = ab * ab;
    const double x0 = remainder(pos[3 * j] - pos[3 * i], box);
    const double y0 = remainder(pos[3 * j + 1] - pos[3 * i + 1], box);
    const double z0 = remainder(pos[3 * j + 2] - pos[3 * i + 2], box);
    double s[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    auto add = [&](double xa, double ya, double ex, double ey, double ez) {
        const double d = xa - ya;
        s[0] += d * ex * ex + ya;
        s[1] += d * ex * ey;
        s[2] += d * ex * ez;
        s[3] += d * ey * ey + ya;
        s[4] += d * ey * ez;
        s[5] += d * ez * ez + ya;
    };
    if (parts & 1) {
        if (i == j) {
            const double self_a =
                1.0 / aa - xi / sqrt(M_PI) * (6.0 - 40.0 / 3.0 * xi * xi * aa2);
            s[0] = s[3] = s[5] = self_a;
        }
        for (int m = 0; m < nr; m++) {
            const double xx = x0 + img[3 * m], yy = y0 + img[3 * m + 1], zz = z0 + img[3 * m + 2];
            const double rr = xx * xx + yy * yy + zz * zz;
            if (rr == 0.0 || rr > rmax2) continue;
            const double r = sqrt(rr);
            const double ex = xx / r, ey = yy / r, ez = zz / r;
            if (r < aa + ab) {
                // overlapping pair: the RPY form for overlapping spheres
                // replaces the far-field one (src/mob/ewald.cc:177-195)
                const double s2 = aa2 + ab2;
                add(-(1.5 - s2 * 0.5 / rr) / r + 2.0 / (aa + ab) - 3.0 * r / 8.0 / s2,
                    -(0.75 + s2 * 0.25 / rr) / r + 2.0 / (aa + ab) - 9.0 * r / 16.0 / s2, ex, ey,
                    ez);
            }
            double xa, ya;
            ewald_scalars<double>(xi, r, aa, ab, &xa, &ya);
            add(xa, ya, ex, ey, ez);
        }
    }
    if (parts & 2) {
        const double s2 = aa2 + ab2;
        for (int m = 0; m < nk; m++) {
            const double* k = ktab + (size_t)m * kEwaldK;
            const double ya = 6.0 * (1.0 - k[6] * s2 / 6.0) * k[7];
            const double cf = cos(k[0] * x0 + k[1] * y0 + k[2] * z0) * ya;
            s[0] += cf * (1.0 - k[3] * k[3]);
            s[1] -= cf * k[3] * k[4];
            s[2] -= cf * k[3] * k[5];
            s[3] += cf * (1.0 - k[4] * k[4]);
            s[4] -= cf * k[4] * k[5];
            s[5] += cf * (1.0 - k[5] * k[5]);
        }
    }
    // the blocks are symmetric, so row-major rows are the column-major columns
    auto put = [&](int r, int c) {
        double* o = mat + (size_t)(3 * r) * ld + 3 * c;
        o[0] = s[0]; o[1] = s[1]; o[2] = s[2];
        o += ld;
        o[0] = s[1]; o[1] = s[3]; o[2] = s[4];
        o += ld;
        o[0] = s[2]; o[1] = s[4]; o[2] = s[5];
    };
    put(i, j);
    if (j != i) put(j, i);
}

void EwaldMatrixGpu::setup(double box_, double xi_, double tol, int mode, hipStream_t s) {
    SDT_REQUIRE(box_ > 0.0 && tol > 0.0 && xi_ > 0.0, SDT_EINVAL,
                "dense Ewald: L, tol and xi must be positive");
    SDT_REQUIRE(mode >= 0 && mode <= 2, SDT_EINVAL, "dense Ewald: mode must be full, real or recip");
    box = box_;
    xi = xi_;
    parts = mode == 0 ? 3 : (mode == 1 ? 1 : 2);
    // periodic images whose shifted cell can hold a pair within rmax
    std::vector<double> img;
    rmax = 0.0;
    if (parts & 1) {
        rmax = ewald_rmax(xi, tol, box);
        const int mi = (int)(rmax / box) + 1;
        const double reach = rmax + sqrt(3.0) * box;
        auto near = [&](int m) { return box * m - (m > 0 ? 0.5 * box : (m < 0 ? -0.5 * box : 0.0)); };
        for (int a = -mi; a <= mi; a++)
            for (int b = -mi; b <= mi; b++)
                for (int c = -mi; c <= mi; c++) {
                    const double dx = near(a), dy = near(b), dz = near(c);
                    if (sqrt(dx * dx + dy * dy + dz * dz) <= reach) {
                        img.push_back(box * a);
                        img.push_back(box * b);
                        img.push_back(box * c);
                    }
                }
    }
    // k vectors 2 pi m / L with 0 < |k| <= kmax
    std::vector<double> kt;
    if (parts & 2) {
        const double kmax = ewald_kmax(xi, tol, box);
        const int mk = (int)(kmax * box / (2.0 * M_PI));
        const double pivol = M_PI / (box * box * box);
        for (int a = -mk; a <= mk; a++)
            for (int b = -mk; b <= mk; b++)
                for (int c = -mk; c <= mk; c++) {
                    if (a == 0 && b == 0 && c == 0) continue;
                    const double k1 = 2.0 * M_PI * a / box, k2 = 2.0 * M_PI * b / box,
                                 k3 = 2.0 * M_PI * c / box;
                    const double kk = k1 * k1 + k2 * k2 + k3 * k3, k = sqrt(kk);
                    if (k > kmax) continue;
                    const double q = kk / (4.0 * xi * xi);
                    const double row[kEwaldK] = {k1, k2, k3, k1 / k, k2 / k, k3 / k, kk,
                                                 pivol * (1.0 + q * (1.0 + 2.0 * q)) / kk * exp(-q)};
                    kt.insert(kt.end(), row, row + kEwaldK);
                }
    }
    nr = (int)(img.size() / 3);
    nk = (int)(kt.size() / kEwaldK);
    d_img.reserve(std::max<size_t>(1, img.size()));
    d_k.reserve(std::max<size_t>(1, kt.size()));
    if (nr)
        SDT_HIP_CHECK(hipMemcpyAsync(d_img.get(), img.data(), sizeof(double) * img.size(),
                                     hipMemcpyHostToDevice, s));
    if (nk)
        SDT_HIP_CHECK(hipMemcpyAsync(d_k.get(), kt.data(), sizeof(double) * kt.size(),
                                     hipMemcpyHostToDevice, s));
    SDT_HIP_CHECK(hipStreamSynchronize(s));
}

void EwaldMatrixGpu::build(int npos, const double* d_pos, const double* d_rdi, long long ld,
                           double* d_mat, hipStream_t s) const {
    SDT_REQUIRE(parts != 0, SDT_ESTATE, "dense Ewald: build before setup");
    SDT_REQUIRE(npos > 0 && npos <= 30000, SDT_EINVAL, "dense Ewald: npos must be in [1, 30000]");
    SDT_REQUIRE(ld >= 3LL * npos, SDT_EINVAL, "dense Ewald: ldm must be at least 3*npos");
    const dim3 grid((unsigned)((npos + 255) / 256), (unsigned)npos);
    ewald_block_kernel<<<grid, 256, 0, s>>>(npos, d_pos, d_rdi, box, xi, rmax * rmax, nr,
                                            d_img.get(), nk, d_k.get(), parts, ld, d_mat);
    SDT_HIP_CHECK(hipGetLastError());
}

// ------------------------------------------------- Ewald rows, no matrix --
// EwaldVectorKernel (src/mob/ewald.cc:333-536) evaluates sum_j M_ij f_j pair
// by pair, both sums per pair. For a sample of rows the reciprocal sum is
// rewritten over structure factors of f, which needs each k vector once per
// particle instead of once per pair:
//   sum_j cos(k.(r_j - r_i)) ya_ij f_j = Re[e^{-i k.r_i} (c0 S0(k) + c2 S2(k))]
// with ya_ij = 6 (1 - k^2 (a_i^2 + a_j^2) / 6) kexp, c0 = (6 - k^2 a_i^2) kexp,
// c2 = -k^2 kexp, S0(k) = sum_j f_j e^{i k.r_j}, S2(k) = sum_j a_j^2 f_j
// e^{i k.r_j}; the projector I - k^ k^T acts on the summed vector. The real
// sum is the per-pair sum of the dense kernel over the same image table.

constexpr int kSk = 12;   // Re S0, Im S0, Re S2, Im S2 (3 components each)

template <int N>
__device__ inline void block_sum_256(double (&a)[N], double* lds) {
#pragma unroll
    for (int c = 0; c < N; c++)
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) a[c] += __shfl_xor(a[c], off, 64);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int c = 0; c < N; c++) lds[w * N + c] = a[c];
    __syncthreads();
    if (threadIdx.x == 0)
#pragma unroll
        for (int c = 0; c < N; c++) a[c] = lds[c] + lds[N + c] + lds[2 * N + c] + lds[3 * N + c];
}

// one workgroup per k vector: its structure factors over all particles
__global__ void __launch_bounds__(256)
ewald_sk_kernel(int npos, const double* __restrict__ pos, const double* __restrict__ rdi,
                const double* __restrict__ f, const double* __restrict__ ktab,
                double* __restrict__ sk) {
    __shared__ double lds[4 * kSk];
    const double* k = ktab + (size_t)blockIdx.x * kEwaldK;
    const double k1 = k[0], k2 = k[1], k3 = k[2];
    double a[kSk];
#pragma unroll
    for (int c = 0; c < kSk; c++) a[c] = 0.0;
    for (int j = threadIdx.x; j < npos; j += 256) {
        double sn, cs;
        sincos(k1 * pos[3 * j] + k2 * pos[3 * j + 1] + k3 * pos[3 * j + 2], &sn, &cs);
        const double a2 = rdi[j] * rdi[j];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            const double fc = f[3 * j + c];
            a[c] += cs * fc;
            a[3 + c] += sn * fc;
            a[6 + c] += a2 * cs * fc;
            a[9 + c] += a2 * sn * fc;
        }
    }
    block_sum_256<kSk>(a, lds);
    if (threadIdx.x == 0)
#pragma unroll
        for (int c = 0; c < kSk; c++) sk[(size_t)blockIdx.x * kSk + c] = a[c];
}

// one workgroup per sampled row
__global__ void __launch_bounds__(256)
ewald_rows_kernel(int npos, const double* __restrict__ pos, const double* __restrict__ rdi,
                  const double* __restrict__ f, double box, double xi, double rmax2, int nr,
                  const double* __restrict__ img, int nk, const double* __restrict__ ktab,
                  const double* __restrict__ sk, const int* __restrict__ rows,
                  double* __restrict__ v) {
    __shared__ double lds[4 * 3];
    const int i = rows[blockIdx.x];
    const double aa = rdi[i], aa2 = aa * aa;
    const double px = pos[3 * i], py = pos[3 * i + 1], pz = pos[3 * i + 2];
    double acc[3] = {0.0, 0.0, 0.0};
    for (int j = threadIdx.x; j < npos; j += 256) {
        const double ab = rdi[j], ab2 = ab * ab;
        const double x0 = remainder(pos[3 * j] - px, box);
        const double y0 = remainder(pos[3 * j + 1] - py, box);
        const double z0 = remainder(pos[3 * j + 2] - pz, box);
        double s[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        auto add = [&](double xa, double ya, double ex, double ey, double ez) {
            const double d = xa - ya;
            s[0] += d * ex * ex + ya;
            s[1] += d * ex * ey;
            s[2] += d * ex * ez;
            s[3] += d * ey * ey + ya;
            s[4] += d * ey * ez;
            s[5] += d * ez * ez + ya;
        };
        if (j == i) {
            const double self_a =
                1.0 / aa - xi / sqrt(M_PI) * (6.0 - 40.0 / 3.0 * xi * xi * aa2);
            s[0] = s[3] = s[5] = self_a;
        }
        for (int m = 0; m < nr; m++) {
            const double xx = x0 + img[3 * m], yy = y0 + img[3 * m + 1], zz = z0 + img[3 * m + 2];
            const double rr = xx * xx + yy * yy + zz * zz;
            if (rr == 0.0 || rr > rmax2) continue;
            const double r = sqrt(rr);
            const double ex = xx / r, ey = yy / r, ez = zz / r;
            if (r < aa + ab) {
                const double s2 = aa2 + ab2;
                add(-(1.5 - s2 * 0.5 / rr) / r + 2.0 / (aa + ab) - 3.0 * r / 8.0 / s2,
                    -(0.75 + s2 * 0.25 / rr) / r + 2.0 / (aa + ab) - 9.0 * r / 16.0 / s2, ex, ey,
                    ez);
            }
            double xa, ya;
            ewald_scalars<double>(xi, r, aa, ab, &xa, &ya);
            add(xa, ya, ex, ey, ez);
        }
        const double f0 = f[3 * j], f1 = f[3 * j + 1], f2 = f[3 * j + 2];
        acc[0] += s[0] * f0 + s[1] * f1 + s[2] * f2;
        acc[1] += s[1] * f0 + s[3] * f1 + s[4] * f2;
        acc[2] += s[2] * f0 + s[4] * f1 + s[5] * f2;
    }
    for (int m = threadIdx.x; m < nk; m += 256) {
        const double* k = ktab + (size_t)m * kEwaldK;
        const double* S = sk + (size_t)m * kSk;
        double sn, cs;
        sincos(k[0] * px + k[1] * py + k[2] * pz, &sn, &cs);
        const double c0 = (6.0 - k[6] * aa2) * k[7], c2 = -k[6] * k[7];
        double w[3];
#pragma unroll
        for (int c = 0; c < 3; c++)
            w[c] = cs * (c0 * S[c] + c2 * S[6 + c]) + sn * (c0 * S[3 + c] + c2 * S[9 + c]);
        const double dot = k[3] * w[0] + k[4] * w[1] + k[5] * w[2];
#pragma unroll
        for (int c = 0; c < 3; c++) acc[c] += w[c] - k[3 + c] * dot;
    }
    block_sum_256<3>(acc, lds);
    if (threadIdx.x == 0)
#pragma unroll
        for (int c = 0; c < 3; c++) v[3 * blockIdx.x + c] = acc[c];
}

void EwaldMatrixGpu::rows_mul(int npos, const double* d_pos, const double* d_rdi, int nrows,
                              const int* d_rows, const double* d_f, double* d_sk, double* d_v,
                              hipStream_t s) const {
    SDT_REQUIRE(parts == 3, SDT_ESTATE, "Ewald rows: set up for the full sum first");
    if (nk > 0) {
        ewald_sk_kernel<<<nk, 256, 0, s>>>(npos, d_pos, d_rdi, d_f, d_k.get(), d_sk);
        SDT_HIP_CHECK(hipGetLastError());
    }
    ewald_rows_kernel<<<nrows, 256, 0, s>>>(npos, d_pos, d_rdi, d_f, box, xi, rmax * rmax, nr,
                                            d_img.get(), nk, d_k.get(), d_sk, d_rows, d_v);
    SDT_HIP_CHECK(hipGetLastError());
}

// The split costs nrows npos n_images distance tests plus the erfc terms of
// the pairs within rmax, against (npos + nrows) n_k phases: scan xi on a
// geometric grid and keep the cheapest.
double ewald_rows_xi(int npos, int nrows, double box, double tol) {
    const double rho = npos / (box * box * box), kcell = box / (2.0 * M_PI);
    double best = 0.0, best_cost = 0.0;
    for (int t = 0; t <= 160; t++) {
        const double xi = 0.3 / box * pow(400.0, t / 160.0);
        const double rmax = ewald_rmax(xi, tol, box), kmax = ewald_kmax(xi, tol, box);
        const double nimg = pow(2.0 * ((int)(rmax / box) + 1) + 1.0, 3.0);
        const double pairs = std::min(rho * 4.0 / 3.0 * M_PI * rmax * rmax * rmax, npos * nimg);
        const double nk = 4.0 / 3.0 * M_PI * pow(kmax * kcell, 3.0);
        const double cost = (double)nrows * npos * nimg + 20.0 * nrows * pairs +
                            10.0 * ((double)npos + nrows) * nk;
        if (t == 0 || cost < best_cost) {
            best = xi;
            best_cost = cost;
        }
    }
    return best;
}

void ewald_rows(int npos, const double* pos, const double* rdi, double box, double xi, double tol,
                int nrows, const int* rows, const double* f, double* v, hipStream_t s) {
    SDT_REQUIRE(npos > 0 && nrows > 0, SDT_EINVAL, "Ewald rows: npos and nrows must be positive");
    for (int r = 0; r < nrows; r++)
        SDT_REQUIRE(rows[r] >= 0 && rows[r] < npos, SDT_EINVAL,
                    "Ewald rows: row index out of range");
    if (xi <= 0.0) xi = ewald_rows_xi(npos, nrows, box, tol);
    EwaldMatrixGpu ew;
    ew.setup(box, xi, tol, 0, s);
    const size_t dim = (size_t)3 * npos;
    DevBuf<double> d_pos, d_rdi, d_f, d_sk, d_v;
    DevBuf<int> d_rows;
    d_pos.reserve(dim);
    d_rdi.reserve(npos);
    d_f.reserve(dim);
    d_sk.reserve((size_t)std::max(1, ew.nk) * kSk);
    d_v.reserve((size_t)3 * nrows);
    d_rows.reserve(nrows);
    SDT_HIP_CHECK(hipMemcpyAsync(d_pos.get(), pos, sizeof(double) * dim, hipMemcpyHostToDevice, s));
    SDT_HIP_CHECK(hipMemcpyAsync(d_rdi.get(), rdi, sizeof(double) * npos, hipMemcpyHostToDevice, s));
    SDT_HIP_CHECK(hipMemcpyAsync(d_f.get(), f, sizeof(double) * dim, hipMemcpyHostToDevice, s));
    SDT_HIP_CHECK(hipMemcpyAsync(d_rows.get(), rows, sizeof(int) * nrows, hipMemcpyHostToDevice, s));
    ew.rows_mul(npos, d_pos.get(), d_rdi.get(), nrows, d_rows.get(), d_f.get(), d_sk.get(),
                d_v.get(), s);
    SDT_HIP_CHECK(hipMemcpyAsync(v, d_v.get(), sizeof(double) * 3 * nrows, hipMemcpyDeviceToHost, s));
    SDT_HIP_CHECK(hipStreamSynchronize(s));
}

void ewald_matrix(int npos, const double* pos, const double* rdi, double box, double xi,
                  double tol, int mode, int ldm, double* mat, hipStream_t s) {
    SDT_REQUIRE(npos > 0 && npos <= 30000, SDT_EINVAL, "dense Ewald: npos must be in [1, 30000]");
    SDT_REQUIRE(ldm >= 3 * npos, SDT_EINVAL, "dense Ewald: ldm must be at least 3*npos");
    EwaldMatrixGpu ew;
    ew.setup(box, xi, tol, mode, s);
    const size_t dim = (size_t)3 * npos;
    DevBuf<double> d_pos, d_rdi, d_mat;
    d_pos.reserve(dim);
    d_rdi.reserve(npos);
    d_mat.reserve(dim * dim);
    SDT_HIP_CHECK(hipMemcpyAsync(d_pos.get(), pos, sizeof(double) * dim, hipMemcpyHostToDevice, s));
    SDT_HIP_CHECK(hipMemcpyAsync(d_rdi.get(), rdi, sizeof(double) * npos, hipMemcpyHostToDevice, s));
    ew.build(npos, d_pos.get(), d_rdi.get(), (long long)dim, d_mat.get(), s);
    SDT_HIP_CHECK(hipMemcpy2DAsync(mat, sizeof(double) * ldm, d_mat.get(), sizeof(double) * dim,
                                   sizeof(double) * dim, dim, hipMemcpyDeviceToHost, s));
    SDT_HIP_CHECK(hipStreamSynchronize(s));
}

}  // namespace sdt

