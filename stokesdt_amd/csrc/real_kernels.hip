// Matrix-free real-space RPY-Ewald matvec for gfx950 (FP32 and FP64).
//
// Reference: the real-space part of MobSpme is a stored 3x3-block CSR built
// once per update (MobSpme::BuildSparseReal, src/mob/mob_spme.cc:135-202,
// with scalars_ewald_F from src/mob/ewald.h:55-85) and applied by SpMV3x3
// (src/util/sparse.cc:79-153). Here the blocks are recomputed inside the
// matvec from FP64 position differences, for a chunk of right-hand sides per
// thread, so a matvec streams only the column indices and the input vector.
#include <cmath>
#include "real_gpu.hpp"

namespace sdt {

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

// Y[(3 i + c) b + r] = beta Y + alpha sum_j B_ij X_j for r in [r0, r0 + RC).
// LPR consecutive lanes share a row, each taking every LPR-th pair, so a
// row's gathers are in flight together; partial sums are combined with
// cross-lane shuffles.
template <typename T, int RC, int LPR>
__global__ void __launch_bounds__(256)
real_matvec_kernel(int npos, const int* __restrict__ rowptr, const int* __restrict__ colidx,
                   const double* __restrict__ pos, const double* __restrict__ rdi, double box,
                   double xi, int b, double alpha, const double* __restrict__ X, double beta,
                   double* __restrict__ Y) {
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = tid / LPR, sub = tid % LPR;
    const bool valid = i < npos;
    const int r0 = blockIdx.y * RC;
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
            const double* xj = X + (size_t)(3 * j) * b + r0;
            double xv[3][RC];
#pragma unroll
            for (int r = 0; r < RC; r++) {
                xv[0][r] = xv[1][r] = xv[2][r] = 0.0;
                if (r < rc) {
                    xv[0][r] = xj[r];
                    xv[1][r] = xj[b + r];
                    xv[2][r] = xj[2 * b + r];
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
                const double rr_d = dx * dx + dy * dy + dz * dz;
                if (rr_d < (aa_d + ab_d) * (aa_d + ab_d)) {
                    // overlapping pair: the overlap correction cancels the
                    // 1/r^3 Ewald terms, so this rare block is formed in FP64
                    const double r = sqrt(rr_d);
                    double xa, ya;
                    ewald_scalars<double>(xi, r, aa_d, ab_d, &xa, &ya);
                    const double s2 = aa_d * aa_d + ab_d * ab_d;
                    xa += -(1.5 - s2 * 0.5 / rr_d) / r + 2.0 / (aa_d + ab_d) - 3.0 * r / 8.0 / s2;
                    ya += -(0.75 + s2 * 0.25 / rr_d) / r + 2.0 / (aa_d + ab_d) - 9.0 * r / 16.0 / s2;
                    const double ex = dx / r, ey = dy / r, ez = dz / r;
                    const double d = xa - ya;
                    bxx = (T)(d * ex * ex + ya); bxy = (T)(d * ex * ey); bxz = (T)(d * ex * ez);
                    byy = (T)(d * ey * ey + ya); byz = (T)(d * ey * ez); bzz = (T)(d * ez * ez + ya);
                } else {
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

template <typename T>
void real_matvec(int npos, const int* d_rowptr, const int* d_colidx, const double* d_pos,
                 const double* d_rdi, double box, double xi, int b, double alpha, const double* X,
                 double beta, double* Y, hipStream_t s) {
    constexpr int LPR = 8;
    if (b >= 8) {
        constexpr int RC = 8;
        dim3 grid(ceil_div((size_t)npos * LPR, 256), ceil_div(b, RC));
        hipLaunchKernelGGL((real_matvec_kernel<T, RC, LPR>), grid, dim3(256), 0, s, npos,
                           d_rowptr, d_colidx, d_pos, d_rdi, box, xi, b, alpha, X, beta, Y);
    } else {
        constexpr int RC = 4;
        dim3 grid(ceil_div((size_t)npos * LPR, 256), ceil_div(b, RC));
        hipLaunchKernelGGL((real_matvec_kernel<T, RC, LPR>), grid, dim3(256), 0, s, npos,
                           d_rowptr, d_colidx, d_pos, d_rdi, box, xi, b, alpha, X, beta, Y);
    }
    SDT_HIP_CHECK(hipGetLastError());
}

template void real_matvec<float>(int, const int*, const int*, const double*, const double*, double,
                                 double, int, double, const double*, double, double*, hipStream_t);
template void real_matvec<double>(int, const int*, const int*, const double*, const double*, double,
                                  double, int, double, const double*, double, double*, hipStream_t);

}  // namespace sdt
