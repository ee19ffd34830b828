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

// Y[(3 i + c) b + r] = beta Y + alpha sum_j B_ij X_j for r in [r0, r0 + RC)
template <typename T, int RC>
__global__ void __launch_bounds__(256)
real_matvec_kernel(int npos, const int* __restrict__ rowptr, const int* __restrict__ colidx,
                   const double* __restrict__ pos, const double* __restrict__ rdi, double box,
                   double xi, int b, double alpha, const double* __restrict__ X, double beta,
                   double* __restrict__ Y) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npos) return;
    const int r0 = blockIdx.y * RC;
    const int rc = min(RC, b - r0);
    const double x1 = pos[3 * i], y1 = pos[3 * i + 1], z1 = pos[3 * i + 2];
    const double aa_d = rdi[i];
    const T aa = (T)aa_d;
    const T txi = (T)xi;
    T acc[3][RC];
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
        for (int r = 0; r < RC; r++) acc[c][r] = T(0);
    // self block: (1/a - xi/sqrt(pi) (6 - 40/3 xi^2 a^2)) I
    const T self_a = (T)(1.0 / aa_d - xi / sqrt(M_PI) * (6.0 - 40.0 / 3.0 * xi * xi * aa_d * aa_d));
    for (int k = rowptr[i]; k < rowptr[i + 1]; k++) {
        const int j = colidx[k];
        T bxx, bxy, bxz, byy, byz, bzz;
        if (j == i) {
            bxx = byy = bzz = self_a;
            bxy = bxz = byz = T(0);
        } else {
            const T rx = (T)remainder(pos[3 * j] - x1, box);
            const T ry = (T)remainder(pos[3 * j + 1] - y1, box);
            const T rz = (T)remainder(pos[3 * j + 2] - z1, box);
            const T rr = rx * rx + ry * ry + rz * rz;
            const T r = sqrt(rr);
            const T ab = (T)rdi[j];
            T xa, ya;
            ewald_scalars<T>(txi, r, aa, ab, &xa, &ya);
            if (r < aa + ab) {
                const T s2 = aa * aa + ab * ab;
                xa += -(T(1.5) - s2 * T(0.5) / rr) / r + T(2) / (aa + ab) - T(3) * r / T(8) / s2;
                ya += -(T(0.75) + s2 * T(0.25) / rr) / r + T(2) / (aa + ab) - T(9) * r / T(16) / s2;
            }
            const T ir = T(1) / r;
            const T ex = rx * ir, ey = ry * ir, ez = rz * ir;
            const T d = xa - ya;
            bxx = d * ex * ex + ya; bxy = d * ex * ey; bxz = d * ex * ez;
            byy = d * ey * ey + ya; byz = d * ey * ez; bzz = d * ez * ez + ya;
        }
        const double* xj = X + (size_t)(3 * j) * b + r0;
#pragma unroll
        for (int r = 0; r < RC; r++) {
            if (r < rc) {
                const T v0 = (T)xj[r], v1 = (T)xj[b + r], v2 = (T)xj[2 * b + r];
                acc[0][r] += bxx * v0 + bxy * v1 + bxz * v2;
                acc[1][r] += bxy * v0 + byy * v1 + byz * v2;
                acc[2][r] += bxz * v0 + byz * v1 + bzz * v2;
            }
        }
    }
#pragma unroll
    for (int c = 0; c < 3; c++) {
        double* yi = Y + (size_t)(3 * i + c) * b + r0;
#pragma unroll
        for (int r = 0; r < RC; r++) {
            if (r < rc) yi[r] = (beta == 0.0 ? 0.0 : beta * yi[r]) + alpha * (double)acc[c][r];
        }
    }
}

template <typename T>
void real_matvec(int npos, const int* d_rowptr, const int* d_colidx, const double* d_pos,
                 const double* d_rdi, double box, double xi, int b, double alpha, const double* X,
                 double beta, double* Y, hipStream_t s) {
    constexpr int RC = 4;
    dim3 grid(ceil_div(npos, 256), ceil_div(b, RC));
    hipLaunchKernelGGL((real_matvec_kernel<T, RC>), grid, dim3(256), 0, s, npos, d_rowptr,
                       d_colidx, d_pos, d_rdi, box, xi, b, alpha, X, beta, Y);
    SDT_HIP_CHECK(hipGetLastError());
}

template void real_matvec<float>(int, const int*, const int*, const double*, const double*, double,
                                 double, int, double, const double*, double, double*, hipStream_t);
template void real_matvec<double>(int, const int*, const int*, const double*, const double*, double,
                                  double, int, double, const double*, double, double*, hipStream_t);

}  // namespace sdt
