// Device helpers of the real-space RPY-Ewald operator shared by the
// matrix-free matvec, the dense Ewald matrix and the stored-block path:
// the min-image squared distance and the scalars xa, ya of one pair
// (scalars_ewald_F, src/mob/ewald.h:55-85).
#pragma once
#include <hip/hip_runtime.h>

namespace sdt {

// squared min-image distance with explicit fused operations, so that every
// kernel that classifies a pair as overlapping rounds identically
__device__ __forceinline__ double pair_rr(double dx, double dy, double dz) {
    return fma(dx, dx, fma(dy, dy, dz * dz));
}

// remainder(d, box) for a position difference whose quotient d / box is not
// within rounding of a half-integer (every pair within a cutoff below box / 2:
// its min image is far from +-box / 2): n = rint(d / box) then equals the
// quotient remainder() uses, and d - n box is exactly representable, so the
// single fused multiply-add returns the same bits without a division
__device__ __forceinline__ double min_image(double d, double box, double ibox) {
    return fma(-rint(d * ibox), box, d);
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

// The same scalars from r and its reciprocal ir = 1 / r, with products by ir
// in place of the divisions by r, r^2 and r^3 (the matvec's gather loop,
// which needs ir for the unit vector anyway); equal to ewald_scalars to a
// few rounding errors.
template <typename T>
__device__ __forceinline__ void ewald_scalars_ir(T xi, T r, T ir, T aa, T ab, T* xa, T* ya) {
    const T xir = xi * r;
    const T xir2 = xir * xir;
    const T ir2 = ir * ir;
    const T ir3 = ir2 * ir;
    const T erfcxir = erfc(xir);
    const T expxir2 = xi * T(0.56418958354775628694807945156077) * exp(-xir2);  // xi/sqrt(pi)
    const T er = erfcxir * ir;
    const T a0 = T(2) * (T(-3) + xir2 * T(2)) * expxir2 + er;
    const T b0 = T(2) * (T(1) + xir2 * T(-2)) * expxir2 + er;
    const T a2 = T(4) * ir2 * (T(1) + xir2 * (T(14) + xir2 * (T(-20) + xir2 * T(4)))) * expxir2 +
                 T(2) * ir3 * erfcxir;
    const T b2 = T(-4) * ir2 * (T(3) + xir2 * (T(2) + xir2 * (T(-16) + xir2 * T(4)))) * expxir2 -
                 T(6) * ir3 * erfcxir;
    const T s6 = (aa * aa + ab * ab) * T(1.0 / 6.0);
    *ya = T(0.75) * (a0 + a2 * s6);
    *xa = T(0.75) * (b0 + b2 * s6) + *ya;
}

}  // namespace sdt
