// Device-side radix-2/3/4/5/8 Stockham FFT on pencils held in LDS.
//
// Self-sorting decimation-in-frequency form: a stage of radix r on a
// sub-sequence length `len` with stride `s` (m = len / r) reads
// src[q + s*(p + j*m)], j < r, applies an r-point DFT and the twiddle
// w_len^(p*t), and writes dst[q + s*(r*p + t)]. After all stages the result
// is in natural order. Twiddles come from one table of K roots of unity
// w_K^j (w_len^(p t) = w_K^((K/len) p t), and (K/len) p t < K).
//
// LDS layout: element e of pencil k sits at k * PS + e + (e >> 4), i.e. one
// pad slot every 16 elements and an odd pencil stride PS, so that strided
// butterfly accesses and the transposed pencil loads of the y/z passes hit
// distinct banks. Index splits use multiply-high reciprocals precomputed on
// the host instead of integer division.
#pragma once
#include "sdt_internal.hpp"

namespace sdt {

struct FFTArgs {
    int K;
    int nst;
    unsigned K_m, Kh_m;            // reciprocals of K and K / 2 + 1
    int radix[16];
    unsigned per[16], per_m[16];   // butterflies per pencil (K / r) and reciprocal
    unsigned str[16], str_m[16];   // sub-sequence stride s = K / len and reciprocal
};

// floor(n / d) = umulhi(n, floor(2^32 / d) + 1) for d >= 2 and n d < 2^32
// (the error term n (d - 2^32 mod d) / 2^32 / d stays below 1 / d); the
// reciprocal of 1 is encoded as 0.
inline unsigned fft_magic(unsigned d) {
    return d == 1 ? 0u : (unsigned)((1ull << 32) / d + 1ull);
}

__device__ __forceinline__ unsigned fdiv(unsigned n, unsigned m) {
    return m ? __umulhi(n, m) : n;
}

__host__ __device__ inline int pencil_stride(int K) { return (K + (K >> 4) + 1) | 1; }

__device__ __forceinline__ int lpad(int e) { return e + (e >> 4); }

template <typename T>
__device__ __forceinline__ cplx<T> cmul(cplx<T> a, cplx<T> b) {
    return cplx<T>{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T>
__device__ __forceinline__ cplx<T> cadd(cplx<T> a, cplx<T> b) {
    return cplx<T>{a.x + b.x, a.y + b.y};
}
template <typename T>
__device__ __forceinline__ cplx<T> csub(cplx<T> a, cplx<T> b) {
    return cplx<T>{a.x - b.x, a.y - b.y};
}
// rot(a) = a * (sign < 0 ? -i : +i)
template <typename T, int SIGN>
__device__ __forceinline__ cplx<T> rot90(cplx<T> a) {
    return SIGN < 0 ? cplx<T>{a.y, -a.x} : cplx<T>{-a.y, a.x};
}

// r-point DFT in registers, kernel e^{SIGN * 2 pi i j t / r}
template <typename T, int R, int SIGN>
__device__ __forceinline__ void small_dft(cplx<T>* a) {
    if constexpr (R == 2) {
        cplx<T> u = a[0], v = a[1];
        a[0] = cadd(u, v);
        a[1] = csub(u, v);
    } else if constexpr (R == 4) {
        cplx<T> s02 = cadd(a[0], a[2]), d02 = csub(a[0], a[2]);
        cplx<T> s13 = cadd(a[1], a[3]), d13 = rot90<T, SIGN>(csub(a[1], a[3]));
        a[0] = cadd(s02, s13);
        a[1] = cadd(d02, d13);
        a[2] = csub(s02, s13);
        a[3] = csub(d02, d13);
    } else if constexpr (R == 3) {
        const T c1 = T(-0.5);
        const T s1 = T(SIGN) * T(0.86602540378443864676372317075294);
        cplx<T> t1 = cadd(a[1], a[2]);
        cplx<T> t2 = csub(a[1], a[2]);
        cplx<T> m = cplx<T>{a[0].x + c1 * t1.x, a[0].y + c1 * t1.y};
        cplx<T> j = cplx<T>{-s1 * t2.y, s1 * t2.x};
        a[0] = cadd(a[0], t1);
        a[1] = cadd(m, j);
        a[2] = csub(m, j);
    } else if constexpr (R == 8) {
        // two 4-point DFTs (even / odd inputs) and the w_8^k combination
        cplx<T> e[4] = {a[0], a[2], a[4], a[6]};
        cplx<T> o[4] = {a[1], a[3], a[5], a[7]};
        small_dft<T, 4, SIGN>(e);
        small_dft<T, 4, SIGN>(o);
        const T h = T(0.70710678118654752440084436210485);
        const T sg = T(SIGN);
        const cplx<T> t1{h * (o[1].x - sg * o[1].y), h * (o[1].y + sg * o[1].x)};
        const cplx<T> t2 = rot90<T, SIGN>(o[2]);
        const cplx<T> t3{h * (-o[3].x - sg * o[3].y), h * (-o[3].y + sg * o[3].x)};
        a[0] = cadd(e[0], o[0]);
        a[4] = csub(e[0], o[0]);
        a[1] = cadd(e[1], t1);
        a[5] = csub(e[1], t1);
        a[2] = cadd(e[2], t2);
        a[6] = csub(e[2], t2);
        a[3] = cadd(e[3], t3);
        a[7] = csub(e[3], t3);
    } else if constexpr (R == 16) {
        // n = 4 n1 + n2: 4-point DFTs over n1, twiddles w_16^(n2 k1), then
        // 4-point DFTs over n2; X[k1 + 4 k2]
        cplx<T> b[4][4];
#pragma unroll
        for (int n2 = 0; n2 < 4; n2++) {
            b[n2][0] = a[n2]; b[n2][1] = a[4 + n2]; b[n2][2] = a[8 + n2]; b[n2][3] = a[12 + n2];
            small_dft<T, 4, SIGN>(b[n2]);
        }
        const T c1 = T(0.92387953251128675612818318939679);   // cos(pi/8)
        const T s1 = T(0.38268343236508977172845998403040);   // sin(pi/8)
        const T h = T(0.70710678118654752440084436210485);
        const T sg = T(SIGN);
        // w_16^e = cos(2 pi e / 16) + SIGN i sin(2 pi e / 16), e = n2 k1
        auto tw16 = [&](cplx<T> v, int e) -> cplx<T> {
            T cr, ci;
            switch (e) {
                case 1: cr = c1; ci = s1; break;
                case 2: cr = h; ci = h; break;
                case 3: cr = s1; ci = c1; break;
                case 4: cr = T(0); ci = T(1); break;
                case 6: cr = -h; ci = h; break;
                case 9: cr = -c1; ci = -s1; break;
                default: return v;
            }
            ci *= sg;
            return cplx<T>{v.x * cr - v.y * ci, v.x * ci + v.y * cr};
        };
#pragma unroll
        for (int n2 = 1; n2 < 4; n2++)
#pragma unroll
            for (int k1 = 1; k1 < 4; k1++) b[n2][k1] = tw16(b[n2][k1], n2 * k1);
#pragma unroll
        for (int k1 = 0; k1 < 4; k1++) {
            cplx<T> c[4] = {b[0][k1], b[1][k1], b[2][k1], b[3][k1]};
            small_dft<T, 4, SIGN>(c);
#pragma unroll
            for (int k2 = 0; k2 < 4; k2++) a[k1 + 4 * k2] = c[k2];
        }
    } else if constexpr (R == 5) {
        const T c1 = T(0.30901699437494742410229341718282);    // cos(2pi/5)
        const T c2 = T(-0.80901699437494742410229341718282);   // cos(4pi/5)
        const T s1 = T(SIGN) * T(0.95105651629515357211643933337938);  // sin(2pi/5)
        const T s2 = T(SIGN) * T(0.58778525229247312916870595463907);  // sin(4pi/5)
        cplx<T> t1 = cadd(a[1], a[4]), t2 = cadd(a[2], a[3]);
        cplx<T> d1 = csub(a[1], a[4]), d2 = csub(a[2], a[3]);
        cplx<T> m1 = cplx<T>{a[0].x + c1 * t1.x + c2 * t2.x, a[0].y + c1 * t1.y + c2 * t2.y};
        cplx<T> m2 = cplx<T>{a[0].x + c2 * t1.x + c1 * t2.x, a[0].y + c2 * t1.y + c1 * t2.y};
        cplx<T> n1 = cplx<T>{-(s1 * d1.y + s2 * d2.y), s1 * d1.x + s2 * d2.x};
        cplx<T> n2 = cplx<T>{-(s2 * d1.y - s1 * d2.y), s2 * d1.x - s1 * d2.x};
        a[0] = cplx<T>{a[0].x + t1.x + t2.x, a[0].y + t1.y + t2.y};
        a[1] = cadd(m1, n1);
        a[4] = csub(m1, n1);
        a[2] = cadd(m2, n2);
        a[3] = csub(m2, n2);
    }
}

template <typename T, int R, int SIGN>
__device__ __forceinline__ void fft_stage(const cplx<T>* __restrict__ src,
                                          cplx<T>* __restrict__ dst, int npen, int PS,
                                          const FFTArgs& fa, int st,
                                          const cplx<T>* __restrict__ tw) {
    const unsigned per = fa.per[st], per_m = fa.per_m[st];
    const unsigned s = fa.str[st], s_m = fa.str_m[st];
    const unsigned m = per / s;            // = len / R
    const unsigned total = (unsigned)npen * per;
    for (unsigned item = threadIdx.x; item < total; item += blockDim.x) {
        const unsigned pen = fdiv(item, per_m);
        const unsigned j = item - pen * per;
        const unsigned p = fdiv(j, s_m);
        const unsigned q = j - p * s;
        const cplx<T>* sp = src + pen * PS;
        cplx<T>* dp = dst + pen * PS;
        cplx<T> a[R];
#pragma unroll
        for (int t = 0; t < R; t++) a[t] = sp[lpad(q + s * (p + t * m))];
        small_dft<T, R, SIGN>(a);
        const unsigned ob = q + s * R * p;
        dp[lpad(ob)] = a[0];
#pragma unroll
        for (int t = 1; t < R; t++) {
            cplx<T> w = tw[s * p * t];    // w_len^(p t) = w_K^(s p t), s p t < K
            if (SIGN > 0) w.y = -w.y;
            dp[lpad(ob + s * t)] = cmul(a[t], w);
        }
    }
}

// In-place form of the same stage: every thread first computes its (at most
// MAXB) butterflies into registers, the workgroup synchronises, then all
// results are stored. One LDS buffer instead of two. Requires
// npen * K / R <= MAXB * blockDim.x.
template <typename T, int R, int SIGN, int MAXB>
__device__ __forceinline__ void fft_stage_inplace(cplx<T>* __restrict__ buf, int npen, int PS,
                                                  const FFTArgs& fa, int st,
                                                  const cplx<T>* __restrict__ tw) {
    const unsigned per = fa.per[st], per_m = fa.per_m[st];
    const unsigned s = fa.str[st], s_m = fa.str_m[st];
    const unsigned m = per / s;
    const unsigned total = (unsigned)npen * per;
    cplx<T> out[MAXB][R];
    unsigned dst[MAXB];
#pragma unroll
    for (int u = 0; u < MAXB; u++) {
        const unsigned item = threadIdx.x + u * blockDim.x;
        dst[u] = 0xffffffffu;
        if (item < total) {
            const unsigned pen = fdiv(item, per_m);
            const unsigned j = item - pen * per;
            const unsigned p = fdiv(j, s_m);
            const unsigned q = j - p * s;
            const cplx<T>* sp = buf + pen * PS;
#pragma unroll
            for (int t = 0; t < R; t++) out[u][t] = sp[lpad(q + s * (p + t * m))];
            small_dft<T, R, SIGN>(out[u]);
#pragma unroll
            for (int t = 1; t < R; t++) {
                cplx<T> w = tw[s * p * t];
                if (SIGN > 0) w.y = -w.y;
                out[u][t] = cmul(out[u][t], w);
            }
            dst[u] = pen * PS + q + s * R * p;   // unpadded offset of output 0
        }
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < MAXB; u++) {
        if (dst[u] != 0xffffffffu) {
            const unsigned pen = fdiv(threadIdx.x + u * blockDim.x, per_m);
            const unsigned ob = dst[u] - pen * PS;
            cplx<T>* dp = buf + pen * PS;
#pragma unroll
            for (int t = 0; t < R; t++) dp[lpad(ob + s * t)] = out[u][t];
        }
    }
    __syncthreads();
}

// In-place transform of npen pencils; MAXI2 bounds the radix-2 butterflies
// per thread (npen * K <= 2 * MAXI2 * blockDim.x), other radices scale.
template <typename T, int SIGN, int MAXI2>
__device__ __forceinline__ void fft_pencils_inplace(cplx<T>* buf, int npen, int PS, const FFTArgs& fa,
                                    const cplx<T>* tw) {
    for (int st = 0; st < fa.nst; st++) {
        const int r = fa.radix[st];
        if (r == 8) fft_stage_inplace<T, 8, SIGN, (2 * MAXI2 + 7) / 8>(buf, npen, PS, fa, st, tw);
        else if (r == 4) fft_stage_inplace<T, 4, SIGN, (2 * MAXI2 + 3) / 4>(buf, npen, PS, fa, st, tw);
        else if (r == 2) fft_stage_inplace<T, 2, SIGN, MAXI2>(buf, npen, PS, fa, st, tw);
        else if (r == 3) fft_stage_inplace<T, 3, SIGN, (2 * MAXI2 + 2) / 3>(buf, npen, PS, fa, st, tw);
        else fft_stage_inplace<T, 5, SIGN, (2 * MAXI2 + 4) / 5>(buf, npen, PS, fa, st, tw);
    }
}

// Two-stage in-place transform for K = R1 R2 (R1, R2 powers of two <= 16):
// stage 1 (radix R1, stride 1): butterfly j < R2 reads elements j + R2 u,
// applies w_K^(j u) and writes R1 j + u; stage 2 (radix R2, stride R1, no
// twiddles): butterfly j < R1 reads and writes elements j + R1 u. Index
// arithmetic is shifts and masks; stage 2 is in place per thread. Requires
// npen * R2 <= MAXB * blockDim.x.
template <typename T, int R1, int R2, int SIGN, int MAXB = 1>
__device__ __forceinline__ void fft2_pencils(cplx<T>* buf, int npen, int PS,
                                             const cplx<T>* __restrict__ tw) {
    static_assert((R1 & (R1 - 1)) == 0 && (R2 & (R2 - 1)) == 0, "powers of two");
    {
        cplx<T> a[MAXB][R1];
#pragma unroll
        for (int v = 0; v < MAXB; v++) {
            const int it = threadIdx.x + v * blockDim.x;
            if (it < npen * R2) {
                const int pen = it / R2, j = it % R2;
                const cplx<T>* base = buf + pen * PS;
#pragma unroll
                for (int u = 0; u < R1; u++) a[v][u] = base[lpad(j + R2 * u)];
                small_dft<T, R1, SIGN>(a[v]);
#pragma unroll
                for (int u = 1; u < R1; u++) {
                    cplx<T> w = tw[j * u];
                    if (SIGN > 0) w.y = -w.y;
                    a[v][u] = cmul(a[v][u], w);
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int v = 0; v < MAXB; v++) {
            const int it = threadIdx.x + v * blockDim.x;
            if (it < npen * R2) {
                const int pen = it / R2, j = it % R2;
                cplx<T>* base = buf + pen * PS;
#pragma unroll
                for (int u = 0; u < R1; u++) base[lpad(R1 * j + u)] = a[v][u];
            }
        }
        __syncthreads();
    }
    for (int it = threadIdx.x; it < npen * R1; it += blockDim.x) {
        const int pen = it / R1, j = it % R1;
        cplx<T>* base = buf + pen * PS;
        cplx<T> a[R2];
#pragma unroll
        for (int u = 0; u < R2; u++) a[u] = base[lpad(j + R1 * u)];
        small_dft<T, R2, SIGN>(a);
#pragma unroll
        for (int u = 0; u < R2; u++) base[lpad(j + R1 * u)] = a[u];
    }
    __syncthreads();
}

// Runs all stages; returns the buffer holding the result (a or b).// Runs all stages; returns the buffer holding the result (a or b).
template <typename T, int SIGN>
__device__ __forceinline__ cplx<T>* fft_pencils(cplx<T>* a, cplx<T>* b, int npen, int PS, const FFTArgs& fa,
                                const cplx<T>* tw) {
    for (int st = 0; st < fa.nst; st++) {
        const int r = fa.radix[st];
        if (r == 8) fft_stage<T, 8, SIGN>(a, b, npen, PS, fa, st, tw);
        else if (r == 4) fft_stage<T, 4, SIGN>(a, b, npen, PS, fa, st, tw);
        else if (r == 2) fft_stage<T, 2, SIGN>(a, b, npen, PS, fa, st, tw);
        else if (r == 3) fft_stage<T, 3, SIGN>(a, b, npen, PS, fa, st, tw);
        else fft_stage<T, 5, SIGN>(a, b, npen, PS, fa, st, tw);
        __syncthreads();
        cplx<T>* t = a;
        a = b;
        b = t;
    }
    return a;
}

}  // namespace sdt
