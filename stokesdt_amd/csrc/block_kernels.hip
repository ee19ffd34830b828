// Tall-skinny FP64 block kernels for block Lanczos (packed row-major n x b
// blocks) and vector layout conversion.
#include "block_gpu.hpp"
#include "fft_lds.hpp"

namespace sdt {

// ---------------------------------------------------------------- Gram -----
// Column groups: a packed row-major n x btot block (leading dimension btot)
// is split into groups of bs consecutive columns (the last may be narrower).
// Each group is an independent Krylov block of width <= 16.
//
// Partial G_g = A_g^T B_g over a chunk of rows per workgroup using the FP64
// MFMA (v_mfma_f64_16x16x4_f64): each wave takes 4 rows per instruction;
// lanes hold A^T (16 x 4) and B (4 x 16) fragments, columns >= b_g are
// zero. Consecutive workgroups take the groups of one row chunk, so the rows
// are read from HBM once and re-read from L2. Partials are reduced by a
// second kernel in fixed order (deterministic).
constexpr int kGramRows = 2048;   // rows per workgroup
constexpr int kGramThreads = 256;

typedef double d4 __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(kGramThreads)
gram_partial_kernel(long long n, int btot, int bs, int ngroup, const double* __restrict__ A,
                    const double* __restrict__ B, double* __restrict__ part) {
    __shared__ double s_red[kGramThreads / 64][256];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    // the groups of one row chunk run on one XCD, so the chunk's rows come
    // from HBM into that XCD's L2 once
    const unsigned lid = xcd_logical_id(blockIdx.x, gridDim.x);
    const int g = (int)(lid % ngroup);
    const int chunk = (int)(lid / ngroup);
    const int c0 = g * bs;
    const int bg = min(bs, btot - c0);
    const long long row0 = (long long)chunk * kGramRows;
    const long long rend = min(row0 + kGramRows, n);
    // fragment layout of v_mfma_f64_16x16x4f64 (measured on gfx950): A
    // operand lane l holds A[m = l % 16][k = l / 16]; B operand lane l holds
    // B[k = l / 16][n = l % 16]; result register v of lane l holds
    // C[4*v + l/16][l % 16].
    const int mcol = lane & 15;
    const int krow = lane >> 4;
    d4 acc = {0.0, 0.0, 0.0, 0.0};
    // 32 rows per wave and step: the 16 loads of a step are in flight together
    constexpr int U = 8;
    for (long long r = row0 + wave * 4 * U; r < rend; r += 4 * U * (kGramThreads / 64)) {
        double a[U], bb[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const long long rr = r + 4 * u + krow;
            a[u] = 0.0;
            bb[u] = 0.0;
            if (rr < rend && mcol < bg) {
                a[u] = A[rr * btot + c0 + mcol];
                bb[u] = B[rr * btot + c0 + mcol];
            }
        }
#pragma unroll
        for (int u = 0; u < U; u++)
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u], bb[u], acc, 0, 0, 0);
    }
    // write this wave's 16x16 tile to LDS, then sum waves
#pragma unroll
    for (int v = 0; v < 4; v++) s_red[wave][(4 * v + (lane >> 4)) * 16 + (lane & 15)] = acc[v];
    __syncthreads();
    const int t = threadIdx.x;   // 256 entries of the 16x16 tile
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < kGramThreads / 64; w++) s += s_red[w][t];
    part[(size_t)lid * 256 + t] = s;
}

// G[g][16][16] = sum over chunks of the partial tiles of group g (four
// interleaved partial sums combined in a fixed order: deterministic)
__global__ void gram_reduce_kernel(int nchunk, int ngroup, const double* __restrict__ part,
                                   double* __restrict__ G) {
    const int g = blockIdx.x, t = threadIdx.x;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    const size_t st = (size_t)ngroup * 256;
    const double* p0 = part + (size_t)g * 256 + t;
    int p = 0;
    for (; p + 4 <= nchunk; p += 4) {
        s0 += p0[(size_t)p * st];
        s1 += p0[(size_t)(p + 1) * st];
        s2 += p0[(size_t)(p + 2) * st];
        s3 += p0[(size_t)(p + 3) * st];
    }
    for (; p < nchunk; p++) s0 += p0[(size_t)p * st];
    G[(size_t)g * 256 + t] = (s0 + s1) + (s2 + s3);
}

void gram_groups(long long n, int btot, int bs, const double* A, const double* B, double* d_G,
                 DevBuf<double>& part, hipStream_t s) {
    SDT_REQUIRE(bs >= 1 && bs <= 16 && btot >= 1, SDT_EINVAL, "group width must be in [1, 16]");
    const int ngroup = (btot + bs - 1) / bs;
    const unsigned nchunk = ceil_div(n, kGramRows);
    part.reserve((size_t)nchunk * ngroup * 256);
    hipLaunchKernelGGL(gram_partial_kernel, dim3(nchunk * ngroup), dim3(kGramThreads), 0, s, n,
                       btot, bs, ngroup, A, B, part.get());
    hipLaunchKernelGGL(gram_reduce_kernel, dim3(ngroup), dim3(256), 0, s, (int)nchunk, ngroup,
                       part.get(), d_G);
    SDT_HIP_CHECK(hipGetLastError());
}

void gram(long long n, int b, const double* A, const double* B, double* d_G, DevBuf<double>& part,
          hipStream_t s) {
    gram_groups(n, b, b, A, B, d_G, part, s);
}

// ------------------------------------------------------- small products ----
// C[i][c] = beta C[i][c] + alpha sum_k sum_j A_k[i][c0 + j] S[g][k][j][c - c0]
// for column c of group g (columns c0 .. c0 + b_g - 1). A is nblk blocks of
// n x btot (block k at A + k * blk_stride); S is [ngroup][nblk][bs][bs].
//
// A workgroup owns kTsRows consecutive rows: for each block k it copies the
// rows (one contiguous run of kTsRows btot values) into LDS with coalesced
// loads, then thread (c, h) adds, for its column c and rows h, h + 2, ...,
// the products with column c - c0 of S_gk held in registers. Every value of
// A is read from HBM once and C is written in whole rows. In LDS each group
// is followed by one unused slot (row length btot + ngroup), so the groups a
// wave reads at once fall on distinct banks. C may alias A when nblk = 1
// (in-place C = C S): a tile is in LDS before any of its rows is written.
constexpr int kTsRows = 32;
constexpr int kTsCols = 128;   // widest packed block of the tile kernel
constexpr int kTsLoads = kTsRows * kTsCols / 256;   // tile elements per thread

__global__ void __launch_bounds__(256)
tsmm_tile_kernel(long long n, int btot, int bs, unsigned bmag, int nblk, const double* A,
                 long long blk_stride, const double* __restrict__ S, double alpha, double beta,
                 double* C) {
    extern __shared__ double tile[];   // [kTsRows][ld]
    const int ld = btot + (btot + bs - 1) / bs;
    const long long row0 = (long long)blockIdx.x * kTsRows;
    const int nr = (int)min((long long)kTsRows, n - row0);
    const int c = threadIdx.x % kTsCols, h = threadIdx.x / kTsCols;
    const bool on = c < btot;
    const int g = c / bs, cl = c - g * bs, c0 = g * bs;
    const int bg = min(bs, btot - c0);
    const int gb = c0 + g;             // LDS offset of the group in a row
    double acc[kTsRows / 2];
#pragma unroll
    for (int m = 0; m < kTsRows / 2; m++) acc[m] = 0.0;
    // flat coalesced copies, all loads of a thread in flight together;
    // (row, column) of element e advance by (dq, dr) per step, the group of
    // a column comes from a magic divide
    const int ne = nr * btot;
    double v[kTsLoads];
    auto load = [&](int k) __attribute__((always_inline)) {
        const double* src = A + k * blk_stride + row0 * btot;
#pragma unroll
        for (int u = 0; u < kTsLoads; u++) {
            const int e = threadIdx.x + u * 256;
            v[u] = e < ne ? src[e] : 0.0;
        }
    };
    for (int k = 0; k < nblk; k++) {
        if (k > 0) __syncthreads();
        load(k);
        {
            int r = threadIdx.x / btot, cc = threadIdx.x - r * btot;
            const int dq = 256 / btot, dr = 256 - dq * btot;
#pragma unroll
            for (int u = 0; u < kTsLoads; u++) {
                if (threadIdx.x + u * 256 < ne) tile[r * ld + cc + (int)fdiv((unsigned)cc, bmag)] = v[u];
                r += dq;
                cc += dr;
                if (cc >= btot) {
                    cc -= btot;
                    r++;
                }
            }
        }
        __syncthreads();
        if (on) {
            double sv[16];
            const double* sk = S + ((size_t)(g * nblk + k) * bs) * bs + cl;
#pragma unroll
            for (int j = 0; j < 16; j++) sv[j] = (j < bg) ? sk[(size_t)j * bs] : 0.0;
#pragma unroll
            for (int m = 0; m < kTsRows / 2; m++) {
                const int r = h + 2 * m;
                if (r < nr) {
                    const double* t = tile + r * ld + gb;
                    double a = 0.0;
#pragma unroll
                    for (int j = 0; j < 16; j++)
                        if (j < bg) a += t[j] * sv[j];
                    acc[m] += a;
                }
            }
        }
    }
    if (!on) return;
    double* cc = C + (row0 + h) * btot + c;
    if (beta != 0.0) {
        // every old value is loaded before the first store
        double old[kTsRows / 2];
#pragma unroll
        for (int m = 0; m < kTsRows / 2; m++)
            old[m] = (h + 2 * m < nr) ? cc[(size_t)(2 * m) * btot] : 0.0;
#pragma unroll
        for (int m = 0; m < kTsRows / 2; m++) acc[m] = beta * old[m] + alpha * acc[m];
    } else {
#pragma unroll
        for (int m = 0; m < kTsRows / 2; m++) acc[m] = alpha * acc[m];
    }
#pragma unroll
    for (int m = 0; m < kTsRows / 2; m++)
        if (h + 2 * m < nr) cc[(size_t)(2 * m) * btot] = acc[m];
}

// wider blocks: one thread per element
__global__ void __launch_bounds__(256)
tsmm_kernel(long long n, int btot, int bs, int nblk, const double* __restrict__ A,
            long long blk_stride, const double* __restrict__ S, double alpha, double beta,
            double* __restrict__ C) {
    const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= n * btot) return;
    const long long i = gid / btot;
    const int c = (int)(gid - i * btot);
    const int g = c / bs, cl = c - g * bs, c0 = g * bs;
    const int bg = min(bs, btot - c0);
    double acc = 0.0;
    for (int k = 0; k < nblk; k++) {
        const double* a = A + k * blk_stride + i * btot + c0;
        const double* sk = S + ((size_t)(g * nblk + k) * bs) * bs + cl;
        for (int j = 0; j < bg; j++) acc += a[j] * sk[(size_t)j * bs];
    }
    double* cp = C + i * btot + c;
    *cp = (beta == 0.0 ? 0.0 : beta * *cp) + alpha * acc;
}

static void launch_tsmm(long long n, int btot, int bs, int nblk, const double* A,
                        long long blk_stride, const double* d_S, double alpha, double beta,
                        double* C, hipStream_t s) {
    if (btot <= kTsCols) {
        const size_t lds = sizeof(double) * kTsRows * (btot + (btot + bs - 1) / bs);
        hipLaunchKernelGGL(tsmm_tile_kernel, dim3(ceil_div(n, kTsRows)), dim3(256), lds, s, n,
                           btot, bs, fft_magic((unsigned)bs), nblk, A, blk_stride, d_S, alpha,
                           beta, C);
    } else {
        // rows are read and written by different threads: no aliasing
        SDT_REQUIRE(C != A, SDT_EINVAL, "tsmm: in-place product needs at most 128 columns");
        hipLaunchKernelGGL(tsmm_kernel, dim3(ceil_div(n * btot, 256)), dim3(256), 0, s, n, btot,
                           bs, nblk, A, blk_stride, d_S, alpha, beta, C);
    }
    SDT_HIP_CHECK(hipGetLastError());
}

void tsmm_groups(long long n, int btot, int bs, int nblk, const double* A, long long blk_stride,
                 const double* d_S, double alpha, double beta, double* C, hipStream_t s) {
    SDT_REQUIRE(bs >= 1 && bs <= 16, SDT_EINVAL, "group width must be in [1, 16]");
    SDT_REQUIRE(n * btot < (1ll << 40), SDT_EINVAL, "tsmm: block too large");
    SDT_REQUIRE(nblk == 1 || C != A, SDT_EINVAL, "tsmm: in-place product needs one block");
    launch_tsmm(n, btot, bs, nblk, A, blk_stride, d_S, alpha, beta, C, s);
}

void tsmm_groups1(long long n, int btot, int bs, const double* A, const double* d_S, double alpha,
                  double beta, double* C, hipStream_t s) {
    SDT_REQUIRE(bs >= 1 && bs <= 16, SDT_EINVAL, "group width must be in [1, 16]");
    launch_tsmm(n, btot, bs, 1, A, 0LL, d_S, alpha, beta, C, s);
}

void tsmm(long long n, int b, int nblk, const double* A, long long blk_stride, const double* d_S,
          double alpha, double beta, double* C, hipStream_t s) {
    tsmm_groups(n, b, b, nblk, A, blk_stride, d_S, alpha, beta, C, s);
}

// ------------------------------------------------------- norms / copies ----
__global__ void sumsq_partial_kernel(long long n, const double* __restrict__ x, double* __restrict__ part) {
    __shared__ double s[256];
    double acc = 0.0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        acc += x[i] * x[i];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

double sumsq(long long n, const double* x, DevBuf<double>& part, hipStream_t s) {
    const int nb = 512;
    part.reserve(nb);
    hipLaunchKernelGGL(sumsq_partial_kernel, dim3(nb), dim3(256), 0, s, n, x, part.get());
    std::vector<double> h(nb);
    SDT_HIP_CHECK(hipMemcpyAsync(h.data(), part.get(), sizeof(double) * nb, hipMemcpyDeviceToHost, s));
    SDT_HIP_CHECK(hipStreamSynchronize(s));
    double t = 0.0;
    for (double v : h) t += v;
    return t;
}

// packed[(3 is + c) b + r] = f[r ldf + 3 perm[is] + c]
__global__ void pack_kernel(int npos, int b, const int* __restrict__ perm, int ldf,
                            const double* __restrict__ f, double* __restrict__ X) {
    const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= (long long)npos * 3 * b) return;
    const int r = (int)(gid % b);
    const long long row = gid / b;
    const int is = (int)(row / 3), c = (int)(row % 3);
    const int i = perm ? perm[is] : is;
    X[gid] = f[(size_t)r * ldf + 3 * (size_t)i + c];
}

// v[r ldv + 3 perm[is] + c] = packed[(3 is + c) b + r]
__global__ void unpack_kernel(int npos, int b, const int* __restrict__ perm, int ldv,
                              const double* __restrict__ X, double* __restrict__ v) {
    const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= (long long)npos * 3 * b) return;
    const int r = (int)(gid % b);
    const long long row = gid / b;
    const int is = (int)(row / 3), c = (int)(row % 3);
    const int i = perm ? perm[is] : is;
    v[(size_t)r * ldv + 3 * (size_t)i + c] = X[gid];
}

// dst[row][k] = src[row][cols[k]] (to_packed) or dst[row][cols[k]] =
// src[row][k] (back): moves a subset of columns between a packed block of
// leading dimension ld and a compact one of nc columns
__global__ void column_subset_kernel(long long nrow, int ld, int nc, const int* __restrict__ cols,
                                     const double* __restrict__ src, double* __restrict__ dst,
                                     int to_compact) {
    const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= nrow * nc) return;
    const long long row = gid / nc;
    const int k = (int)(gid - row * nc);
    if (to_compact) dst[gid] = src[row * ld + cols[k]];
    else dst[row * ld + cols[k]] = src[gid];
}

void column_subset(long long nrow, int ld, int nc, const int* d_cols, const double* src,
                   double* dst, bool to_compact, hipStream_t s) {
    const long long n = nrow * nc;
    hipLaunchKernelGGL(column_subset_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, s, nrow, ld, nc,
                       d_cols, src, dst, to_compact ? 1 : 0);
    SDT_HIP_CHECK(hipGetLastError());
}

void pack(int npos, int b, const int* perm, int ldf, const double* f, double* X, hipStream_t s) {
    const long long n = (long long)npos * 3 * b;
    hipLaunchKernelGGL(pack_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, s, npos, b, perm, ldf, f, X);
    SDT_HIP_CHECK(hipGetLastError());
}

void unpack(int npos, int b, const int* perm, int ldv, const double* X, double* v, hipStream_t s) {
    const long long n = (long long)npos * 3 * b;
    hipLaunchKernelGGL(unpack_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, s, npos, b, perm, ldv, X, v);
    SDT_HIP_CHECK(hipGetLastError());
}

// ------------------------------------------- Lanczos coefficient steps -----
// One workgroup per column group; the Gram tiles are 16 x 16 row-major, the
// coefficient blocks S[g] bs x bs row-major (zero for a masked group and
// beyond a narrower last group's width).

// mode 0: S[g] = H_g = G[g] (and the raw tile into the record);
// mode 1: S[g] = B_g^T from the previous iteration's record
__global__ void __launch_bounds__(256)
lz_coef_kernel(int bs, int btot, const int* __restrict__ mask, const double* __restrict__ G,
               const double* __restrict__ prev, double* __restrict__ rec, double* __restrict__ S,
               int mode) {
    const int g = blockIdx.x, t = threadIdx.x;
    const int i = t >> 4, j = t & 15;
    const int b = min(bs, btot - g * bs);
    double v;
    if (mode == 0) {
        v = G[(size_t)g * 256 + t];
        rec[(size_t)g * kLzRec + t] = v;
    } else {
        v = prev[(size_t)g * kLzRec + 256 + j * 16 + i];
    }
    if (i < bs && j < bs)
        S[((size_t)g * bs + i) * bs + j] = (mask[g] && i < b && j < b) ? v : 0.0;
}

// One CholQR pass over the Gram tiles G[g] = W_g^T W_g: R = chol(sym G)
// (upper; host_linalg.cpp chol_upper), S[g] = R^{-1} (inv_upper). Pass 0
// keeps R in R1[g]; pass 1 writes B_g = R R1[g] (mul_upper) into the record.
// A masked group, or one whose pivot is not positive (numerically rank
// deficient: the pass zeroes its columns), gets S = 0, B = 0 and ok = 0.
// Lane c (< 16) holds column c; the width is padded to 16 with the identity.
__global__ void __launch_bounds__(64)
lz_chol_kernel(int bs, int btot, const int* __restrict__ mask, const double* __restrict__ G,
               double* __restrict__ R1, double* __restrict__ rec, double* __restrict__ S,
               int pass) {
    const int g = blockIdx.x, c = threadIdx.x & 15;
    const bool lead = threadIdx.x < 16;
    const int b = min(bs, btot - g * bs);
    double* rg = rec + (size_t)g * kLzRec;
    bool ok = mask[g] != 0;
    if (pass == 1) ok = ok && rg[512] != 0.0;
    const double* Gg = G + (size_t)g * 256;
    double a[16];
#pragma unroll
    for (int r = 0; r < 16; r++)
        a[r] = (r < b && c < b) ? 0.5 * (Gg[r * 16 + c] + Gg[c * 16 + r]) : (r == c ? 1.0 : 0.0);
    double rc[16];   // column c of R
#pragma unroll
    for (int j = 0; j < 16; j++) {
        const double p = __shfl(a[j], j);
        if (!(p > 0.0)) ok = false;
        const double rjj = p > 0.0 ? sqrt(p) : 1.0;
        const double rjc = c == j ? rjj : (c > j ? a[j] / rjj : 0.0);
        rc[j] = rjc;
#pragma unroll
        for (int i = j + 1; i < 16; i++) a[i] -= __shfl(rjc, i) * rjc;
    }
    // column c of R^{-1} by back substitution
    double x[16];
#pragma unroll
    for (int r = 15; r >= 0; r--) {
        double t = r == c ? 1.0 : 0.0;
#pragma unroll
        for (int k = r + 1; k < 16; k++) t -= __shfl(rc[r], k) * x[k];
        x[r] = r <= c ? t / __shfl(rc[r], r) : 0.0;
    }
    if (lead) {
        if (c < bs)
#pragma unroll
            for (int i = 0; i < 16; i++)
                if (i < bs) S[((size_t)g * bs + i) * bs + c] = (ok && i < b && c < b) ? x[i] : 0.0;
        if (pass == 0) {
#pragma unroll
            for (int r = 0; r < 16; r++) R1[(size_t)g * 256 + r * 16 + c] = rc[r];
        }
    }
    if (pass == 1) {
        double r1[16];
#pragma unroll
        for (int k = 0; k < 16; k++) r1[k] = R1[(size_t)g * 256 + k * 16 + c];
#pragma unroll
        for (int i = 0; i < 16; i++) {
            double t = 0.0;
#pragma unroll
            for (int k = i; k < 16; k++) t += __shfl(rc[i], k) * r1[k];
            if (lead) rg[256 + i * 16 + c] = (ok && c >= i) ? t : 0.0;
        }
    }
    if (threadIdx.x == 0) rg[512] = ok ? 1.0 : 0.0;
}

void lz_coef(int ng, int bs, int btot, const int* d_mask, const double* d_G, const double* d_prev,
             double* d_rec, double* d_S, int mode, hipStream_t s) {
    lz_coef_kernel<<<ng, 256, 0, s>>>(bs, btot, d_mask, d_G, d_prev, d_rec, d_S, mode);
    SDT_HIP_CHECK(hipGetLastError());
}

void lz_chol(int ng, int bs, int btot, const int* d_mask, const double* d_G, double* d_R1,
             double* d_rec, double* d_S, int pass, hipStream_t s) {
    lz_chol_kernel<<<ng, 64, 0, s>>>(bs, btot, d_mask, d_G, d_R1, d_rec, d_S, pass);
    SDT_HIP_CHECK(hipGetLastError());
}

__global__ void axpby_kernel(long long n, double a, const double* __restrict__ x, double b,
                             double* __restrict__ y) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = a * x[i] + (b == 0.0 ? 0.0 : b * y[i]);
}

void axpby(long long n, double a, const double* x, double b, double* y, hipStream_t s) {
    hipLaunchKernelGGL(axpby_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, s, n, a, x, b, y);
    SDT_HIP_CHECK(hipGetLastError());
}

__global__ void fill_double_kernel(double* p, long long n, double v) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

void fill(long long n, double v, double* p, hipStream_t s) {
    hipLaunchKernelGGL(fill_double_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, s, p, n, v);
    SDT_HIP_CHECK(hipGetLastError());
}

}  // namespace sdt
