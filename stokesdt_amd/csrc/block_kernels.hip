// Tall-skinny FP64 block kernels for block Lanczos (packed row-major n x b
// blocks) and vector layout conversion.
#include "block_gpu.hpp"

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
    const int g = blockIdx.x % ngroup;
    const int chunk = blockIdx.x / ngroup;
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
    for (long long r = row0 + wave * 4; r < rend; r += 4 * (kGramThreads / 64)) {
        const long long rr = r + krow;
        double a = 0.0, bb = 0.0;
        if (rr < rend && mcol < bg) {
            a = A[rr * btot + c0 + mcol];
            bb = B[rr * btot + c0 + mcol];
        }
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bb, acc, 0, 0, 0);
    }
    // write this wave's 16x16 tile to LDS, then sum waves
#pragma unroll
    for (int v = 0; v < 4; v++) s_red[wave][(4 * v + (lane >> 4)) * 16 + (lane & 15)] = acc[v];
    __syncthreads();
    const int t = threadIdx.x;   // 256 entries of the 16x16 tile
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < kGramThreads / 64; w++) s += s_red[w][t];
    part[(size_t)blockIdx.x * 256 + t] = s;
}

// G[g][16][16] = sum over chunks of the partial tiles of group g
__global__ void gram_reduce_kernel(int nchunk, int ngroup, const double* __restrict__ part,
                                   double* __restrict__ G) {
    const int g = blockIdx.x, t = threadIdx.x;
    double s = 0.0;
    for (int p = 0; p < nchunk; p++) s += part[((size_t)p * ngroup + g) * 256 + t];
    G[(size_t)g * 256 + t] = s;
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
// n x btot (block k at A + k * blk_stride); S is [ngroup][nblk][bs][bs] on
// the device, staged in LDS.
__global__ void __launch_bounds__(256)
tsmm_kernel(long long n, int btot, int bs, int nblk, const double* __restrict__ A,
            long long blk_stride, const double* __restrict__ S, double alpha, double beta,
            double* __restrict__ C) {
    extern __shared__ double s_S[];
    const int ngroup = (btot + bs - 1) / bs;
    const int ns = ngroup * nblk * bs * bs;
    for (int t = threadIdx.x; t < ns; t += blockDim.x) s_S[t] = S[t];
    __syncthreads();
    const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= n * btot) return;
    const long long i = gid / btot;
    const int c = (int)(gid - i * btot);
    const int g = c / bs, cl = c - g * bs, c0 = g * bs;
    const int bg = min(bs, btot - c0);
    double acc = 0.0;
    for (int k = 0; k < nblk; k++) {
        const double* a = A + k * blk_stride + i * btot + c0;
        const double* sk = s_S + ((size_t)(g * nblk + k) * bs) * bs + cl;
        for (int j = 0; j < bg; j++) acc += a[j] * sk[(size_t)j * bs];
    }
    double* cp = C + i * btot + c;
    *cp = (beta == 0.0 ? 0.0 : beta * *cp) + alpha * acc;
}

void tsmm_groups(long long n, int btot, int bs, int nblk, const double* A, long long blk_stride,
                 const double* d_S, double alpha, double beta, double* C, hipStream_t s) {
    const int ngroup = (btot + bs - 1) / bs;
    const size_t lds = sizeof(double) * (size_t)ngroup * nblk * bs * bs;
    SDT_REQUIRE(lds <= 64 * 1024, SDT_ECAP, "tsmm: coefficient block too large for LDS");
    SDT_REQUIRE(n * btot < (1ll << 40), SDT_EINVAL, "tsmm: block too large");
    hipLaunchKernelGGL(tsmm_kernel, dim3(ceil_div(n * btot, 256)), dim3(256), lds, s, n, btot, bs,
                       nblk, A, blk_stride, d_S, alpha, beta, C);
    SDT_HIP_CHECK(hipGetLastError());
}

// One coefficient block per group, one thread per (row, group): the thread
// reads the row's b_g values of A, then writes the b_g results, so C may
// alias A (in-place C = C S).
__global__ void __launch_bounds__(256)
tsmm_rowgroup_kernel(long long n, int btot, int bs, const double* A, const double* __restrict__ S,
                     double alpha, double beta, double* C) {
    extern __shared__ double s_S[];
    const int ngroup = (btot + bs - 1) / bs;
    for (int t = threadIdx.x; t < ngroup * bs * bs; t += blockDim.x) s_S[t] = S[t];
    __syncthreads();
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * ngroup) return;
    const long long i = t / ngroup;
    const int g = (int)(t - i * ngroup);
    const int c0 = g * bs, bg = min(bs, btot - c0);
    const double* sg = s_S + (size_t)g * bs * bs;
    double a[16], out[16];
#pragma unroll
    for (int j = 0; j < 16; j++) a[j] = (j < bg) ? A[i * btot + c0 + j] : 0.0;
#pragma unroll
    for (int c = 0; c < 16; c++) {
        double acc = 0.0;
#pragma unroll
        for (int j = 0; j < 16; j++)
            if (j < bg && c < bg) acc += a[j] * sg[j * bs + c];
        out[c] = acc;
    }
#pragma unroll
    for (int c = 0; c < 16; c++) {
        if (c < bg) {
            double* cp = C + i * btot + c0 + c;
            *cp = (beta == 0.0 ? 0.0 : beta * *cp) + alpha * out[c];
        }
    }
}

void tsmm_groups1(long long n, int btot, int bs, const double* A, const double* d_S, double alpha,
                  double beta, double* C, hipStream_t s) {
    SDT_REQUIRE(bs >= 1 && bs <= 16, SDT_EINVAL, "group width must be in [1, 16]");
    const int ngroup = (btot + bs - 1) / bs;
    const size_t lds = sizeof(double) * (size_t)ngroup * bs * bs;
    SDT_REQUIRE(lds <= 64 * 1024, SDT_ECAP, "tsmm: coefficient block too large for LDS");
    hipLaunchKernelGGL(tsmm_rowgroup_kernel, dim3(ceil_div(n * ngroup, 256)), dim3(256), lds, s, n,
                       btot, bs, A, d_S, alpha, beta, C);
    SDT_HIP_CHECK(hipGetLastError());
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
