// Integer scan and deterministic bucket sort (particles by cell or tile).
#include "sort_gpu.hpp"

namespace sdt {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;

__global__ void fill_int_kernel(int* p, long long n, int v) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// block-wide exclusive scan of one value per thread; returns the total
__device__ int block_exclusive_scan(int v, int* s_tmp, int* total) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int nw = blockDim.x >> 6;
    // inclusive wave scan
    int x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) s_tmp[wid] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        int acc = 0;
        for (int w = 0; w < nw; w++) {
            int t = s_tmp[w];
            s_tmp[w] = acc;
            acc += t;
        }
        s_tmp[nw] = acc;
    }
    __syncthreads();
    int res = x - v + s_tmp[wid];
    *total = s_tmp[nw];
    __syncthreads();
    return res;
}

__global__ void scan_tile_sums_kernel(const int* __restrict__ in, long long n, int* __restrict__ bsum) {
    __shared__ int s_tmp[kScanThreads / 64 + 1];
    const long long base = (long long)blockIdx.x * kScanTile + (long long)threadIdx.x * kScanItems;
    int s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; k++) s += (base + k < n) ? in[base + k] : 0;
    int total;
    block_exclusive_scan(s, s_tmp, &total);
    if (threadIdx.x == 0) bsum[blockIdx.x] = total;
}

__global__ void scan_block_sums_kernel(int* bsum, int nb, int* total_out) {
    __shared__ int s_tmp[kScanThreads / 64 + 1];
    int carry = 0;
    for (int b0 = 0; b0 < nb; b0 += kScanThreads) {
        int i = b0 + threadIdx.x;
        int v = (i < nb) ? bsum[i] : 0;
        int total;
        int ex = block_exclusive_scan(v, s_tmp, &total);
        if (i < nb) bsum[i] = ex + carry;
        carry += total;
    }
    if (threadIdx.x == 0) *total_out = carry;
}

__global__ void scan_apply_kernel(const int* __restrict__ in, long long n, const int* __restrict__ bsum,
                                  int* __restrict__ out) {
    __shared__ int s_tmp[kScanThreads / 64 + 1];
    const long long base = (long long)blockIdx.x * kScanTile + (long long)threadIdx.x * kScanItems;
    int v[kScanItems];
    int s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        v[k] = (base + k < n) ? in[base + k] : 0;
        s += v[k];
    }
    int total;
    int ex = block_exclusive_scan(s, s_tmp, &total) + bsum[blockIdx.x];
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
        if (base + k < n) out[base + k] = ex;
        ex += v[k];
    }
}

void exclusive_scan(const int* d_in, long long n, int* d_out, DevBuf<int>& scratch, hipStream_t s) {
    const unsigned nb = ceil_div(n, kScanTile);
    scratch.reserve(nb + 1);
    hipLaunchKernelGGL(scan_tile_sums_kernel, dim3(nb), dim3(kScanThreads), 0, s, d_in, n,
                       scratch.get());
    hipLaunchKernelGGL(scan_block_sums_kernel, dim3(1), dim3(kScanThreads), 0, s, scratch.get(),
                       (int)nb, d_out + n);
    hipLaunchKernelGGL(scan_apply_kernel, dim3(nb), dim3(kScanThreads), 0, s, d_in, n,
                       scratch.get(), d_out);
    SDT_HIP_CHECK(hipGetLastError());
}

__global__ void bucket_count_kernel(const int* __restrict__ key, int n, int* __restrict__ cnt) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) atomicAdd(&cnt[key[i]], 1);
}

__global__ void bucket_scatter_kernel(const int* __restrict__ key, int n, int* __restrict__ fill,
                                      int* __restrict__ perm) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) perm[atomicAdd(&fill[key[i]], 1)] = i;
}

// sort each bucket's members by particle index (buckets are small)
__global__ void bucket_sort_members_kernel(const int* __restrict__ start, int nkeys,
                                           int* __restrict__ perm) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nkeys) return;
    const int b = start[k], e = start[k + 1];
    for (int i = b + 1; i < e; i++) {
        int v = perm[i];
        int j = i - 1;
        while (j >= b && perm[j] > v) {
            perm[j + 1] = perm[j];
            j--;
        }
        perm[j + 1] = v;
    }
}

__global__ void invert_perm_kernel(const int* __restrict__ perm, int n, int* __restrict__ iperm) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) iperm[perm[i]] = i;
}

void bucket_sort(const int* d_key, int n, int nkeys, int* d_perm, int* d_start, int* d_iperm,
                 SortScratch& sc, hipStream_t s) {
    sc.cnt.reserve((size_t)nkeys + 1);
    sc.fill.reserve((size_t)nkeys + 1);
    hipLaunchKernelGGL(fill_int_kernel, dim3(ceil_div(nkeys, 256)), dim3(256), 0, s, sc.cnt.get(),
                       (long long)nkeys, 0);
    hipLaunchKernelGGL(bucket_count_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, s, d_key, n,
                       sc.cnt.get());
    exclusive_scan(sc.cnt.get(), nkeys, d_start, sc.scan, s);
    SDT_HIP_CHECK(hipMemcpyAsync(sc.fill.get(), d_start, sizeof(int) * nkeys,
                                 hipMemcpyDeviceToDevice, s));
    hipLaunchKernelGGL(bucket_scatter_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, s, d_key, n,
                       sc.fill.get(), d_perm);
    hipLaunchKernelGGL(bucket_sort_members_kernel, dim3(ceil_div(nkeys, 256)), dim3(256), 0, s,
                       d_start, nkeys, d_perm);
    if (d_iperm)
        hipLaunchKernelGGL(invert_perm_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, s, d_perm, n,
                           d_iperm);
    SDT_HIP_CHECK(hipGetLastError());
}

}  // namespace sdt
