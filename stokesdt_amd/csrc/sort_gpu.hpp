// Device scan and deterministic bucket sort.
#pragma once
#include "sdt_internal.hpp"

namespace sdt {

struct SortScratch {
    DevBuf<int> cnt, fill, scan;
};

// out[0..n-1] = exclusive prefix sums of in, out[n] = total
void exclusive_scan(const int* d_in, long long n, int* d_out, DevBuf<int>& scratch, hipStream_t s);

// perm[0..n) lists the items by ascending key, items of one key in
// ascending index; start[0..nkeys] is the CSR offset of each key.
// iperm (optional) is the inverse permutation.
void bucket_sort(const int* d_key, int n, int nkeys, int* d_perm, int* d_start, int* d_iperm,
                 SortScratch& sc, hipStream_t s);

__global__ void fill_int_kernel(int* p, long long n, int v);

}  // namespace sdt
