// Pair lists and steric forces for gfx950.
//
// Pair membership is decided with exactly the reference's FP64 operations
// (src/pairlist/pair_base.cc:94-103 cell assignment, pair_r.cc:140-157 and
// pair_s.cc:148-178 tests), so the GPU sets are bit-identical to the CPU
// ones. This translation unit is compiled with -ffp-contract=off so that no
// multiply-add is fused differently from the reference.
#include <algorithm>
#include <cmath>
#include "pair_gpu.hpp"

namespace sdt {

__global__ void cell_assign_kernel(int npos, const double* __restrict__ pos, double box, int nc1,
                                   double* __restrict__ pos0, int* __restrict__ cidx) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npos) return;
    int ic[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        double x = fmod(pos[3 * i + d], box);
        x = (x >= 0.0 ? x : box + x);
        int c = (int)(x / box * nc1);
        // guards the rounding case x/box*nc1 == nc1 (out of range in the
        // reference as well)
        c = c < nc1 ? c : nc1 - 1;
        ic[d] = c;
        pos0[3 * i + d] = x;
    }
    cidx[i] = (ic[0] * nc1 + ic[1]) * nc1 + ic[2];
}

__device__ __forceinline__ int wrapc(int c, int nc1) {
    c = c >= nc1 ? c - nc1 : c;
    return c < 0 ? c + nc1 : c;
}

// One thread per row. PASS 0 counts, PASS 1 writes column ids.
template <char MODE, int PASS>
__global__ void find_pairs_kernel(int npos, const int* __restrict__ row_map,
                                  const int* __restrict__ col_map, const double* __restrict__ pos0,
                                  const double* __restrict__ rdi, const int* __restrict__ cidx,
                                  const int* __restrict__ cell_start,
                                  const int* __restrict__ cell_perm, const int* __restrict__ sidx,
                                  const double* __restrict__ ssum, int nsidx, int nc1, double box,
                                  double cutoff, double min_radius, double max_radius,
                                  int* __restrict__ cnt, const int* __restrict__ rowptr,
                                  int* __restrict__ colidx) {
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= npos) return;
    const int i = row_map ? row_map[row] : row;
    const int c1 = cidx[i];
    const int ix1 = c1 / (nc1 * nc1), iy1 = (c1 / nc1) % nc1, iz1 = c1 % nc1;
    const double x1 = pos0[3 * i], y1 = pos0[3 * i + 1], z1 = pos0[3 * i + 2];
    const double r1 = rdi[i];
    const double half = box / 2.0;
    int n = 0;
    int out = (PASS == 1) ? rowptr[row] : 0;
    for (int k = 0; k < nsidx; k++) {
        const int ix2 = wrapc(ix1 + sidx[4 * k], nc1);
        const int iy2 = wrapc(iy1 + sidx[4 * k + 1], nc1);
        const int iz2 = wrapc(iz1 + sidx[4 * k + 2], nc1);
        const int c2 = (ix2 * nc1 + iy2) * nc1 + iz2;
        const int jb = cell_start[c2], je = cell_start[c2 + 1];
        if (MODE == 'R') {
            const bool all = sidx[4 * k + 3] != 0;
            for (int t = jb; t < je; t++) {
                const int j = cell_perm[t];
                bool take = all;
                if (!all) {
                    double rx = fabs(pos0[3 * j] - x1);
                    double ry = fabs(pos0[3 * j + 1] - y1);
                    double rz = fabs(pos0[3 * j + 2] - z1);
                    rx = (rx > half ? rx - box : rx);
                    ry = (ry > half ? ry - box : ry);
                    rz = (rz > half ? rz - box : rz);
                    double rr = rx * rx + ry * ry + rz * rz;
                    take = sqrt(rr) <= cutoff;
                }
                if (take) {
                    if (PASS == 1) colidx[out++] = col_map ? col_map[j] : j;
                    n++;
                }
            }
        } else {
            const double min_sum = ssum[2 * k], max_sum = ssum[2 * k + 1];
            if (r1 + min_radius > max_sum) {
                for (int t = jb; t < je; t++) {
                    const int j = cell_perm[t];
                    if (PASS == 1) colidx[out++] = col_map ? col_map[j] : j;
                    n++;
                }
            } else if (r1 + max_radius > min_sum) {
                for (int t = jb; t < je; t++) {
                    const int j = cell_perm[t];
                    const double r2 = rdi[j];
                    bool take = false;
                    if (r1 + r2 > max_sum) {
                        take = true;
                    } else if (r1 + r2 > min_sum) {
                        double rx = fabs(pos0[3 * j] - x1);
                        double ry = fabs(pos0[3 * j + 1] - y1);
                        double rz = fabs(pos0[3 * j + 2] - z1);
                        rx = (rx > half ? rx - box : rx);
                        ry = (ry > half ? ry - box : ry);
                        rz = (rz > half ? rz - box : rz);
                        double rr = rx * rx + ry * ry + rz * rz;
                        if (rr == 0.0) rr = 1.0e-10;
                        double s = 2.0 * sqrt(rr) / (r1 + r2);
                        take = s < cutoff;
                    }
                    if (take) {
                        if (PASS == 1) colidx[out++] = col_map ? col_map[j] : j;
                        n++;
                    }
                }
            }
        }
    }
    if (PASS == 0) cnt[row] = n;
}

__global__ void sort_rows_kernel(int nrows, const int* __restrict__ rowptr, int* __restrict__ colidx) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrows) return;
    const int b = rowptr[r], e = rowptr[r + 1];
    for (int i = b + 1; i < e; i++) {
        int v = colidx[i];
        int j = i - 1;
        while (j >= b && colidx[j] > v) {
            colidx[j + 1] = colidx[j];
            j--;
        }
        colidx[j + 1] = v;
    }
}

int PairListGpu::init(int npos_, const double* h_rdi, double box_, double cutoff_, char mode_) {
    npos = npos_; box = box_; cutoff = cutoff_; mode = mode_;
    if (npos <= 0 || box <= 0.0 || cutoff <= 0.0) return 0;
    if (mode != 'R' && mode != 'S') return 0;
    max_radius = *std::max_element(h_rdi, h_rdi + npos);
    min_radius = *std::min_element(h_rdi, h_rdi + npos);
    if (min_radius <= 0.0) return 0;
    // NumCells (src/pairlist/pair_r.cc:15-22, pair_s.cc:16-24)
    if (mode == 'R') nc1 = (int)(box / (cutoff / 2.0));
    else nc1 = (int)(box / (cutoff * max_radius / 2.0));
    nc1 = nc1 == 0 ? 1 : nc1;
    SDT_REQUIRE((double)nc1 * nc1 * nc1 < 2.0e9, SDT_EINVAL, "pair list: too many cells");
    cell_size = box / nc1;
    // InitSearchIdx (src/pairlist/pair_r.cc:25-68, pair_s.cc:27-71)
    std::vector<int> sidx;
    std::vector<double> ssum;
    const long long nc2 = (long long)nc1 * nc1, nc3 = nc2 * nc1;
    for (long long i = 0; i < nc3; i++) {
        int ix = (int)(i / nc2), iy = (int)((i % nc2) / nc1), iz = (int)(i % nc1);
        ix = ix > nc1 / 2 ? ix - nc1 : ix;
        iy = iy > nc1 / 2 ? iy - nc1 : iy;
        iz = iz > nc1 / 2 ? iz - nc1 : iz;
        double dx = abs(ix) - 1 > 0 ? abs(ix) - 1 : 0;
        double dy = abs(iy) - 1 > 0 ? abs(iy) - 1 : 0;
        double dz = abs(iz) - 1 > 0 ? abs(iz) - 1 : 0;
        double rmin = sqrt(dx * dx + dy * dy + dz * dz) * cell_size;
        dx = abs(ix) + 1 > nc1 / 2 ? abs(ix) + 0.5 : abs(ix) + 1;
        dy = abs(iy) + 1 > nc1 / 2 ? abs(iy) + 0.5 : abs(iy) + 1;
        dz = abs(iz) + 1 > nc1 / 2 ? abs(iz) + 0.5 : abs(iz) + 1;
        double rmax = sqrt(dx * dx + dy * dy + dz * dz) * cell_size;
        int all = 0;
        double mn = 0.0, mx = 0.0;
        if (mode == 'R') {
            if (rmin > cutoff) continue;
            all = (rmax <= cutoff) ? 1 : 0;
        } else {
            mn = 2.0 * rmin / cutoff;
            if (mn > 2.0 * max_radius) continue;
            mx = 2.0 * rmax / cutoff;
        }
        sidx.push_back(ix); sidx.push_back(iy); sidx.push_back(iz); sidx.push_back(all);
        ssum.push_back(mn); ssum.push_back(mx);
    }
    nsidx = (int)(sidx.size() / 4);
    d_sidx.reserve(sidx.size());
    d_ssum.reserve(ssum.size());
    SDT_HIP_CHECK(sync_memcpy(d_sidx.get(), sidx.data(), sizeof(int) * sidx.size(), hipMemcpyHostToDevice));
    SDT_HIP_CHECK(sync_memcpy(d_ssum.get(), ssum.data(), sizeof(double) * ssum.size(), hipMemcpyHostToDevice));
    d_rdi.reserve(npos);
    SDT_HIP_CHECK(sync_memcpy(d_rdi.get(), h_rdi, sizeof(double) * npos, hipMemcpyHostToDevice));
    d_pos0.reserve(3 * (size_t)npos);
    d_cidx.reserve(npos);
    d_cell_perm.reserve(npos);
    d_cell_start.reserve((size_t)nc3 + 1);
    d_cnt.reserve(npos);
    d_rowptr.reserve((size_t)npos + 1);
    // estimated nnz (src/pairlist/pair_r.cc:71-92, pair_s.cc:74-96)
    long long est = 0;
    for (int i = 0; i < npos; i++) {
        double t1 = (mode == 'R') ? pow(cutoff + max_radius, 3.0)
                                  : pow((h_rdi[i] + max_radius) * cutoff / 2.0 + max_radius, 3.0);
        double t2 = pow(h_rdi[i], 3.0), t3 = pow(min_radius, 3.0);
        long long sz = (long long)((t1 - t2) / t3 * 0.6) + 1;
        est += std::min<long long>(sz, npos);
    }
    return (int)std::min<long long>(est, 2000000000LL);
}

int PairListGpu::build(const double* d_pos, const int* row_map, const int* col_map,
                       bool sort_rows, hipStream_t s) {
    const unsigned g = ceil_div(npos, 256);
    const int nc3 = nc1 * nc1 * nc1;
    hipLaunchKernelGGL(cell_assign_kernel, dim3(g), dim3(256), 0, s, npos, d_pos, box, nc1,
                       d_pos0.get(), d_cidx.get());
    bucket_sort(d_cidx.get(), npos, nc3, d_cell_perm.get(), d_cell_start.get(), nullptr, sc, s);
#define SDT_FIND(M, PS)                                                                        \
    hipLaunchKernelGGL((find_pairs_kernel<M, PS>), dim3(g), dim3(256), 0, s, npos, row_map,     \
                       col_map, d_pos0.get(), d_rdi.get(), d_cidx.get(), d_cell_start.get(),    \
                       d_cell_perm.get(), d_sidx.get(), d_ssum.get(), nsidx, nc1, box, cutoff,  \
                       min_radius, max_radius, d_cnt.get(), d_rowptr.get(), d_colidx.get())
    if (mode == 'R') SDT_FIND('R', 0); else SDT_FIND('S', 0);
    exclusive_scan(d_cnt.get(), npos, d_rowptr.get(), sc.scan, s);
    int h_nnz = 0;
    SDT_HIP_CHECK(hipMemcpyAsync(&h_nnz, d_rowptr.get() + npos, sizeof(int), hipMemcpyDeviceToHost, s));
    SDT_HIP_CHECK(hipStreamSynchronize(s));
    nnz = h_nnz;
    if ((long long)nnz > colcap) {
        d_colidx.reserve((size_t)nnz + nnz / 8 + 16);
        colcap = (long long)d_colidx.n;
    }
    if (mode == 'R') SDT_FIND('R', 1); else SDT_FIND('S', 1);
#undef SDT_FIND
    if (sort_rows)
        hipLaunchKernelGGL(sort_rows_kernel, dim3(g), dim3(256), 0, s, npos, d_rowptr.get(),
                           d_colidx.get());
    SDT_HIP_CHECK(hipGetLastError());
    return nnz;
}

__global__ void steric_kernel(int npos, const double* __restrict__ pos, const double* __restrict__ rdi,
                              double box, double r0, double k0, const int* __restrict__ rowptr,
                              const int* __restrict__ colidx, double* __restrict__ f) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npos) return;
    const double x1 = pos[3 * i], y1 = pos[3 * i + 1], z1 = pos[3 * i + 2];
    const double aa = rdi[i];
    double f0 = 0.0, f1 = 0.0, f2 = 0.0;
    for (int k = rowptr[i]; k < rowptr[i + 1]; k++) {
        const int j = colidx[k];
        if (j == i) continue;
        const double rx = remainder(x1 - pos[3 * j], box);
        const double ry = remainder(y1 - pos[3 * j + 1], box);
        const double rz = remainder(z1 - pos[3 * j + 2], box);
        const double r = sqrt(rx * rx + ry * ry + rz * rz);
        const double force = k0 * ((aa + rdi[j]) * r0 / 2.0 - r);
        f0 += force * (rx / r);
        f1 += force * (ry / r);
        f2 += force * (rz / r);
    }
    f[3 * i] += f0;
    f[3 * i + 1] += f1;
    f[3 * i + 2] += f2;
}

// Steric forces straight from a cell list whose cells are at least as wide
// as the largest contact distance cutoff * max_radius, so every partner lies
// in the 27 cells around a particle's own. A pair is taken when
// 2 r / (a_i + a_j) < cutoff with r from the wrapped positions exactly as in
// find_pairs_kernel<'S'> (pair_s.cc:148-178; the reference's whole-cell
// shortcuts only accept pairs that pass this test), and the force of every
// taken j != i is the one steric_kernel adds. Eight lanes share a particle,
// each visiting every eighth of the 27 cells (members in ascending order);
// the lanes' sums are combined in a fixed tree. No pair list is stored and
// the host never waits for its size.
constexpr int kStericLanes = 8;   // lanes per particle, each taking every 8th cell

__global__ void steric_cell_kernel(int npos, const double* __restrict__ pos,
                                   const double* __restrict__ pos0,
                                   const double* __restrict__ rdi, const int* __restrict__ cidx,
                                   const int* __restrict__ cell_start,
                                   const int* __restrict__ cell_perm, int nc1, double box,
                                   double cutoff, double r0, double k0, double* __restrict__ f) {
    const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int i = (int)(tid / kStericLanes), sub = (int)(tid % kStericLanes);
    const bool on = i < npos;
    double f0 = 0.0, f1 = 0.0, f2 = 0.0;
    if (on) {
        const int c1 = cidx[i];
        const int ix1 = c1 / (nc1 * nc1), iy1 = (c1 / nc1) % nc1, iz1 = c1 % nc1;
        const double x1 = pos0[3 * i], y1 = pos0[3 * i + 1], z1 = pos0[3 * i + 2];
        const double px = pos[3 * i], py = pos[3 * i + 1], pz = pos[3 * i + 2];
        const double r1 = rdi[i];
        const double half = box / 2.0;
        for (int d = sub; d < 27; d += kStericLanes) {
            const int dx = d / 9 - 1, dy = (d / 3) % 3 - 1, dz = d % 3 - 1;
            const int c2 = (wrapc(ix1 + dx, nc1) * nc1 + wrapc(iy1 + dy, nc1)) * nc1 +
                           wrapc(iz1 + dz, nc1);
            const int jb = cell_start[c2], je = cell_start[c2 + 1];
            for (int t = jb; t < je; t++) {
                const int j = cell_perm[t];
                if (j == i) continue;
                const double r2 = rdi[j];
                double rx = fabs(pos0[3 * j] - x1);
                double ry = fabs(pos0[3 * j + 1] - y1);
                double rz = fabs(pos0[3 * j + 2] - z1);
                rx = (rx > half ? rx - box : rx);
                ry = (ry > half ? ry - box : ry);
                rz = (rz > half ? rz - box : rz);
                double rr = rx * rx + ry * ry + rz * rz;
                if (rr == 0.0) rr = 1.0e-10;
                const double sv = 2.0 * sqrt(rr) / (r1 + r2);
                if (!(sv < cutoff)) continue;
                const double qx = remainder(px - pos[3 * j], box);
                const double qy = remainder(py - pos[3 * j + 1], box);
                const double qz = remainder(pz - pos[3 * j + 2], box);
                const double r = sqrt(qx * qx + qy * qy + qz * qz);
                const double force = k0 * ((r1 + r2) * r0 / 2.0 - r);
                f0 += force * (qx / r);
                f1 += force * (qy / r);
                f2 += force * (qz / r);
            }
        }
    }
    // a particle's lanes are consecutive: sum them in a fixed tree
#pragma unroll
    for (int off = kStericLanes / 2; off > 0; off >>= 1) {
        f0 += __shfl_xor(f0, off, kStericLanes);
        f1 += __shfl_xor(f1, off, kStericLanes);
        f2 += __shfl_xor(f2, off, kStericLanes);
    }
    if (on && sub == 0) {
        f[3 * i] += f0;
        f[3 * i + 1] += f1;
        f[3 * i + 2] += f2;
    }
}

void PairListGpu::steric_direct(const double* d_pos, double r0, double k0, double* d_f,
                                hipStream_t s) {
    SDT_REQUIRE(mode == 'S', SDT_EINVAL, "steric_direct needs a steric ('S') cell list");
    const int ncd = (int)(box / (cutoff * max_radius));
    if (ncd < 3) {
        // fewer than 3 cells per side would visit a cell twice: use the list
        build(d_pos, nullptr, nullptr, false, s);
        steric_accumulate(npos, d_pos, d_rdi.get(), box, r0, k0, d_rowptr.get(), d_colidx.get(),
                          d_f, s);
        return;
    }
    const unsigned g = ceil_div(npos, 256);
    const int nc3 = ncd * ncd * ncd;
    d_dcidx.reserve(npos);
    d_dperm.reserve(npos);
    d_dstart.reserve((size_t)nc3 + 1);
    hipLaunchKernelGGL(cell_assign_kernel, dim3(g), dim3(256), 0, s, npos, d_pos, box, ncd,
                       d_pos0.get(), d_dcidx.get());
    bucket_sort(d_dcidx.get(), npos, nc3, d_dperm.get(), d_dstart.get(), nullptr, sc, s);
    hipLaunchKernelGGL(steric_cell_kernel, dim3(ceil_div((size_t)npos * kStericLanes, 256)),
                       dim3(256), 0, s, npos, d_pos, d_pos0.get(),
                       d_rdi.get(), d_dcidx.get(), d_dstart.get(), d_dperm.get(), ncd, box, cutoff,
                       r0, k0, d_f);
    SDT_HIP_CHECK(hipGetLastError());
}

void steric_accumulate(int npos, const double* d_pos, const double* d_rdi, double box, double r0,
                       double k0, const int* d_rowptr, const int* d_colidx, double* d_f,
                       hipStream_t s) {
    hipLaunchKernelGGL(steric_kernel, dim3(ceil_div(npos, 256)), dim3(256), 0, s, npos, d_pos,
                       d_rdi, box, r0, k0, d_rowptr, d_colidx, d_f);
    SDT_HIP_CHECK(hipGetLastError());
}

}  // namespace sdt
