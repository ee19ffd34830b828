// SPME reciprocal-space kernels for gfx950 (FP32 and FP64 instantiations).
//
// Reference algorithm: src/mob/spme.cc (UpdateSpmeEngine_ :366-509, Spread
// :313-363, ApplyInfluence :201-310, Interpolate :163-198, ComputeSpline
// :71-160, ComputeSpmeRecip_ :714-839). The GPU form recomputes B-spline
// weights from per-particle fractional offsets instead of storing the
// N x p^3 P/ind tables, fuses all b right-hand sides into each launch,
// spreads by owner-computes mesh tiles accumulated in LDS (plain stores, no
// global atomics, no zero-fill pass), and fuses the influence multiply
// between the forward and inverse z-transforms.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include "fft_lds.hpp"
#include "real_gpu.hpp"
#include "spme_gpu.hpp"

namespace sdt {

// ----------------------------------------------------------- B-splines -----
// w[x] = M_P(x + 1 - frac), x = 0..P-1, by raising the order of
// a[k] = M_n(u + k), u = 1 - frac, from n = 1 to P (Cox-de Boor). Same values
// as BSpline(P, floor(c) - c + x) in src/mob/spme.cc:38-68,469-473.
template <typename T, int P>
__device__ __forceinline__ void bspline_weights(T frac, T* w) {
    const T u = T(1) - frac;
    w[0] = T(1);
#pragma unroll
    for (int k = 1; k < P; k++) w[k] = T(0);
#pragma unroll
    for (int n = 2; n <= P; n++) {
        const T inv = T(1) / T(n - 1);
#pragma unroll
        for (int k = n - 1; k >= 1; k--)
            w[k] = ((u + T(k)) * w[k] + (T(n) - u - T(k)) * w[k - 1]) * inv;
        w[0] = u * w[0] * inv;
    }
}

// the weights from the FP64 fractional offset, evaluated in FP64 and rounded
// to the arithmetic type once (FP32 weights carry no error from the offset)
template <typename T, int P>
__device__ __forceinline__ void bspline_weights_r(double frac, T* w) {
    double wd[P];
    bspline_weights<double, P>(frac, wd);
#pragma unroll
    for (int k = 0; k < P; k++) w[k] = (T)wd[k];
}

// ------------------------------------------------------------ prepare ------
// Wrap into [0, L) (pair-list convention, src/pairlist/pair_base.cc:94-99),
// scaled coordinate c = x / L * K, base index as src/mob/spme.cc:419-443.
template <typename T, int P>
__global__ void spme_prepare_kernel(int npos, const double* __restrict__ pos, double box, int K,
                                    int ts, int ntile, double* __restrict__ frac_u,
                                    int* __restrict__ base_u, int* __restrict__ key) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npos) return;
    int t[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        double x = fmod(pos[3 * i + d], box);
        x = (x >= 0.0 ? x : box + x);
        double c = x / box * K;
        double g = floor(c);
        int b0 = (P % 2 == 0) ? (int)(g - P / 2 + 1) : (int)(g - (P - 1) / 2);
        int b = ((b0 % K) + K) % K;
        frac_u[3 * i + d] = c - g;
        base_u[3 * i + d] = b;
        int tt = b / ts;
        t[d] = tt < ntile ? tt : ntile - 1;
    }
    key[i] = (t[2] * ntile + t[1]) * ntile + t[0];
}

template <typename T>
__global__ void spme_gather_kernel(int npos, const int* __restrict__ perm,
                                   const double* __restrict__ frac_u, const int* __restrict__ base_u,
                                   const double* __restrict__ rdi, double* __restrict__ frac,
                                   int* __restrict__ base, double* __restrict__ a2) {
    const int is = blockIdx.x * blockDim.x + threadIdx.x;
    if (is >= npos) return;
    const int i = perm[is];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        frac[3 * is + d] = frac_u[3 * i + d];
        base[3 * is + d] = base_u[3 * i + d];
    }
    if (a2) a2[is] = rdi[i] * rdi[i];
}

// ------------------------------------------------------------- spread ------
// Gather form: each workgroup owns a super-tile of 4 ts x ts x ts mesh tiles
// along x (one per wave) for RB right-hand sides. Lane l of a wave owns the
// x-row (y = l % ts, z = l / ts) of its tile and accumulates its ts points
// in registers. The particles whose B-spline stencils reach the super-tile
// (base points in the 4+1 x 2 x 2 surrounding tiles, filtered by overlap) are
// staged in LDS with their weights and forces; every lane then walks the
// staged list. No atomics: every mesh point is written once, by its owner.

// exclusive prefix over the workgroup of a 0/1 predicate; returns this
// thread's slot and the total (all threads must call)
__device__ __forceinline__ int wg_compact(bool pred, int* s_wave, int* total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const unsigned long long m = __ballot(pred);
    const int before = __popcll(m & ((1ull << lane) - 1ull));
    if (lane == 0) s_wave[wave] = __popcll(m);
    __syncthreads();
    int off = 0, tot = 0;
    const int nw = blockDim.x >> 6;
    for (int w = 0; w < nw; w++) {
        const int c = s_wave[w];
        off += (w < wave) ? c : 0;
        tot += c;
    }
    __syncthreads();
    *total = tot;
    return off + before;
}

__device__ __forceinline__ int wrapk(int d, int K) {
    d += (d < 0) ? K : 0;
    return d - ((d >= K) ? K : 0);
}

template <typename T, int P, int RBNC>
__global__ void __launch_bounds__(256)
spread_gather_kernel(int K, int ts, int ntile, int nsx, const int* __restrict__ tile_start,
                     const double* __restrict__ frac, const int* __restrict__ base,
                     const double* __restrict__ a2, const double* __restrict__ X, int b, int RB,
                     int NC, int cap, T* __restrict__ mesh) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    // record: x weights in a zero-padded window wxp[0 .. P+13] (wxp[7 + k] =
    // wx[k]), wy[P], wz[P], bx, by, bz, f[RBNC]. A wave reads the weight of
    // its tile point x at wxp[7 - sd + x], sd = base - tile start in
    // [-(P-1), ts-1]: no per-point selects or range tests.
    constexpr int OWY = P + 14, OWZ = 2 * P + 14, OB = 3 * P + 14, OF = 3 * P + 17;
    constexpr int REC = OF + RBNC;
    T* stage = reinterpret_cast<T*>(smem_raw);
    __shared__ int s_wave[4];
    __shared__ int s_off[21];                   // candidate offsets of the base tiles
    __shared__ int s_beg[20];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    // block -> (super-tile x, tile y, tile z), rhs chunk; XCD-contiguous
    const unsigned lid = xcd_logical_id(blockIdx.x + blockIdx.y * gridDim.x, gridDim.x * gridDim.y);
    const int bxl = (int)(lid % gridDim.x), byl = (int)(lid / gridDim.x);
    const int sxi = bxl % nsx;
    const int tyz = bxl / nsx;
    const int ty = tyz % ntile, tz = tyz / ntile;
    const int r0 = byl * RB;
    const int rb = min(RB, b - r0);
    const int rbnc = rb * NC;
    const int stx = sxi * 4;                    // first x tile of the super-tile
    const int nxt = min(4, ntile - stx);        // x tiles present
    const int sx0 = stx * ts, sxlen = nxt * ts;
    const int sy0 = ty * ts, sz0 = tz * ts;
    // this lane's row and x range
    const int wx0 = sx0 + wave * ts;
    const bool wave_on = wave < nxt;
    const bool lane_on = wave_on && lane < ts * ts;
    const int yl = sy0 + (lane % ts), zl = sz0 + (lane / ts);
    T acc[8][RBNC];
#pragma unroll
    for (int x = 0; x < 8; x++)
#pragma unroll
        for (int q = 0; q < RBNC; q++) acc[x][q] = T(0);

    // contributor base tiles: x tiles stx-1 .. stx+nxt-1, y ty-1..ty, z tz-1..tz
    const int nbx = min(nxt + 1, ntile);
    const int nbt = 4 * nbx;
    if (threadIdx.x == 0) {
        int off = 0;
        for (int k = 0; k < nbt; k++) {
            const int jx = k % nbx, jy = (k / nbx) & 1, jz = k / (2 * nbx);
            const int btz = (tz - 1 + jz + ntile) % ntile;
            const int bty = (ty - 1 + jy + ntile) % ntile;
            const int btx = (stx - 1 + jx + ntile) % ntile;
            const int bt = (btz * ntile + bty) * ntile + btx;
            const int pb = tile_start[bt], pe = tile_start[bt + 1];
            s_beg[k] = pb;
            s_off[k] = off;
            off += pe - pb;
        }
        s_off[nbt] = off;
    }
    __syncthreads();
    const int ncand = s_off[nbt];

    auto accumulate = [&](int n) __attribute__((always_inline)) {
        if (!lane_on) return;
        for (int sidx = 0; sidx < n; sidx++) {
            const T* rec = stage + (size_t)sidx * REC;
            const int bx = (int)rec[OB], by = (int)rec[OB + 1], bz = (int)rec[OB + 2];
            const int dxw = wrapk(bx - wx0, K);
            if (!(dxw < ts || K - dxw < P)) continue;          // wave-uniform
            const int ly = wrapk(yl - by, K), lz = wrapk(zl - bz, K);
            if (ly >= P || lz >= P) continue;
            const int sd = dxw < ts ? dxw : dxw - K;
            const T* wxp = rec + (7 - sd);
            // every load of the candidate is issued before the first use;
            // points x >= ts accumulate window zeros / unused values and are
            // never stored
            T f[RBNC], wx8[8];
#pragma unroll
            for (int q = 0; q < RBNC; q++) f[q] = rec[OF + q];
#pragma unroll
            for (int x = 0; x < 8; x++) wx8[x] = wxp[x];
            const T wyz = rec[OWY + ly] * rec[OWZ + lz];
#pragma unroll
            for (int x = 0; x < 8; x++) {
                const T w = wx8[x] * wyz;
#pragma unroll
                for (int q = 0; q < RBNC; q++) acc[x][q] += w * f[q];
            }
        }
    };

    // candidates in rounds of 2 per thread: all loads of a round are in flight
    // together; takers are compacted in candidate order (deterministic sums)
    constexpr int U = 2;
    int nst = 0;
    for (int c0 = 0; c0 < ncand; c0 += U * 256) {
        int is_u[U];
        bool take_u[U];
        int b3u[U][3];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int ci = c0 + u * 256 + (int)threadIdx.x;
            is_u[u] = -1;
            take_u[u] = false;
            if (ci < ncand) {
                int k = 0;
                while (k + 1 < nbt && s_off[k + 1] <= ci) k++;
                is_u[u] = s_beg[k] + (ci - s_off[k]);
            }
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            if (is_u[u] >= 0) {
                const int is = is_u[u];
                b3u[u][0] = base[3 * is]; b3u[u][1] = base[3 * is + 1]; b3u[u][2] = base[3 * is + 2];
            }
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            if (is_u[u] >= 0) {
                const int dx = wrapk(b3u[u][0] - sx0, K);
                const int dy = wrapk(b3u[u][1] - sy0, K);
                const int dz = wrapk(b3u[u][2] - sz0, K);
                take_u[u] = (dx < sxlen || K - dx < P) && (dy < ts || K - dy < P) &&
                            (dz < ts || K - dz < P);
            }
        }
#pragma unroll
        for (int u = 0; u < U; u++) {
            int tot;
            const int slot = wg_compact(take_u[u], s_wave, &tot);
            // the takers fill the stage in candidate order; a full stage is
            // accumulated and emptied (tot may exceed the capacity)
            for (int done = 0; done < tot;) {
                if (nst == cap) {
                    __syncthreads();
                    accumulate(nst);
                    __syncthreads();
                    nst = 0;
                }
                const int take = min(cap - nst, tot - done);
                if (take_u[u] && slot >= done && slot < done + take) {
                    const int is = is_u[u];
                    T* rec = stage + (size_t)(nst + slot - done) * REC;
                    // forces: slot q holds rhs r, component c (mono: q = 3 r + c;
                    // poly: q = 6 r + c for f, 6 r + 3 + c for a^2 f); every
                    // index is a compile-time constant (no register arrays
                    // indexed at run time)
                    const double sc = (NC == 6) ? a2[is] : 1.0;
#pragma unroll
                    for (int q = 0; q < RBNC; q++) {
                        const int r = (NC == 6) ? q / 6 : q / 3;
                        const bool scaled = (NC == 6) && (q % 6) >= 3;
                        T fv = T(0);
                        if (r < rb) {
                            const double x = X[(size_t)(3 * is + q % 3) * b + r0 + r];
                            fv = (T)(scaled ? x * sc : x);
                        }
                        rec[OF + q] = fv;
                    }
                    T w[P];
                    bspline_weights_r<T, P>(frac[3 * is], w);
#pragma unroll
                    for (int k = 0; k < 7; k++) rec[k] = T(0);
#pragma unroll
                    for (int k = 0; k < P; k++) rec[7 + k] = w[k];
#pragma unroll
                    for (int k = 7 + P; k < P + 14; k++) rec[k] = T(0);
                    bspline_weights_r<T, P>(frac[3 * is + 1], w);
#pragma unroll
                    for (int k = 0; k < P; k++) rec[OWY + k] = w[k];
                    bspline_weights_r<T, P>(frac[3 * is + 2], w);
#pragma unroll
                    for (int k = 0; k < P; k++) rec[OWZ + k] = w[k];
#pragma unroll
                    for (int d = 0; d < 3; d++) rec[OB + d] = (T)b3u[u][d];
                }
                nst += take;
                done += take;
            }
        }
        __syncthreads();
    }
    accumulate(nst);
    if (!lane_on) return;
    const size_t K3 = (size_t)K * K * K;
    const size_t row = ((size_t)zl * K + yl) * K + wx0;
#pragma unroll
    for (int q = 0; q < RBNC; q++) {
        if (q < rbnc) {
            T* m = mesh + (size_t)(r0 * NC + q) * K3 + row;
#pragma unroll
            for (int x = 0; x < 8; x++)
                if (x < ts) m[x] = acc[x][q];
        }
    }
}

// logical column block of a workgroup (the y part of its XCD-remapped id)
__device__ __forceinline__ unsigned byl_of(unsigned linear, unsigned n, unsigned gx) {
    return xcd_logical_id(linear, n) / gx;
}

// Point-owner form (K % 4 == 0): a workgroup of 8 waves owns a 32 x 4 x 4
// block of mesh points (wave w: x in [4w, 4w+4)), lane l one point
// (x = l % 4, y = (l / 4) % 4, z = l / 16), and NQ meshes (NQ / NC
// right-hand sides) accumulated in registers. Candidates (base points whose
// stencil reaches the block) are staged in LDS in candidate order; every wave
// lists the candidates that reach its 4 x 4 x 4 sub-block, and each lane
// adds w = wx[x - bx] wy[y - by] wz[z - bz] (0 outside the stencil) times the
// NQ forces. Deterministic, no atomics, small per-candidate overhead.
constexpr int kPtThreads = 512;
constexpr int kPtMaxTiles = 200;   // contributor bucket tiles of one block (host-checked)

//
// BLK: the forces come from the column-blocked copy Xb[cb][j][c][r] of the
// packed input (NQ / 3 columns per block, arithmetic type T, monodisperse):
// a candidate's NQ values are one contiguous run, staged with vector loads,
// and accumulator q holds component q / (NQ / 3) of column q % (NQ / 3).
//
// WRAP: K < 32 + P - 1, a stencil can reach both ends of a block along x and
// stencil offsets are taken modulo K.
// SGF (with BLK): the forces are not staged; every wave reads a candidate's
// NQ values at a wave-uniform address, through the scalar data cache into
// scalar registers, so the FMAs take them as scalar operands and LDS serves
// only the per-lane weights.
// MF (FP32, BLK, 48 meshes, staged forces): the sums run on the matrix
// cores. A wave's sub-block is the product W F of its 64 points x listed
// candidates weight matrix and the candidates' 48 forces, taken 4 candidates
// at a time by v_mfma_f32_16x16x4f32: m-tile m = the 16 points of z plane m
// (row i = 4 y + x), n-tile t = component t of the 16 columns. Each lane
// forms its A entry (point l % 16, candidate l / 16) from the staged 1D
// weights, so no lane is spent on points outside a stencil, and its result
// registers hold 4 consecutive x points of one mesh (one 16-byte store).
// MA (with MF): one workgroup per mesh block takes every 16-column block in
// turn (grid y = 1). When all its candidates fit the stage, they are found
// and their weights staged once; each further column block then only loads
// the candidates' forces into the stage before its sums and stores. The
// meshes leave through LDS tiles after the stage (16 columns x 4 rows x 32
// points of one z plane and component at a time, rows padded to 132 floats
// so that the waves' 16-byte pieces land in distinct banks; two tiles used
// in turn): every store instruction then writes whole 128-byte rows.
constexpr int kMaTileStride = 132;                       // floats per column in the tile
constexpr int kMaTileBytes = 2 * 16 * kMaTileStride * 4;     // two tiles, 16896
template <typename T, int P, int NQ, bool BLK, bool WRAP, bool SGF = false, bool MF = false,
          bool MA = false>
__global__ void __launch_bounds__(kPtThreads) __attribute__((amdgpu_waves_per_eu(NQ <= 12 ? 6 : (MA ? 4 : 1))))
spread_pt_kernel(int K, int ts, int ntile, int nsx, const int* __restrict__ tile_start,
                 const double* __restrict__ frac, const int* __restrict__ base,
                 const double* __restrict__ a2, const double* __restrict__ X,
                 const T* __restrict__ Xb, int npos, int b, int RB, int NC, int cap,
                 T* __restrict__ mesh, int tiled_out) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    // record: wx[P] 0, wy[P] 0, wz[P] 0, base - block origin (3 ints in the
    // slots of bx by bz), pad, f[NQ] (16-byte aligned)
    constexpr int OWY = P + 1, OWZ = 2 * (P + 1), OB = 3 * (P + 1);
    constexpr int OF = (OB + 3 + 3) & ~3;
    constexpr int REC = SGF ? OF : OF + NQ;
    static_assert(NQ % 4 == 0, "NQ must be a multiple of 4");
    static_assert(!SGF || BLK, "scalar force reads need the blocked copy");
    static_assert(!MF || (BLK && !SGF && !WRAP && NQ == 48 && sizeof(T) == 4),
                  "matrix-core sums: FP32, blocked staged forces, 48 meshes, no wrap");
    static_assert(!MA || (BLK && !SGF && !WRAP), "column blocks in turn: blocked staged forces");
    const T* __restrict__ xbb = Xb + (size_t)byl_of(blockIdx.x + blockIdx.y * gridDim.x,
                                                    gridDim.x * gridDim.y, gridDim.x) * npos * NQ;
    T* stage = reinterpret_cast<T*>(smem_raw);
    int* s_list = reinterpret_cast<int*>(stage + (size_t)cap * REC);   // [8][cap]
    int* s_pid = s_list + 8 * cap;                                       // [cap]
    // MA: output tile after the stage (16-byte aligned)
    [[maybe_unused]] T* ma_tile =
        reinterpret_cast<T*>(smem_raw + ((((size_t)cap * (REC * sizeof(T) + 9 * sizeof(int))) + 15) & ~(size_t)15));
    __shared__ int s_wave[kPtThreads / 64];
    __shared__ int s_off[kPtMaxTiles + 1];
    __shared__ int s_beg[kPtMaxTiles];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const unsigned lid = xcd_logical_id(blockIdx.x + blockIdx.y * gridDim.x, gridDim.x * gridDim.y);
    const int bxl = (int)(lid % gridDim.x), byl = (int)(lid / gridDim.x);
    const int K4 = K >> 2;
    const int sxi = bxl % nsx;
    const int yz = bxl / nsx;
    const int x0 = sxi * 32, y0 = (yz % K4) * 4, z0 = (yz / K4) * 4;
    const int xlen = min(32, K - x0);
    int r0 = byl * (NQ / NC);
    int rb = min(NQ / NC, b - r0);
    int nq = rb * NC;
    const bool wave_on = 4 * wave < xlen;
    // this lane's point, relative to the block origin
    const int lx = 4 * wave + (lane & 3), ly = (lane >> 2) & 3, lz = lane >> 4;
    const int px = x0 + lx, py = y0 + ly, pz = z0 + lz;
    T acc[NQ];
#pragma unroll
    for (int q = 0; q < NQ; q++) acc[q] = T(0);
    // MF: result tiles [z plane m][component t]
    typedef float f4 __attribute__((ext_vector_type(4)));
    f4 accm[4][3];
#pragma unroll
    for (int m = 0; m < 4; m++)
#pragma unroll
        for (int t = 0; t < 3; t++) accm[m][t] = f4{0.f, 0.f, 0.f, 0.f};

    // contributor base tiles (distinct, with wrap): bases in [lo, lo + len);
    // wave 0 reads the bucket bounds of 64 tiles at a time and scans the
    // counts into the candidate offsets
    if (wave == 0) {
        int first[3], cnt[3];
        const int lo[3] = {x0 - (P - 1), y0 - (P - 1), z0 - (P - 1)};
        const int len[3] = {xlen + P - 1, 4 + P - 1, 4 + P - 1};
#pragma unroll
        for (int d = 0; d < 3; d++) {
            if (len[d] >= K) { first[d] = 0; cnt[d] = ntile; continue; }
            const int a = (lo[d] >= 0) ? lo[d] / ts : -((-lo[d] + ts - 1) / ts);
            const int e = lo[d] + len[d] - 1;
            const int bt = (e >= 0) ? e / ts : -((-e + ts - 1) / ts);
            cnt[d] = min(bt - a + 1, ntile);
            first[d] = ((a % ntile) + ntile) % ntile;
        }
        const int ntot = cnt[0] * cnt[1] * cnt[2];
        int run = 0;
        for (int k0 = 0; k0 < ntot; k0 += 64) {
            const int k = k0 + lane;
            int pb = 0, n = 0;
            if (k < ntot) {
                const int jx = k % cnt[0], jyz = k / cnt[0];
                const int jy = jyz % cnt[1], jz = jyz / cnt[1];
                const int tt = (((first[2] + jz) % ntile) * ntile + (first[1] + jy) % ntile) *
                                   ntile + (first[0] + jx) % ntile;
                pb = tile_start[tt];
                n = tile_start[tt + 1] - pb;
            }
            int inc = n;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const int t = __shfl_up(inc, o);
                if (lane >= o) inc += t;
            }
            if (k < ntot) {
                s_beg[k] = pb;
                s_off[k] = run + inc - n;
            }
            run += __shfl(inc, 63);
        }
        if (lane == 0) {
            s_off[ntot] = run;
            s_wave[0] = ntot;   // tile count, read below before s_wave is reused
        }
    }
    __syncthreads();
    const int nbt = s_wave[0];
    const int ncand = s_off[nbt];
    __syncthreads();

    // accumulate the first n staged candidates
    auto accumulate = [&](int n) __attribute__((always_inline)) {
        int* list = s_list + wave * cap;
        int cnt = 0;
        if (wave_on) {
            for (int c0 = 0; c0 < n; c0 += 64) {
                const int c = c0 + lane;
                bool hit = false;
                if (c < n) {
                    const int* rl = reinterpret_cast<const int*>(stage + (size_t)c * REC + OB);
                    const int dx = wrapk(rl[0] - 4 * wave, K);
                    const int dy = wrapk(rl[1], K);
                    const int dz = wrapk(rl[2], K);
                    hit = (dx < 4 || K - dx < P) && (dy < 4 || K - dy < P) && (dz < 4 || K - dz < P);
                }
                const unsigned long long m = __ballot(hit);
                if (hit) list[cnt + __popcll(m & ((1ull << lane) - 1ull))] = c;
                cnt += __popcll(m);
            }
        }
        // the list is written and read by this wave only (LDS accesses of a
        // wave complete in order)
        __builtin_amdgcn_wave_barrier();
        if constexpr (MF) {
            // A entry of lane l: point (x, y) = (l % 4, (l / 4) % 4) of the
            // sub-block, candidate l / 16 of the step; B entry: force
            // component t of column l % 16 of the same candidate. Past the
            // list end both are 0.
            const int kq = lane >> 4, jj = lane & 15;
            const int ax = 4 * wave + (lane & 3), ay = (lane >> 2) & 3;
            for (int k0 = 0; k0 < cnt; k0 += 4) {
                const int k = k0 + kq;
                T wxy = T(0), wz[4] = {T(0), T(0), T(0), T(0)}, bf[3] = {T(0), T(0), T(0)};
                if (k < cnt) {
                    const T* rec = stage + (size_t)list[k] * REC;
                    const int* rl = reinterpret_cast<const int*>(rec + OB);
                    const int dx = (int)min((unsigned)(ax - rl[0]), (unsigned)P);
                    const int dy = (int)min((unsigned)(ay - rl[1]), (unsigned)P);
                    wxy = rec[dx] * rec[OWY + dy];
#pragma unroll
                    for (int m = 0; m < 4; m++)
                        wz[m] = rec[OWZ + (int)min((unsigned)(m - rl[2]), (unsigned)P)];
#pragma unroll
                    for (int t = 0; t < 3; t++) bf[t] = rec[OF + 16 * t + jj];
                }
#pragma unroll
                for (int m = 0; m < 4; m++) {
                    const T a = wxy * wz[m];
#pragma unroll
                    for (int t = 0; t < 3; t++)
                        accm[m][t] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, bf[t], accm[m][t], 0, 0, 0);
                }
            }
        } else if constexpr (!WRAP) {
            // no stencil reaches both ends of the block: the offset of this
            // lane's point in a candidate's stencil is a plain difference
            // (out of the stencil: P, a zero weight)
            for (int k = 0; k < cnt; k++) {
                const T* rec = stage + (size_t)list[k] * REC;
                const int* rl = reinterpret_cast<const int*>(rec + OB);
                const int dx = (int)min((unsigned)(lx - rl[0]), (unsigned)P);
                const int dy = (int)min((unsigned)(ly - rl[1]), (unsigned)P);
                const int dz = (int)min((unsigned)(lz - rl[2]), (unsigned)P);
                const T w = rec[dx] * rec[OWY + dy] * rec[OWZ + dz];
                if constexpr (SGF) {
                    const int pid = __builtin_amdgcn_readfirstlane(s_pid[list[k]]);
                    const T* fp = xbb + (size_t)pid * NQ;
#pragma unroll
                    for (int q = 0; q < NQ; q++) acc[q] += w * fp[q];
                } else {
#pragma unroll
                    for (int q = 0; q < NQ; q++) acc[q] += w * rec[OF + q];
                }
            }
        } else {
            for (int k = 0; k < cnt; k++) {
                const T* rec = stage + (size_t)list[k] * REC;
                const int* rl = reinterpret_cast<const int*>(rec + OB);
                const int dx = min(wrapk(lx - rl[0], K), P);
                const int dy = min(wrapk(ly - rl[1], K), P);
                const int dz = min(wrapk(lz - rl[2], K), P);
                const T w = rec[dx] * rec[OWY + dy] * rec[OWZ + dz];
                if constexpr (SGF) {
                    const int pid = __builtin_amdgcn_readfirstlane(s_pid[list[k]]);
                    const T* fp = xbb + (size_t)pid * NQ;
#pragma unroll
                    for (int q = 0; q < NQ; q++) acc[q] += w * fp[q];
                } else {
#pragma unroll
                    for (int q = 0; q < NQ; q++) acc[q] += w * rec[OF + q];
                }
            }
        }
    };

    int nst = 0;
    bool staged_all = false;   // MA: the stage holds every candidate
    [[maybe_unused]] const int ncb = MA ? (b + NQ / NC - 1) / (NQ / NC) : 1;
    int cb = 0;
    // MA: the next column block starts here (a jump back, not a loop, so
    // that the other instances compile without one)
next_column_block:
    {
    const int cbl = MA ? cb : byl;
    if constexpr (MA) {
        r0 = cb * (NQ / NC);
        rb = min(NQ / NC, b - r0);
        nq = rb * NC;
        __syncthreads();   // every wave is done with the previous block's forces
    }
    if (staged_all) {
        // the staged candidates' forces of this column block
        constexpr int VW = 16 / sizeof(T);
        constexpr int NV = NQ / VW;
        constexpr int U = 4;
        typedef T VT __attribute__((ext_vector_type(16 / sizeof(T))));
        const VT* src = reinterpret_cast<const VT*>(Xb + (size_t)cbl * npos * NQ);
        for (int e0 = threadIdx.x; e0 < nst * NV; e0 += kPtThreads * U) {
            VT v[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int e = e0 + u * kPtThreads;
                v[u] = VT(0);
                if (e < nst * NV) v[u] = src[(size_t)s_pid[e / NV] * NV + (e - (e / NV) * NV)];
            }
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int e = e0 + u * kPtThreads;
                if (e < nst * NV)
                    *reinterpret_cast<VT*>(stage + (size_t)(e / NV) * REC + OF + (e - (e / NV) * NV) * VW) = v[u];
            }
        }
        __syncthreads();
        accumulate(nst);
    } else {
    nst = 0;
    bool flushed = false;
    for (int c0 = 0; c0 < ncand; c0 += kPtThreads) {
        const int ci = c0 + (int)threadIdx.x;
        int is = -1, b3[3] = {0, 0, 0};
        bool take = false;
        if (ci < ncand) {
            // the last tile whose offset is <= ci (empty tiles skipped)
            int k = 0, hi = nbt - 1;
            while (k < hi) {
                const int mid = (k + hi + 1) >> 1;
                if (s_off[mid] <= ci) k = mid;
                else hi = mid - 1;
            }
            is = s_beg[k] + (ci - s_off[k]);
            b3[0] = base[3 * is]; b3[1] = base[3 * is + 1]; b3[2] = base[3 * is + 2];
            const int dx = wrapk(b3[0] - x0, K);
            const int dy = wrapk(b3[1] - y0, K);
            const int dz = wrapk(b3[2] - z0, K);
            take = (dx < xlen || K - dx < P) && (dy < 4 || K - dy < P) && (dz < 4 || K - dz < P);
        }
        int tot;
        const int slot = wg_compact(take, s_wave, &tot);
        for (int done = 0; done < tot;) {
            if (nst == cap) {
                __syncthreads();
                accumulate(nst);
                __syncthreads();
                nst = 0;
                flushed = true;
            }
            const int taken = min(cap - nst, tot - done);
            const int first = nst;
            if (take && slot >= done && slot < done + taken) {
                const int sl = first + slot - done;
                T* rec = stage + (size_t)sl * REC;
                T w[P];
                // base relative to the block origin, as an int in (-P, K - P]
                int* rl = reinterpret_cast<int*>(rec + OB);
                const int org[3] = {x0, y0, z0};
#pragma unroll
                for (int d = 0; d < 3; d++) {
                    bspline_weights_r<T, P>(frac[3 * is + d], w);
#pragma unroll
                    for (int k = 0; k < P; k++) rec[d * (P + 1) + k] = w[k];
                    rec[d * (P + 1) + P] = T(0);
                    const int r = wrapk(b3[d] - org[d], K);
                    rl[d] = r > K - P ? r - K : r;
                }
                s_pid[sl] = is;
            }
            __syncthreads();
            if constexpr (SGF) {
                // forces are read in the accumulation
            } else if constexpr (BLK) {
                // NQ contiguous values per slot, 16 bytes per load, the loads
                // of a thread in flight together
                constexpr int VW = 16 / sizeof(T);
                constexpr int NV = NQ / VW;
                constexpr int U = 4;
                typedef T VT __attribute__((ext_vector_type(16 / sizeof(T))));
                const VT* src = reinterpret_cast<const VT*>(Xb + (size_t)cbl * npos * NQ);
                for (int e0 = threadIdx.x; e0 < taken * NV; e0 += kPtThreads * U) {
                    VT v[U];
#pragma unroll
                    for (int u = 0; u < U; u++) {
                        const int e = e0 + u * kPtThreads;
                        v[u] = VT(0);
                        if (e < taken * NV) {
                            const int sl = first + e / NV;
                            v[u] = src[(size_t)s_pid[sl] * NV + (e - (e / NV) * NV)];
                        }
                    }
#pragma unroll
                    for (int u = 0; u < U; u++) {
                        const int e = e0 + u * kPtThreads;
                        if (e < taken * NV) {
                            const int sl = first + e / NV, vv = e - (e / NV) * NV;
                            *reinterpret_cast<VT*>(stage + (size_t)sl * REC + OF + vv * VW) = v[u];
                        }
                    }
                }
            } else {
                // forces of the new slots: slot q holds rhs r, component c
                // (mono: q = 3 r + c; poly: q = 6 r + c for f, 6 r + 3 + c
                // for a^2 f); the loads of a thread in flight together
                constexpr int U = 4;
                for (int e0 = threadIdx.x; e0 < taken * NQ; e0 += kPtThreads * U) {
                    T fv[U];
#pragma unroll
                    for (int u = 0; u < U; u++) {
                        const int idx = e0 + u * kPtThreads;
                        fv[u] = T(0);
                        if (idx < taken * NQ) {
                            const int sl = first + idx / NQ, q = idx % NQ;
                            const int p = s_pid[sl];
                            const int r = q / NC, cc = q - r * NC;
                            if (r < rb) {
                                const double x = X[(size_t)(3 * p + cc % 3) * b + r0 + r];
                                fv[u] = (T)((cc >= 3) ? x * a2[p] : x);
                            }
                        }
                    }
#pragma unroll
                    for (int u = 0; u < U; u++) {
                        const int idx = e0 + u * kPtThreads;
                        if (idx < taken * NQ)
                            stage[(size_t)(first + idx / NQ) * REC + OF + idx % NQ] = fv[u];
                    }
                }
            }
            nst += taken;
            done += taken;
        }
        __syncthreads();
    }
    __syncthreads();
    accumulate(nst);
    // the mesh tiles below reuse the stage
    staged_all = MA && !flushed && !tiled_out;
    }
    // Mesh stores. Default: each lane stores its own point (a wave writes
    // 16-byte pieces of 16 rows; the eight waves' pieces of a row reach HBM
    // about 1.4x). tiled_out: the meshes leave through LDS (the stage is free
    // now), QC at a time (up to 48 KB): each wave writes its 4 x 4 x 4
    // sub-block into a 16-row x 32 tile per mesh, then every 32 threads store
    // one whole x row (128 bytes for xlen = 32). Exactly the mesh bytes reach
    // HBM, but the extra LDS round trip and barriers cost about 5 % of the
    // kernel at K = 128 and 100 columns, so it is not the default.
    constexpr int QC = (NQ * 512 * (int)sizeof(T) <= 49152) ? NQ : NQ / 2;
    static_assert(NQ % QC == 0, "mesh groups must divide NQ");
    T* tile = stage;
    const size_t K3 = (size_t)K * K * K;
    if constexpr (MF) {
        // result register v of lane l holds row 4 (l / 16) + v of a tile:
        // x = v, y = l / 16 of z plane m; column l % 16 of component t
        if (MA && !tiled_out) {
            // through the LDS tiles: wave w writes x in [4w, 4w + 4) of every
            // (column, row); thread e then stores 16 bytes of row e / 8. The
            // rounds alternate between two tiles, so one barrier per round
            // suffices: a thread writes a tile again only after the next
            // round's barrier, which every thread reaches after its reads.
            const int jj = lane & 15, sy = lane >> 4;
            const int e = threadIdx.x;
            const int sj = e >> 5, sr = (e >> 3) & 3, sx = (e & 7) * 4;
#pragma unroll
            for (int m = 0; m < 4; m++)
#pragma unroll
                for (int t = 0; t < 3; t++) {
                    T* tb = ma_tile + ((m * 3 + t) & 1) * (16 * kMaTileStride);
                    if (wave_on)
                        *reinterpret_cast<f4*>(tb + jj * kMaTileStride + sy * 32 + 4 * wave) =
                            accm[m][t];
                    __syncthreads();
                    if (sj < rb && sx < xlen) {
                        const f4 v =
                            *reinterpret_cast<const f4*>(tb + sj * kMaTileStride + sr * 32 + sx);
                        *reinterpret_cast<f4*>(mesh + ((size_t)(r0 + sj) * 3 + t) * K3 +
                                               ((size_t)(z0 + m) * K + (y0 + sr)) * K + x0 + sx) = v;
                    }
                }
#pragma unroll
            for (int m = 0; m < 4; m++)
#pragma unroll
                for (int t = 0; t < 3; t++) accm[m][t] = f4{0.f, 0.f, 0.f, 0.f};
            if (++cb < ncb) goto next_column_block;
            return;
        }
        if (!tiled_out) {
            const int jj = lane & 15, sy = lane >> 4;
            if (wave_on && jj < rb) {
#pragma unroll
                for (int m = 0; m < 4; m++)
#pragma unroll
                    for (int t = 0; t < 3; t++) {
                        T* dst = mesh + ((size_t)(r0 + jj) * 3 + t) * K3 +
                                 ((size_t)(z0 + m) * K + (y0 + sy)) * K + x0 + 4 * wave;
                        *reinterpret_cast<f4*>(dst) = accm[m][t];
                    }
            }
            if constexpr (MA) {
#pragma unroll
                for (int m = 0; m < 4; m++)
#pragma unroll
                    for (int t = 0; t < 3; t++) accm[m][t] = f4{0.f, 0.f, 0.f, 0.f};
                if (++cb < ncb) goto next_column_block;
            }
            return;
        }
    }
    if constexpr (MA && !MF) {
        // vector-ALU sums (FP64): per-lane stores, then the next column block
        if (!tiled_out) {
            if (wave_on) {
                const size_t pt = ((size_t)pz * K + py) * K + px;
                constexpr int RBB = NQ / 3;
#pragma unroll
                for (int q = 0; q < NQ; q++) {
                    const int c = q / RBB, r = q % RBB;
                    if (r < rb) mesh[((size_t)(r0 + r) * 3 + c) * K3 + pt] = acc[q];
                }
            }
#pragma unroll
            for (int q = 0; q < NQ; q++) acc[q] = T(0);
            if (++cb < ncb) goto next_column_block;
            return;
        }
    }
    if (!tiled_out) {
        if (!wave_on) return;
        const size_t pt = ((size_t)pz * K + py) * K + px;
#pragma unroll
        for (int q = 0; q < NQ; q++) {
            size_t mi;
            bool on;
            if constexpr (BLK) {
                constexpr int RBB = NQ / 3;
                const int c = q / RBB, r = q % RBB;
                on = r < rb;
                mi = (size_t)(r0 + r) * 3 + c;
            } else {
                on = q < nq;
                mi = (size_t)r0 * NC + q;
            }
            if (on) mesh[mi * K3 + pt] = acc[q];
        }
        return;
    }
    __syncthreads();   // every wave is done reading the stage
#pragma unroll
    for (int q0 = 0; q0 < NQ; q0 += QC) {
        if (wave_on) {
            if constexpr (MF) {
                const int jj = lane & 15, sy = lane >> 4;
#pragma unroll
                for (int m = 0; m < 4; m++)
#pragma unroll
                    for (int t = 0; t < 3; t++) {
                        const int qq = 16 * t + jj - q0;
                        if (qq >= 0 && qq < QC) {
#pragma unroll
                            for (int v = 0; v < 4; v++)
                                tile[(qq * 16 + m * 4 + sy) * 32 + 4 * wave + v] = accm[m][t][v];
                        }
                    }
            } else {
#pragma unroll
                for (int qq = 0; qq < QC; qq++) tile[(qq * 16 + lz * 4 + ly) * 32 + lx] = acc[q0 + qq];
            }
        }
        __syncthreads();
        for (int e = threadIdx.x; e < QC * 512; e += kPtThreads) {
            const int x = e & 31, rw = (e >> 5) & 15, qq = e >> 9;
            const int q = q0 + qq;
            size_t mi;
            bool on;
            if constexpr (BLK) {
                constexpr int RBB = NQ / 3;
                const int c = q / RBB, r = q % RBB;
                on = r < rb;
                mi = (size_t)(r0 + r) * 3 + c;
            } else {
                on = q < nq;
                mi = (size_t)r0 * NC + q;
            }
            if (on && x < xlen)
                mesh[mi * K3 + ((size_t)(z0 + (rw >> 2)) * K + (y0 + (rw & 3))) * K + x0 + x] =
                    tile[e];
        }
        __syncthreads();
    }
    if constexpr (MA) {
        if constexpr (MF) {
#pragma unroll
            for (int m = 0; m < 4; m++)
#pragma unroll
                for (int t = 0; t < 3; t++) accm[m][t] = f4{0.f, 0.f, 0.f, 0.f};
        } else {
#pragma unroll
            for (int q = 0; q < NQ; q++) acc[q] = T(0);
        }
        if (++cb < ncb) goto next_column_block;
    }
    }
}

// ------------------------------------------------------ x transforms -------
// Real rows (2 per complex FFT: z = a + i b). Each workgroup takes NR rows.
// LDS: K twiddles, then two buffers of NR / 2 pencils of stride PS.
template <typename T>
__device__ __forceinline__ void load_twiddles(cplx<T>* tw, const double* __restrict__ tw_g, int K) {
    for (int i = threadIdx.x; i < K; i += blockDim.x)
        tw[i] = cplx<T>{(T)tw_g[2 * i], (T)tw_g[2 * i + 1]};
}

template <typename T>
__global__ void __launch_bounds__(256)
fft_x_r2c_kernel(FFTArgs fa, long long nrows, const T* __restrict__ real,
                 cplx<T>* __restrict__ spec, const double* __restrict__ tw_g, int NR) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int K = fa.K, Kh = K / 2 + 1, PS = pencil_stride(K);
    const int npen = NR / 2;
    cplx<T>* tw = reinterpret_cast<cplx<T>*>(smem_raw);
    cplx<T>* bufa = tw + K;
    cplx<T>* bufb = bufa + npen * PS;
    load_twiddles(tw, tw_g, K);
    const long long row0 = (long long)blockIdx.x * NR;
    // the loads of a thread (XB rows pairs at a time) are in flight together
    constexpr int XB = 8;
    for (int i0 = threadIdx.x; i0 < npen * K; i0 += XB * blockDim.x) {
        T va[XB], vb[XB];
#pragma unroll
        for (int u = 0; u < XB; u++) {
            const int i = i0 + u * (int)blockDim.x;
            const int pen = (int)fdiv(i, fa.K_m), x = i - pen * K;
            const long long ra = row0 + 2 * pen;
            va[u] = vb[u] = T(0);
            if (i < npen * K && ra < nrows) {
                va[u] = real[ra * K + x];
                vb[u] = real[(ra + 1) * K + x];
            }
        }
#pragma unroll
        for (int u = 0; u < XB; u++) {
            const int i = i0 + u * (int)blockDim.x;
            const int pen = (int)fdiv(i, fa.K_m), x = i - pen * K;
            if (i < npen * K) bufa[pen * PS + lpad(x)] = cplx<T>{va[u], vb[u]};
        }
    }
    __syncthreads();
    const cplx<T>* res = fft_pencils<T, -1>(bufa, bufb, npen, PS, fa, tw);
    for (int i = threadIdx.x; i < npen * Kh; i += blockDim.x) {
        const int pen = (int)fdiv(i, fa.Kh_m), k = i - pen * Kh;
        const long long ra = row0 + 2 * pen;
        if (ra >= nrows) continue;
        const cplx<T> zk = res[pen * PS + lpad(k)];
        const cplx<T> zm = res[pen * PS + lpad(k == 0 ? 0 : K - k)];
        const T h = T(0.5);
        spec[ra * Kh + k] = cplx<T>{h * (zk.x + zm.x), h * (zk.y - zm.y)};
        spec[(ra + 1) * Kh + k] = cplx<T>{h * (zk.y + zm.y), h * (zm.x - zk.x)};
    }
}

template <typename T>
__global__ void __launch_bounds__(256)
fft_x_c2r_kernel(FFTArgs fa, long long nrows, const cplx<T>* __restrict__ spec,
                 T* __restrict__ real, const double* __restrict__ tw_g, int NR) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int K = fa.K, Kh = K / 2 + 1, PS = pencil_stride(K);
    const int npen = NR / 2;
    cplx<T>* tw = reinterpret_cast<cplx<T>*>(smem_raw);
    cplx<T>* bufa = tw + K;
    cplx<T>* bufb = bufa + npen * PS;
    load_twiddles(tw, tw_g, K);
    const long long row0 = (long long)blockIdx.x * NR;
    // the loads of a thread (XB at a time) are in flight together
    constexpr int XB = 8;
    for (int i0 = threadIdx.x; i0 < npen * K; i0 += XB * blockDim.x) {
        cplx<T> A[XB], B[XB];
#pragma unroll
        for (int u = 0; u < XB; u++) {
            const int i = i0 + u * (int)blockDim.x;
            const int pen = (int)fdiv(i, fa.K_m), k = i - pen * K;
            const long long ra = row0 + 2 * pen;
            A[u] = B[u] = cplx<T>{T(0), T(0)};
            if (i < npen * K && ra < nrows) {
                const int kk = (k <= K / 2) ? k : K - k;
                A[u] = spec[ra * Kh + kk];
                B[u] = spec[(ra + 1) * Kh + kk];
            }
        }
#pragma unroll
        for (int u = 0; u < XB; u++) {
            const int i = i0 + u * (int)blockDim.x;
            const int pen = (int)fdiv(i, fa.K_m), k = i - pen * K;
            const long long ra = row0 + 2 * pen;
            cplx<T> z{T(0), T(0)};
            if (i < npen * K && ra < nrows) {
                const int kk = (k <= K / 2) ? k : K - k;
                cplx<T> a = A[u], bb = B[u];
                if (kk == 0 || 2 * kk == K) { a.y = T(0); bb.y = T(0); }
                if (k <= K / 2) z = cplx<T>{a.x - bb.y, a.y + bb.x};
                else z = cplx<T>{a.x + bb.y, bb.x - a.y};
            }
            if (i < npen * K) bufa[pen * PS + lpad(k)] = z;
        }
    }
    __syncthreads();
    const cplx<T>* res = fft_pencils<T, +1>(bufa, bufb, npen, PS, fa, tw);
    for (int i = threadIdx.x; i < npen * K; i += blockDim.x) {
        const int pen = (int)fdiv(i, fa.K_m), x = i - pen * K;
        const long long ra = row0 + 2 * pen;
        if (ra >= nrows) continue;
        const cplx<T> v = res[pen * PS + lpad(x)];
        real[ra * K + x] = v.x;
        real[(ra + 1) * K + x] = v.y;
    }
}

// ------------------------------------------------------- y transform -------
// In place along y (element stride Kh) for NP = 2^lnp consecutive kx of one
// (mesh, z) plane.
template <typename T, int SIGN>
__global__ void __launch_bounds__(256)
fft_y_kernel(FFTArgs fa, cplx<T>* __restrict__ spec, const double* __restrict__ tw_g, int lnp) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int K = fa.K, Kh = K / 2 + 1, PS = pencil_stride(K);
    const int NP = 1 << lnp;
    cplx<T>* tw = reinterpret_cast<cplx<T>*>(smem_raw);
    cplx<T>* bufa = tw + K;
    cplx<T>* bufb = bufa + NP * PS;
    load_twiddles(tw, tw_g, K);
    const int nkb = (Kh + NP - 1) >> lnp;
    const int kx0 = (int)(blockIdx.x % nkb) * NP;
    const long long plane = blockIdx.x / nkb;    // mesh * K + z
    cplx<T>* base = spec + plane * K * Kh;
    // the loads of a thread (YB at a time) are in flight together
    constexpr int YB = 8;
    for (int i0 = threadIdx.x; i0 < NP * K; i0 += YB * blockDim.x) {
        cplx<T> v[YB];
#pragma unroll
        for (int u = 0; u < YB; u++) {
            const int i = i0 + u * (int)blockDim.x;
            const int y = i >> lnp, j = i & (NP - 1);
            const int kx = kx0 + j;
            v[u] = (i < NP * K && kx < Kh) ? base[(size_t)y * Kh + kx] : cplx<T>{T(0), T(0)};
        }
#pragma unroll
        for (int u = 0; u < YB; u++) {
            const int i = i0 + u * (int)blockDim.x;
            const int y = i >> lnp, j = i & (NP - 1);
            if (i < NP * K) bufa[j * PS + lpad(y)] = v[u];
        }
    }
    __syncthreads();
    const cplx<T>* res = fft_pencils<T, SIGN>(bufa, bufb, NP, PS, fa, tw);
    for (int i = threadIdx.x; i < NP * K; i += blockDim.x) {
        const int y = i >> lnp, j = i & (NP - 1);
        const int kx = kx0 + j;
        if (kx < Kh) base[(size_t)y * Kh + kx] = res[j * PS + lpad(y)];
    }
}

// K = R1 R2 (both powers of two): the two-stage transform; otherwise the
// planned radix-8/4/2/3/5 stages
template <typename T, int SIGN, int MAXI2, int R1, int R2>
__device__ __forceinline__ void fft_any(cplx<T>* buf, int npen, int PS, const FFTArgs& fa,
                                        const cplx<T>* tw) {
    if constexpr (R1 > 0) fft2_pencils<T, R1, R2, SIGN, (R2 >= 16 ? 2 : 1)>(buf, npen, PS, tw);
    else fft_pencils_inplace<T, SIGN, MAXI2>(buf, npen, PS, fa, tw);
}

// Mesh side and divisions by K and K/2+1: compile-time when the two-stage
// transform fixes K = R1 R2, magic reciprocals otherwise
template <int R1, int R2>
struct KDiv {
    static constexpr int KC = R1 * R2;
    int K, Kh;
    unsigned K_m, Kh_m;
    __device__ __forceinline__ explicit KDiv(const FFTArgs& fa) {
        if constexpr (KC > 0) {
            K = KC;
            Kh = KC / 2 + 1;
        } else {
            K = fa.K;
            Kh = fa.K / 2 + 1;
        }
        K_m = fa.K_m;
        Kh_m = fa.Kh_m;
    }
    __device__ __forceinline__ int div_k(int i) const {
        if constexpr (KC > 0) return (int)((unsigned)i / (unsigned)KC);
        else return (int)fdiv((unsigned)i, K_m);
    }
    __device__ __forceinline__ int div_kh(int i) const {
        if constexpr (KC > 0) return (int)((unsigned)i / (unsigned)(KC / 2 + 1));
        else return (int)fdiv((unsigned)i, Kh_m);
    }
};

// ------------------------------------------------ fused x + y plane -------
// One workgroup per z-plane of one mesh: the K x K real plane is read once,
// transformed along x (two real rows per complex FFT) and along y in LDS
// (in place, padded pencils), and the K x (K/2+1) half-spectrum plane is
// written once. The inverse kernel mirrors it. Used when one plane fits LDS.
// Register bounds are template parameters: MAXI2 radix-2 butterflies and
// MAXP pack/unpack items per thread (K <= 128: 5 and 8; K <= 180: 8 and 16).
constexpr int kPlaneThreads = 1024;

template <typename T, int MAXI2, int MAXP, int R1, int R2>
__global__ void __launch_bounds__(kPlaneThreads)
fft_plane_fwd_kernel(FFTArgs fa, const T* __restrict__ real, cplx<T>* __restrict__ spec,
                     const double* __restrict__ tw_g) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const KDiv<R1, R2> kd(fa);
    const int K = kd.K, Kh = kd.Kh, PS = pencil_stride(K), H = K / 2;
    cplx<T>* tw = reinterpret_cast<cplx<T>*>(smem_raw);
    cplx<T>* buf = tw + K;
    load_twiddles(tw, tw_g, K);
    const size_t plane = blockIdx.x;
    const T* src = real + plane * K * K;
    // row pairs (2j, 2j+1) -> pencil j; the loads of a thread (8 at a time)
    // are in flight together
    for (int i0 = threadIdx.x; i0 < H * K; i0 += 8 * kPlaneThreads) {
        cplx<T> v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int i = min(i0 + u * kPlaneThreads, H * K - 1);
            const int j = kd.div_k(i), x = i - j * K;
            v[u] = cplx<T>{src[(size_t)(2 * j) * K + x], src[(size_t)(2 * j + 1) * K + x]};
        }
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int i = i0 + u * kPlaneThreads;
            const int j = kd.div_k(i), x = i - j * K;
            if (i < H * K) buf[j * PS + lpad(x)] = v[u];
        }
    }
    __syncthreads();
    fft_any<T, -1, MAXI2, R1, R2>(buf, H, PS, fa, tw);
    // unpack Z_j = A + i B into the half spectra of rows 2j and 2j+1, stored
    // as y-pencils (pencil kx, element y)
    {
        cplx<T> va[MAXP], vb[MAXP];
        const int nit = H * Kh;
#pragma unroll
        for (int u = 0; u < MAXP; u++) {
            const int i = threadIdx.x + u * blockDim.x;
            if (i < nit) {
                const int j = kd.div_kh(i), k = i - j * Kh;
                const cplx<T> zk = buf[j * PS + lpad(k)];
                const cplx<T> zm = buf[j * PS + lpad(k == 0 ? 0 : K - k)];
                const T h = T(0.5);
                va[u] = cplx<T>{h * (zk.x + zm.x), h * (zk.y - zm.y)};
                vb[u] = cplx<T>{h * (zk.y + zm.y), h * (zm.x - zk.x)};
            }
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < MAXP; u++) {
            const int i = threadIdx.x + u * blockDim.x;
            if (i < nit) {
                const int j = kd.div_kh(i), k = i - j * Kh;
                buf[k * PS + lpad(2 * j)] = va[u];
                buf[k * PS + lpad(2 * j + 1)] = vb[u];
            }
        }
        __syncthreads();
    }
    fft_any<T, -1, MAXI2, R1, R2>(buf, Kh, PS, fa, tw);
    cplx<T>* dst = spec + plane * K * Kh;
    for (int i = threadIdx.x; i < K * Kh; i += blockDim.x) {
        const int y = kd.div_kh(i), kx = i - y * Kh;
        dst[i] = buf[kx * PS + lpad(y)];
    }
}

template <typename T, int MAXI2, int MAXP, int R1, int R2>
__global__ void __launch_bounds__(kPlaneThreads)
fft_plane_inv_kernel(FFTArgs fa, const cplx<T>* __restrict__ spec, T* __restrict__ real,
                     const double* __restrict__ tw_g) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const KDiv<R1, R2> kd(fa);
    const int K = kd.K, Kh = kd.Kh, PS = pencil_stride(K), H = K / 2;
    cplx<T>* tw = reinterpret_cast<cplx<T>*>(smem_raw);
    cplx<T>* buf = tw + K;
    load_twiddles(tw, tw_g, K);
    const size_t plane = blockIdx.x;
    const cplx<T>* src = spec + plane * K * Kh;
    // the loads of a thread (8 at a time) are in flight together
    for (int i0 = threadIdx.x; i0 < K * Kh; i0 += 8 * kPlaneThreads) {
        cplx<T> v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) v[u] = src[min(i0 + u * kPlaneThreads, K * Kh - 1)];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int i = i0 + u * kPlaneThreads;
            const int y = kd.div_kh(i), kx = i - y * Kh;
            if (i < K * Kh) buf[kx * PS + lpad(y)] = v[u];
        }
    }
    __syncthreads();
    fft_any<T, +1, MAXI2, R1, R2>(buf, Kh, PS, fa, tw);
    // pack the half spectra of rows 2j, 2j+1 into Z_j = A + i B (full x):
    // the pair A, B at kx = kk gives Z_j[kk] and Z_j[K - kk]
    {
        cplx<T> vz[MAXP], vw[MAXP];
        const int nit = H * Kh;
#pragma unroll
        for (int u = 0; u < MAXP; u++) {
            const int i = threadIdx.x + u * blockDim.x;
            if (i < nit) {
                const int j = kd.div_kh(i), kk = i - j * Kh;
                cplx<T> A = buf[kk * PS + lpad(2 * j)];
                cplx<T> B = buf[kk * PS + lpad(2 * j + 1)];
                if (kk == 0 || 2 * kk == K) { A.y = T(0); B.y = T(0); }
                vz[u] = cplx<T>{A.x - B.y, A.y + B.x};
                vw[u] = cplx<T>{A.x + B.y, B.x - A.y};
            }
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < MAXP; u++) {
            const int i = threadIdx.x + u * blockDim.x;
            if (i < nit) {
                const int j = kd.div_kh(i), kk = i - j * Kh;
                buf[j * PS + lpad(kk)] = vz[u];
                if (kk > 0 && 2 * kk < K) buf[j * PS + lpad(K - kk)] = vw[u];
            }
        }
        __syncthreads();
    }
    fft_any<T, +1, MAXI2, R1, R2>(buf, H, PS, fa, tw);
    T* dst = real + plane * K * K;
    for (int i = threadIdx.x; i < H * K; i += blockDim.x) {
        const int j = kd.div_k(i), x = i - j * K;
        const cplx<T> v = buf[j * PS + lpad(x)];
        dst[(size_t)(2 * j) * K + x] = v.x;
        dst[(size_t)(2 * j + 1) * K + x] = v.y;
    }
}

// ------------------------------------------ z transform + influence -------
// A workgroup takes NP = 2^lnp consecutive (y, kx) positions of the flat
// K x (K/2+1) half-plane index j = y (K/2+1) + kx (contiguous in memory, so
// every z-plane read is one NP-element run) for one right-hand side: forward
// z-FFT of all components in place, the RPY influence tensor
// (src/mob/spme.cc:201-310, spme.h:114-123), inverse z-FFT, store.
// Mono (NC = 3): B = lm2 (1 - k^2/3)(I - k k^T / k^2) applied to the
// 3-vector. Poly (NC = 6, spreads of f and a^2 f): the reference's three
// mesh round trips (flags 1, 2, 2) reduce to two outputs
//   U = G (F f - k^2/6 F[a^2 f]),  W = -k^2/6 G F f,  G = lm2 (I - k k^T / k^2)
// interpolated as interp(U) + a_i^2 interp(W).
constexpr int kZMaxI2 = 8;

template <typename T, int R1, int R2, int NT>
__global__ void __launch_bounds__(NT)
fft_z_influence_kernel(FFTArgs fa, cplx<T>* __restrict__ spec, const double* __restrict__ tw_g,
                       const double* __restrict__ bsp2, int lnp, int NC, double box, double xi) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const KDiv<R1, R2> kd(fa);
    const int K = kd.K, Kh = kd.Kh, PS = pencil_stride(K);
    const int NP = 1 << lnp;
    const int nj = K * Kh;
    cplx<T>* tw = reinterpret_cast<cplx<T>*>(smem_raw);
    cplx<T>* buf = tw + K;
    load_twiddles(tw, tw_g, K);
    // |b|^2 factors in arithmetic T, after the spectrum buffer
    T* s_b2 = reinterpret_cast<T*>(buf + (size_t)NC * NP * PS);
    for (int i = threadIdx.x; i < K; i += blockDim.x) s_b2[i] = (T)bsp2[i];
    const int nch = (nj + NP - 1) >> lnp;
    const int j0 = (int)(blockIdx.x % nch) * NP;
    const int rhs = (int)(blockIdx.x / nch);
    const size_t mesh_sz = (size_t)nj * K;
    cplx<T>* base = spec + (size_t)rhs * NC * mesh_sz + j0;
    const int nvalid = min(NP, nj - j0);
    // the loads of a thread (ZB at a time) are in flight together
    constexpr int ZB = 8;
    for (int i0 = threadIdx.x; i0 < NC * NP * K; i0 += ZB * NT) {
        cplx<T> v[ZB];
#pragma unroll
        for (int u = 0; u < ZB; u++) {
            const int i = i0 + u * NT;
            const int jl = i & (NP - 1);
            const int cz = i >> lnp;
            const int c = kd.div_k(cz), z = cz - c * K;
            v[u] = (i < NC * NP * K && jl < nvalid) ? base[c * mesh_sz + (size_t)z * nj + jl]
                                                    : cplx<T>{T(0), T(0)};
        }
#pragma unroll
        for (int u = 0; u < ZB; u++) {
            const int i = i0 + u * NT;
            const int jl = i & (NP - 1);
            const int cz = i >> lnp;
            const int c = kd.div_k(cz), z = cz - c * K;
            if (i < NC * NP * K) buf[(c * NP + jl) * PS + lpad(z)] = v[u];
        }
    }
    __syncthreads();
    fft_any<T, -1, kZMaxI2, R1, R2>(buf, NC * NP, PS, fa, tw);
    // influence
    const T twopi_l = (T)(2.0 * M_PI / box);
    const T inv_xi2 = (T)(1.0 / (xi * xi));
    const int cs = NP * PS;    // component stride in LDS
    for (int i = threadIdx.x; i < NP * K; i += blockDim.x) {
        const int jl = kd.div_k(i), z = i - jl * K;
        if (jl >= nvalid) continue;
        const int j = j0 + jl;
        const int y = kd.div_kh(j), kxi = j - y * Kh;
        const int my = (y <= K / 2) ? y : y - K;
        const int mz = (z <= K / 2) ? z : z - K;
        const T kx = twopi_l * (T)kxi;
        const T ky = twopi_l * (T)my;
        const T kz = twopi_l * (T)mz;
        const T kk = kx * kx + ky * ky + kz * kz;
        T m2 = T(0), ex = T(0), ey = T(0), ez = T(0);
        if (kk > T(0)) {
            const T kx2 = kk * inv_xi2;
            m2 = T(6.0 * M_PI) * exp(-kx2 * T(0.25)) / kk *
                 (T(1) + kx2 * (T(0.25) + kx2 * T(0.125)));
            m2 *= s_b2[kxi] * s_b2[y] * s_b2[z];
            const T rk = T(1) / sqrt(kk);
            ex = kx * rk; ey = ky * rk; ez = kz * rk;
        }
        const int o = jl * PS + lpad(z);
        if (NC == 3) {
            const T s = m2 * (T(1) - kk / T(3));
            cplx<T> g0 = buf[o], g1 = buf[cs + o], g2 = buf[2 * cs + o];
            // (I - e e^T) g, scaled
            const T dr = ex * g0.x + ey * g1.x + ez * g2.x;
            const T di = ex * g0.y + ey * g1.y + ez * g2.y;
            buf[o] = cplx<T>{s * (g0.x - ex * dr), s * (g0.y - ex * di)};
            buf[cs + o] = cplx<T>{s * (g1.x - ey * dr), s * (g1.y - ey * di)};
            buf[2 * cs + o] = cplx<T>{s * (g2.x - ez * dr), s * (g2.y - ez * di)};
        } else {
            const T q = -kk / T(6);
            cplx<T> f[3], a[3];
#pragma unroll
            for (int c = 0; c < 3; c++) {
                f[c] = buf[c * cs + o];
                a[c] = buf[(3 + c) * cs + o];
            }
            cplx<T> u[3], w[3];
#pragma unroll
            for (int c = 0; c < 3; c++) {
                u[c] = cplx<T>{f[c].x + q * a[c].x, f[c].y + q * a[c].y};
                w[c] = cplx<T>{q * f[c].x, q * f[c].y};
            }
            const T ur = ex * u[0].x + ey * u[1].x + ez * u[2].x;
            const T ui = ex * u[0].y + ey * u[1].y + ez * u[2].y;
            const T wr = ex * w[0].x + ey * w[1].x + ez * w[2].x;
            const T wi = ex * w[0].y + ey * w[1].y + ez * w[2].y;
            const T e3[3] = {ex, ey, ez};
#pragma unroll
            for (int c = 0; c < 3; c++) {
                buf[c * cs + o] = cplx<T>{m2 * (u[c].x - e3[c] * ur), m2 * (u[c].y - e3[c] * ui)};
                buf[(3 + c) * cs + o] =
                    cplx<T>{m2 * (w[c].x - e3[c] * wr), m2 * (w[c].y - e3[c] * wi)};
            }
        }
    }
    __syncthreads();
    fft_any<T, +1, kZMaxI2, R1, R2>(buf, NC * NP, PS, fa, tw);
    for (int i = threadIdx.x; i < NC * NP * K; i += blockDim.x) {
        const int jl = i & (NP - 1);
        const int cz = i >> lnp;
        const int c = kd.div_k(cz), z = cz - c * K;
        if (jl < nvalid) base[c * mesh_sz + (size_t)z * nj + jl] = buf[(c * NP + jl) * PS + lpad(z)];
    }
}

// ------------------------------------------------------------- interp ------
// A workgroup takes the particles whose base points lie in a strip of sxt
// tiles along x (ts x ts x ts each) for RB right-hand sides, i.e. the
// consecutive meshes r0 NC .. (r0 + rb) NC - 1. The meshes go through LDS MV
// at a time (one MV-vector per mesh point): each pass
// stages the region their stencils cover, (sxt ts + P - 1) x (ts + P - 1) x
// (ts + P - 1) points, while the next pass's region is loaded into
// registers. Items (particle, z line) sum their P x P stencil slice of all MV
// meshes with vector LDS reads, and each particle adds its P slices in fixed
// order (src/mob/spme.h:219-265 without the stored P/ind tables).
// Y = beta Y + alpha (sum_U + a^2 sum_W).
template <typename T, int MV>
struct InterpVec {
    typedef T type __attribute__((ext_vector_type(MV)));
};

template <typename T, int P, int RREG, int MV>
__global__ void __launch_bounds__(256)
interp_tile_kernel(int K, int ts, int ntile, int nsx, int sxt, int ch,
                   const int* __restrict__ tile_start, const double* __restrict__ frac,
                   const int* __restrict__ base, const double* __restrict__ a2,
                   const T* __restrict__ mesh, int b, int RB, int NC, double alpha, double beta,
                   double* __restrict__ Y) {
    typedef typename InterpVec<T, MV>::type VT;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const unsigned lid = xcd_logical_id(blockIdx.x + blockIdx.y * gridDim.x, gridDim.x * gridDim.y);
    const int bxl = (int)(lid % gridDim.x), byl = (int)(lid / gridDim.x);
    const int sxi = bxl % nsx;
    const int tyz = bxl / nsx;
    const int ty = tyz % ntile, tz = tyz / ntile;
    const int r0 = byl * RB;
    const int rb = min(RB, b - r0);
    const int nmesh = rb * NC;
    const int stx = sxi * sxt;
    const int nxt = min(sxt, ntile - stx);
    const int sx0 = stx * ts, sy0 = ty * ts, sz0 = tz * ts;
    const int RX = nxt * ts + P - 1, RY = ts + P - 1, RZ = ts + P - 1;
    const int nreg = RX * RY * RZ;
    // LDS: region VT[RZ][RY][RX] | partial sums VT[ch][P] | weights T[ch][3][P] |
    //      bases int[ch][3] | results double[ch][RB][3]
    VT* region = reinterpret_cast<VT*>(smem_raw);
    VT* s_part = region + (size_t)(sxt * ts + P - 1) * RY * RZ;
    T* s_w = reinterpret_cast<T*>(s_part + (size_t)ch * P);
    int* s_b = reinterpret_cast<int*>(s_w + ch * 3 * P);
    double* s_res = reinterpret_cast<double*>(
        (reinterpret_cast<uintptr_t>(s_b + ch * 3) + 7) & ~(uintptr_t)7);
    const size_t K3 = (size_t)K * K * K;
    // byte offsets within one mesh of this thread's region points (the same
    // for every mesh; 32-bit, added to the wave-uniform mesh base): point u is
    // region index threadIdx.x + 256 u
    unsigned offb[RREG];
    {
        const unsigned rx_m = 0xffffffffu / (unsigned)RX + 1u;
        const unsigned ry_m = 0xffffffffu / (unsigned)RY + 1u;
#pragma unroll
        for (int u = 0; u < RREG; u++) {
            const int i = min((int)threadIdx.x + 256 * u, nreg - 1);
            const int row = (int)fdiv(i, rx_m), lx = i - row * RX;
            const int lz = (int)fdiv(row, ry_m), ly = row - lz * RY;
            offb[u] = (unsigned)((wrapk(sz0 + lz, K) * K + wrapk(sy0 + ly, K)) * K +
                                 wrapk(sx0 + lx, K)) * (unsigned)sizeof(T);
        }
    }
    const unsigned char* mbase =
        reinterpret_cast<const unsigned char*>(mesh + (size_t)r0 * NC * K3);
    VT v[RREG];
    auto prefetch = [&](int pi) __attribute__((always_inline)) {
#pragma unroll
        for (int j = 0; j < MV; j++) {
            const int q = pi * MV + j;
            const unsigned char* m = mbase + (size_t)min(q, nmesh - 1) * K3 * sizeof(T);
            const bool on = q < nmesh;
#pragma unroll
            for (int u = 0; u < RREG; u++) {
                const T x = *reinterpret_cast<const T*>(m + offb[u]);
                v[u][j] = on ? x : T(0);
            }
        }
    };
    // own particles: base tiles stx .. stx+nxt-1 of row (ty, tz) are contiguous
    const int t_first = (tz * ntile + ty) * ntile + stx;
    const int pb = tile_start[t_first], pe = tile_start[t_first + nxt];
    const int npass = (nmesh + MV - 1) / MV;
    for (int c0 = pb; c0 < pe; c0 += ch) {
        const int cnt = min(ch, pe - c0);
        prefetch(0);
        for (int t = threadIdx.x; t < cnt; t += blockDim.x) {
            const int is = c0 + t;
            T w[P];
#pragma unroll
            for (int d = 0; d < 3; d++) {
                bspline_weights_r<T, P>(frac[3 * is + d], w);
#pragma unroll
                for (int k = 0; k < P; k++) s_w[(t * 3 + d) * P + k] = w[k];
            }
            // stencil origin relative to the region
            s_b[t * 3 + 0] = wrapk(base[3 * is + 0] - sx0, K);
            s_b[t * 3 + 1] = wrapk(base[3 * is + 1] - sy0, K);
            s_b[t * 3 + 2] = wrapk(base[3 * is + 2] - sz0, K);
            for (int q = 0; q < rb * 3; q++) s_res[t * RB * 3 + q] = 0.0;
        }
        for (int pi = 0; pi < npass; pi++) {
            __syncthreads();          // the previous pass no longer reads the region
#pragma unroll
            for (int u = 0; u < RREG; u++) {
                const int i = (int)threadIdx.x + 256 * u;
                if (i < nreg) region[i] = v[u];
            }
            __syncthreads();
            if (pi + 1 < npass) prefetch(pi + 1);
            for (int item = threadIdx.x; item < cnt * P; item += blockDim.x) {
                const int t = item / P, iz = item - t * P;
                const T* wx = s_w + (t * 3 + 0) * P;
                const T* wy = s_w + (t * 3 + 1) * P;
                const VT* rp = region + ((size_t)(s_b[t * 3 + 2] + iz) * RY + s_b[t * 3 + 1]) * RX +
                               s_b[t * 3 + 0];
                VT acc = VT(0);
                if constexpr (MV == 1) {
                    // rows in pairs: their reads are issued ahead of the sums
#pragma unroll 2
                    for (int iy = 0; iy < P; iy++) {
                        VT row = VT(0);
#pragma unroll
                        for (int ix = 0; ix < P; ix++) row += wx[ix] * rp[iy * RX + ix];
                        acc += wy[iy] * row;
                    }
                } else {
                    // one stencil row at a time keeps the live vector reads to P
#pragma unroll 1
                    for (int iy = 0; iy < P; iy++) {
                        VT row = VT(0);
#pragma unroll
                        for (int ix = 0; ix < P; ix++) row += wx[ix] * rp[iy * RX + ix];
                        acc += wy[iy] * row;
                    }
                }
                s_part[t * P + iz] = acc * s_w[(t * 3 + 2) * P + iz];
            }
            __syncthreads();
            for (int t = threadIdx.x; t < cnt; t += blockDim.x) {
                VT sum = VT(0);
#pragma unroll
                for (int iz = 0; iz < P; iz++) sum += s_part[t * P + iz];
#pragma unroll
                for (int j = 0; j < MV; j++) {
                    const int q = pi * MV + j;
                    if (q < nmesh) {
                        const int r = q / NC, c = q - r * NC;
                        double vv = (double)sum[j];
                        if (c >= 3) vv *= a2[c0 + t];
                        s_res[t * RB * 3 + r * 3 + (c % 3)] += vv;
                    }
                }
            }
        }
        __syncthreads();
        for (int i = threadIdx.x; i < cnt * rb * 3; i += blockDim.x) {
            const int t = i / (rb * 3), rc = i - t * rb * 3;
            const int r = rc / 3, c = rc - r * 3;
            double* y = &Y[(size_t)(3 * (c0 + t) + c) * b + r0 + r];
            *y = (beta == 0.0 ? 0.0 : beta * *y) + alpha * s_res[t * RB * 3 + r * 3 + c];
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------ host side ----
void bspline_moduli(int K, int P, std::vector<double>& out) {
    // |b(m)|^2 = 1 / |sum_{i<P-1} M_P(i+1) e^{2 pi i m i / K}|^2
    // (src/mob/spme.cc:80-128; Essmann et al. 1995)
    std::vector<double> sv(P - 1);
    for (int i = 0; i < P - 1; i++) {
        // M_P(i+1) by Cox-de Boor on integer knots
        std::vector<double> a(P + 1, 0.0);
        a[0] = 1.0;  // M_1(u + k) with u = 0 -> use u = 1 piece boundary
        double u = 1.0;
        for (int n = 2; n <= P; n++) {
            for (int k = n - 1; k >= 1; k--)
                a[k] = ((u + k) * a[k] + (n - u - k) * a[k - 1]) / (n - 1);
            a[0] = u * a[0] / (n - 1);
        }
        sv[i] = a[i];   // M_P(1 + i)
    }
    std::vector<double> den2(K);
    for (int m = 0; m < K; m++) {
        int mm = (m <= K / 2) ? m : m - K;
        double re = 0.0, im = 0.0;
        for (int i = 0; i < P - 1; i++) {
            double ph = 2.0 * M_PI * mm * i / K;
            re += sv[i] * cos(ph);
            im += sv[i] * sin(ph);
        }
        den2[m] = re * re + im * im;
    }
    // odd P: the sum vanishes at the Nyquist index (M_P is symmetric), and
    // rounding leaves |den|^2 ~ 1e-33; take the mean of the neighbours there
    // instead (Essmann et al. 1995). Even P is unaffected (|den|^2 > 1e-3).
    for (int m = 0; m < K; m++)
        if (den2[m] < 1e-7) den2[m] = 0.5 * (den2[(m + K - 1) % K] + den2[(m + 1) % K]);
    out.resize(K);
    for (int m = 0; m < K; m++) out[m] = 1.0 / den2[m];
}

bool make_fft_plan(int K, FFTPlan* plan, int order) {
    plan->K = K;
    plan->nst = 0;
    int m = K;
    auto take = [&](int r) {
        while (m % r == 0 && plan->nst < 16) { plan->radix[plan->nst++] = r; m /= r; }
    };
    if (order == 2) { take(3); take(5); }
    if (order != 1) take(8);
    take(4);
    take(2);
    take(3);
    take(5);
    return m == 1;
}

template <typename T>
void SpmeGpu<T>::set_tuning(const SpmeTuning& t, int width) {
    FFTPlan p;
    SDT_REQUIRE(t.fft_order >= 0 && t.fft_order <= 2 && make_fft_plan(K, &p, t.fft_order),
                SDT_EINVAL, "tuning: invalid FFT stage order");
    SDT_REQUIRE(t.spread_stage_kb == 0 || (t.spread_stage_kb >= 16 && t.spread_stage_kb <= 64),
                SDT_EINVAL, "tuning: spread stage must be 16 .. 64 KB");
    SDT_REQUIRE(t.interp_sxt == 0 || t.interp_sxt == 1 || t.interp_sxt == 2 || t.interp_sxt == 4,
                SDT_EINVAL, "tuning: interp strip width must be 1, 2 or 4 tiles");
    SDT_REQUIRE(t.interp_ch == 0 || (t.interp_ch >= 32 && t.interp_ch <= 128), SDT_EINVAL,
                "tuning: interp chunk must be 32 .. 128");
    SDT_REQUIRE(width == 0 || width == 1, SDT_EINVAL, "tuning: width class must be 0 or 1");
    plans[width] = p;
    tune[width] = t;
}

static FFTArgs fft_args(const FFTPlan& p) {
    FFTArgs a;
    memset(&a, 0, sizeof(a));
    a.K = p.K;
    a.nst = p.nst;
    a.K_m = fft_magic(p.K);
    a.Kh_m = fft_magic(p.K / 2 + 1);
    int len = p.K;
    for (int i = 0; i < p.nst; i++) {
        const int r = p.radix[i];
        a.radix[i] = r;
        a.per[i] = p.K / r;
        a.per_m[i] = fft_magic(a.per[i]);
        a.str[i] = p.K / len;
        a.str_m[i] = fft_magic(a.str[i]);
        len /= r;
    }
    return a;
}

template <typename T>
void SpmeGpu<T>::create(int npos_, double box_, double xi_, int K_, int porder_, bool poly_) {
    npos = npos_; box = box_; xi = xi_; K = K_; porder = porder_; poly = poly_;
    // narrower than the reference (porder > 1 with 4 porder <= dim,
    // src/mob/mob_spme.cc:79-90): the kernels are instantiated for these
    SDT_REQUIRE(porder >= 4 && porder <= 8, SDT_EINVAL,
                "porder must be in [4, 8] on the GPU path (the reference accepts any porder > 1 "
                "with 4 porder <= dim)");
    SDT_REQUIRE(K % 2 == 0, SDT_EINVAL,
                "mesh dimension must be even on the GPU path (the reference's odd-dim wavenumber "
                "map is incomplete)");
    SDT_REQUIRE(make_fft_plan(K, &plan), SDT_EINVAL,
                "mesh dimension must factor into 2, 3 and 5 on the GPU path");
    SDT_REQUIRE(K <= 1024, SDT_EINVAL, "mesh dimension must be <= 1024");
    plans[0] = plans[1] = plan;
    tune[0] = tune[1] = SpmeTuning();
    // tile side: a divisor of K in [P-1, 8], preferring 8 (one wave owns
    // ts x ts mesh rows)
    ts = 0;
    const int pref[] = {8, 6, 7, 5, 4};
    for (int c : pref)
        if (c >= porder - 1 && K % c == 0 && K / c >= 3) { ts = c; break; }
    SDT_REQUIRE(ts > 0, SDT_EINVAL, "no spread tile size divides the mesh dimension");
    ntile = K / ts;
    std::vector<double> tw(2 * K);
    for (int j = 0; j < K; j++) {
        tw[2 * j] = cos(-2.0 * M_PI * j / K);
        tw[2 * j + 1] = sin(-2.0 * M_PI * j / K);
    }
    d_tw.reserve(2 * K);
    SDT_HIP_CHECK(sync_memcpy(d_tw.get(), tw.data(), sizeof(double) * 2 * K, hipMemcpyHostToDevice));
    std::vector<double> bs;
    bspline_moduli(K, porder, bs);
    d_bsp2.reserve(K);
    SDT_HIP_CHECK(sync_memcpy(d_bsp2.get(), bs.data(), sizeof(double) * K, hipMemcpyHostToDevice));
    d_frac.reserve(3 * (size_t)npos);
    d_base.reserve(3 * (size_t)npos);
    d_frac_u.reserve(3 * (size_t)npos);
    d_base_u.reserve(3 * (size_t)npos);
    d_a2.reserve(npos);
    d_tile_start.reserve((size_t)ntile * ntile * ntile + 1);
}

template <typename T>
void SpmeGpu<T>::reserve_rhs(int b) {
    if (b <= bcap) return;
    const size_t K3 = (size_t)K * K * K;
    const size_t Ksp = (size_t)K * K * (K / 2 + 1);
    d_real.reserve((size_t)nmesh(b) * K3);
    d_spec.reserve((size_t)nmesh(b) * Ksp);
    bcap = b;
}

template <typename T>
void SpmeGpu<T>::prepare(const double* d_pos, int* d_key, hipStream_t s) {
    const unsigned g = ceil_div(npos, 256);
#define SDT_PREP(PP)                                                                 \
    case PP:                                                                         \
        hipLaunchKernelGGL((spme_prepare_kernel<T, PP>), dim3(g), dim3(256), 0, s, npos, \
                           d_pos, box, K, ts, ntile, d_frac_u.get(), d_base_u.get(), d_key); \
        break;
    switch (porder) { SDT_PREP(4) SDT_PREP(5) SDT_PREP(6) SDT_PREP(7) SDT_PREP(8) }
#undef SDT_PREP
    SDT_HIP_CHECK(hipGetLastError());
}

template <typename T>
void SpmeGpu<T>::gather_sorted(const int* d_perm, const double* d_pos, const double* d_rdi,
                               const int* d_tile_start_in, hipStream_t s) {
    (void)d_pos;
    hipLaunchKernelGGL((spme_gather_kernel<T>), dim3(ceil_div(npos, 256)), dim3(256), 0, s, npos,
                       d_perm, d_frac_u.get(), d_base_u.get(), d_rdi, d_frac.get(), d_base.get(),
                       poly ? d_a2.get() : (double*)nullptr);
    SDT_HIP_CHECK(hipGetLastError());
    const size_t nt3 = (size_t)ntile * ntile * ntile + 1;
    SDT_HIP_CHECK(hipMemcpyAsync(d_tile_start.get(), d_tile_start_in, sizeof(int) * nt3,
                                 hipMemcpyDeviceToDevice, s));
}

template <typename T>
bool SpmeGpu<T>::point_owner() const {
    // contributor bucket tiles of a 32 x 4 x 4 block: bases in intervals of
    // length min(32, K) + P - 1 (x) and P + 3 (y, z)
    const int pt_tiles = std::min(ntile, (std::min(32, K) + porder - 1 + ts - 1) / ts + 1) *
                         std::min(ntile, (porder + 3 + ts - 1) / ts + 1) *
                         std::min(ntile, (porder + 3 + ts - 1) / ts + 1);
    return K % 4 == 0 && K >= 16 && pt_tiles <= kPtMaxTiles;
}

// point-owner meshes per workgroup: 12 when few are needed
template <typename T>
static int pt_meshes(int b, int NC) {
    if (b * NC <= 4) return 4;
    return (b * NC <= 12) ? 12 : (sizeof(T) == 4 ? 48 : 24);
}

template <typename T>
int SpmeGpu<T>::blocked_cols(int b) const {
    if (poly || !point_owner()) return 0;
    const int NQ = pt_meshes<T>(b, 3);
    return NQ <= 12 ? 0 : NQ / 3;
}

template <typename T>
void SpmeGpu<T>::spread(int b, const double* X, hipStream_t s, const T* xb) {
    SDT_REQUIRE(b <= bcap, SDT_ECAP, "spread: rhs count exceeds mesh capacity");
    const int NC = poly ? 6 : 3;
    SDT_REQUIRE(xb == nullptr || blocked_cols(b) > 0, SDT_EINVAL,
                "spread: no blocked input for this configuration");
    if (point_owner()) {
        const SpmeTuning& tn = tune[width_class(b)];
        constexpr int NQL = sizeof(T) == 4 ? 48 : 24;
        const int NQ = pt_meshes<T>(b, NC);
        // Wide FP32 launches on the blocked copy sum on the matrix cores, one
        // workgroup per mesh block over every column block (K = 128, 100
        // columns: 1.08 ms against 2.17 ms for the vector-ALU sums with
        // scalar force reads, the same sums to the bit); wide FP64 launches
        // keep the vector-ALU sums over every column block in turn (2.85
        // against 3.97 ms); where the stencil can wrap, FP32 reads the
        // blocked forces through the scalar path and FP64 stages them in
        // LDS per column block. Forms override this for tests.
        const int fr = spread_form & 3;
        const bool mf_ok = xb != nullptr && sizeof(T) == 4 && NQ == NQL && K >= 32 + porder - 1;
        const bool mf = mf_ok && (fr == 0 || fr == 3);
        const bool ma64 = sizeof(T) == 8 && xb != nullptr && NQ == NQL && K >= 32 + porder - 1 &&
                          (fr == 0 || (spread_form & 8) != 0);
        const bool ma = (mf && (fr == 0 || (spread_form & 8) != 0)) || ma64;
        const bool sgf = xb != nullptr && sizeof(T) == 4 && !mf && fr != 1;
        const int rec = ((3 * (porder + 1) + 3 + 3) & ~3) + (sgf ? 0 : NQ);
        // few meshes: a 48 KB stage leaves room for three workgroups per CU
        // 60 KB for the matrix-core form over every column block, whose two
        // output tiles (16.5 KB) also fit two workgroups per CU
        const int stage_bytes =
            tn.spread_stage_kb > 0 ? tn.spread_stage_kb * 1024
                                   : (NQ <= 12 ? 48 * 1024 : (mf && ma ? 60 * 1024 : 65536));
        const int cap = (int)(stage_bytes / (rec * sizeof(T) + 9 * sizeof(int)));
        // the output tiles reuse the stage's LDS: all meshes, or half of
        // them above 48 KB (the allocation is at least that)
        const size_t tile = (size_t)NQ * 512 * sizeof(T);
        size_t lds = std::max((size_t)cap * (rec * sizeof(T) + 9 * sizeof(int)),
                              tile <= 49152 ? tile : tile / 2);
        if (mf && ma && !(spread_form & 4))
            lds = (((size_t)cap * (rec * sizeof(T) + 9 * sizeof(int)) + 15) & ~(size_t)15) +
                  kMaTileBytes;
        const int nsx = (K + 31) / 32;
        dim3 grid(nsx * (K / 4) * (K / 4), ma ? 1 : ceil_div(b, NQ / NC));
#define SDT_SPREAD_PWS(PP, QQ, BB, WW, SS, MM, AA)                                             \
    hipLaunchKernelGGL((spread_pt_kernel<T, PP, QQ, BB, WW, SS, MM, AA>), grid, dim3(kPtThreads), \
                       lds, s,                                                                   \
                       K, ts, ntile, nsx, d_tile_start.get(), d_frac.get(), d_base.get(),      \
                       poly ? d_a2.get() : (double*)nullptr, X, xb, npos, b, NQ / NC, NC, cap, \
                       d_real.get(), (spread_form & 4) ? 1 : 0)
#define SDT_SPREAD_PW(PP, QQ, BB, WW)                                                          \
    do {                                                                                      \
        if constexpr (BB && sizeof(T) == 4) {                                                 \
            if (sgf) SDT_SPREAD_PWS(PP, QQ, BB, WW, true, false, false);                       \
            else if constexpr (QQ == 48 && !WW) {                                             \
                if (ma) SDT_SPREAD_PWS(PP, QQ, BB, WW, false, true, true);                     \
                else if (mf) SDT_SPREAD_PWS(PP, QQ, BB, WW, false, true, false);               \
                else SDT_SPREAD_PWS(PP, QQ, BB, WW, false, false, false);                      \
            } else SDT_SPREAD_PWS(PP, QQ, BB, WW, false, false, false);                        \
        } else if constexpr (BB && QQ == NQL && !WW) {                                      \
            if (ma) SDT_SPREAD_PWS(PP, QQ, BB, WW, false, false, true);                        \
            else SDT_SPREAD_PWS(PP, QQ, BB, WW, false, false, false);                          \
        } else {                                                                              \
            SDT_SPREAD_PWS(PP, QQ, BB, WW, false, false, false);                               \
        }                                                                                     \
    } while (0)
#define SDT_SPREAD_PT(PP, QQ, BB)                                                              \
    do {                                                                                      \
        if (K < 32 + PP - 1) SDT_SPREAD_PW(PP, QQ, BB, true);                                  \
        else SDT_SPREAD_PW(PP, QQ, BB, false);                                                 \
    } while (0)
#define SDT_SPREAD_P(PP)                                                                       \
    case PP:                                                                                   \
        if (NQ == 4) SDT_SPREAD_PT(PP, 4, false);                                              \
        else if (NQ == 12) SDT_SPREAD_PT(PP, 12, false);                                       \
        else if (xb) SDT_SPREAD_PT(PP, NQL, true);                                             \
        else SDT_SPREAD_PT(PP, NQL, false);                                                    \
        break;
        switch (porder) { SDT_SPREAD_P(4) SDT_SPREAD_P(5) SDT_SPREAD_P(6) SDT_SPREAD_P(7) SDT_SPREAD_P(8) }
#undef SDT_SPREAD_P
#undef SDT_SPREAD_PT
#undef SDT_SPREAD_PW
#undef SDT_SPREAD_PWS
        SDT_HIP_CHECK(hipGetLastError());
        return;
    }
    // right-hand sides per workgroup: RB * NC register accumulators per point
    const int RBNC = poly ? 12 : 12;
    const int RB = RBNC / NC;
    const int nsx = (ntile + 3) / 4;
    const int rec = 3 * porder + 17 + RBNC;
    // staging list of up to 64 KB (any capacity works: a full stage is
    // accumulated and emptied)
    const int cap = (int)(65536 / (rec * sizeof(T)));
    SDT_REQUIRE(cap >= 64, SDT_EINVAL, "spread: staging buffer too small");
    const size_t lds = (size_t)cap * rec * sizeof(T);
    dim3 grid(nsx * ntile * ntile, ceil_div(b, RB));
#define SDT_SPREAD(PP)                                                                        \
    case PP:                                                                                  \
        hipLaunchKernelGGL((spread_gather_kernel<T, PP, 12>), grid, dim3(256), lds, s, K, ts, \
                           ntile, nsx, d_tile_start.get(), d_frac.get(), d_base.get(),        \
                           poly ? d_a2.get() : (double*)nullptr, X, b, RB, NC, cap,          \
                           d_real.get());                                                     \
        break;
    switch (porder) { SDT_SPREAD(4) SDT_SPREAD(5) SDT_SPREAD(6) SDT_SPREAD(7) SDT_SPREAD(8) }
#undef SDT_SPREAD
    SDT_HIP_CHECK(hipGetLastError());
}

template <typename T>
static int pick_lnp(int K, int ncomp_pen, size_t budget) {
    // log2 of the pencils per workgroup so that 2 padded buffers + twiddles
    // fit the budget
    int lnp = 4;
    while (lnp > 0 &&
           (2 * (size_t)ncomp_pen * ((size_t)1 << lnp) * pencil_stride(K) + K) * sizeof(cplx<T>) >
               budget)
        lnp--;
    return lnp;
}

template <typename T>
static size_t fft_lds_bytes(int K, int npen) {
    return (K + 2 * (size_t)npen * pencil_stride(K)) * sizeof(cplx<T>);
}

// largest dynamic LDS allocation of one workgroup on the current device
static size_t max_lds_bytes() {
    static size_t v = 0;
    if (!v) {
        int dev = 0, x = 0;
        SDT_HIP_CHECK(hipGetDevice(&dev));
        SDT_HIP_CHECK(hipDeviceGetAttribute(&x, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
        v = (size_t)x;
    }
    return v;
}

template <typename T>
static size_t plane_lds_bytes(int K) {
    return (K + (size_t)(K / 2 + 1) * pencil_stride(K)) * sizeof(cplx<T>);
}

// the fused x + y plane kernels apply when one half-spectrum plane fits the
// LDS of a workgroup and the per-thread butterfly / pack bounds hold:
// 0 = not applicable, 1 = the K <= 128 variant, 2 = the K <= 180 variant
template <typename T>
static int plane_path(int K) {
    if (plane_lds_bytes<T>(K) > max_lds_bytes()) return 0;
    const int Kh = K / 2 + 1, H = K / 2, nt = kPlaneThreads;
    if (Kh * K <= 2 * 5 * nt && H * K <= 8 * nt && H * Kh <= 8 * nt) return 1;
    if (sizeof(T) == 4 && Kh * K <= 2 * 8 * nt && H * K <= 16 * nt && H * Kh <= 16 * nt) return 2;
    return 0;
}

template <typename T>
static void allow_lds(const void* fn, size_t bytes) {
    if (bytes <= 64 * 1024) return;
    // raised once per kernel and size: later launches (and graph captures of
    // them) make no attribute call
    static std::mutex mu;
    static std::map<const void*, size_t> raised;
    std::lock_guard<std::mutex> lk(mu);
    size_t& have = raised[fn];
    if (have >= bytes) return;
    SDT_HIP_CHECK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    have = bytes;
}

template <typename T>
void SpmeGpu<T>::fft_forward(int b, hipStream_t s) {
    const FFTArgs fa = fft_args(plans[width_class(b)]);
    const int Kh = K / 2 + 1;
    const long long nrows = (long long)nmesh(b) * K * K;
    if (const int pv = fft_form == 1 ? 0 : plane_path<T>(K)) {
        const size_t lds = plane_lds_bytes<T>(K);
        const dim3 grid((unsigned)(nmesh(b) * K));
        if (sizeof(T) == 4 && K == 128) {
            // two-stage 16 x 8 transforms (FP32: 16 complex registers)
            allow_lds<T>((const void*)fft_plane_fwd_kernel<T, 5, 8, 16, 8>, lds);
            hipLaunchKernelGGL((fft_plane_fwd_kernel<T, 5, 8, 16, 8>), grid, dim3(kPlaneThreads), lds, s,
                               fa, d_real.get(), d_spec.get(), d_tw.get());
        } else if (K == 64) {
            allow_lds<T>((const void*)fft_plane_fwd_kernel<T, 5, 8, 8, 8>, lds);
            hipLaunchKernelGGL((fft_plane_fwd_kernel<T, 5, 8, 8, 8>), grid, dim3(kPlaneThreads), lds, s,
                               fa, d_real.get(), d_spec.get(), d_tw.get());
        } else if (pv == 1) {
            allow_lds<T>((const void*)fft_plane_fwd_kernel<T, 5, 8, 0, 0>, lds);
            hipLaunchKernelGGL((fft_plane_fwd_kernel<T, 5, 8, 0, 0>), grid, dim3(kPlaneThreads), lds, s,
                               fa, d_real.get(), d_spec.get(), d_tw.get());
        } else if constexpr (sizeof(T) == 4) {
            allow_lds<T>((const void*)fft_plane_fwd_kernel<T, 8, 16, 0, 0>, lds);
            hipLaunchKernelGGL((fft_plane_fwd_kernel<T, 8, 16, 0, 0>), grid, dim3(kPlaneThreads), lds, s,
                               fa, d_real.get(), d_spec.get(), d_tw.get());
        }
        SDT_HIP_CHECK(hipGetLastError());
        return;
    }
    const int NR = 2 << pick_lnp<T>(K, 1, 48 * 1024);
    hipLaunchKernelGGL((fft_x_r2c_kernel<T>), dim3(ceil_div(nrows, NR)), dim3(256),
                       fft_lds_bytes<T>(K, NR / 2), s, fa, nrows, d_real.get(), d_spec.get(),
                       d_tw.get(), NR);
    SDT_HIP_CHECK(hipGetLastError());
    const int lnp = pick_lnp<T>(K, 1, 48 * 1024);
    hipLaunchKernelGGL((fft_y_kernel<T, -1>), dim3(ceil_div(Kh, 1 << lnp) * nmesh(b) * K),
                       dim3(256), fft_lds_bytes<T>(K, 1 << lnp), s, fa, d_spec.get(), d_tw.get(),
                       lnp);
    SDT_HIP_CHECK(hipGetLastError());
}

template <typename T>
void SpmeGpu<T>::influence(int b, hipStream_t s) {
    const FFTArgs fa = fft_args(plans[width_class(b)]);
    const int Kh = K / 2 + 1;
    const int NC = poly ? 6 : 3;
    const int nj = K * Kh;
    // threads per workgroup and at most 2^lnp pencils per component, one
    // padded buffer of NC * 2^lnp pencils within ~110 KB and the butterfly
    // bounds; z_form > 0 picks (threads, largest lnp) explicitly. Built-in
    // (measured at K = 128, 100 columns): FP32 512 threads x 16 pencils,
    // FP64 512 x 8 (three workgroups per CU instead of one)
    static const int forms[][2] = {{1024, 4}, {1024, 4}, {512, 4}, {512, 3}, {256, 3}, {256, 2}};
    SDT_REQUIRE(z_form >= 0 && z_form <= 5, SDT_EINVAL, "z transform: form must be in [0, 5]");
    const int zf = z_form > 0 ? z_form : (sizeof(T) == 4 ? 2 : 3);
    const int nt = forms[zf][0];
    int lnp = forms[zf][1];
    const size_t budget = std::min<size_t>(max_lds_bytes(), 112 * 1024);
    auto lds_of = [&](int l) {
        return (K + (size_t)NC * ((size_t)1 << l) * pencil_stride(K)) * sizeof(cplx<T>) +
               (size_t)K * sizeof(T);
    };
    while (lnp > 0 && (lds_of(lnp) > budget || NC * (K << lnp) > 2 * kZMaxI2 * nt)) lnp--;
    SDT_REQUIRE(lds_of(lnp) <= max_lds_bytes() && NC * K <= 2 * kZMaxI2 * nt, SDT_EINVAL,
                "z transform: mesh dimension too large");
    const size_t lds = lds_of(lnp);
    const unsigned nch = ceil_div(nj, (size_t)1 << lnp);
    const dim3 grid(nch * (unsigned)b);
#define SDT_Z_LAUNCH(R1, R2, NT)                                                              \
    do {                                                                                      \
        allow_lds<T>((const void*)fft_z_influence_kernel<T, R1, R2, NT>, lds);                \
        hipLaunchKernelGGL((fft_z_influence_kernel<T, R1, R2, NT>), grid, dim3(NT), lds, s, fa, \
                           d_spec.get(), d_tw.get(), d_bsp2.get(), lnp, NC, box, xi);         \
    } while (0)
#define SDT_Z_SHAPES(NT)                                                                      \
    do {                                                                                      \
        /* two-stage transforms need one stage-1 butterfly per thread */                      \
        if (sizeof(T) == 4 && K == 128 && (NC << lnp) * 8 <= NT) SDT_Z_LAUNCH(16, 8, NT);     \
        else if (K == 64 && (NC << lnp) * 8 <= NT) SDT_Z_LAUNCH(8, 8, NT);                    \
        else SDT_Z_LAUNCH(0, 0, NT);                                                          \
    } while (0)
    if (nt == 1024) SDT_Z_SHAPES(1024);
    else if (nt == 512) SDT_Z_SHAPES(512);
    else SDT_Z_SHAPES(256);
#undef SDT_Z_SHAPES
#undef SDT_Z_LAUNCH
    SDT_HIP_CHECK(hipGetLastError());
}

template <typename T>
void SpmeGpu<T>::fft_backward(int b, hipStream_t s) {
    const FFTArgs fa = fft_args(plans[width_class(b)]);
    const int Kh = K / 2 + 1;
    const long long nrows = (long long)nmesh(b) * K * K;
    if (const int pv = fft_form == 1 ? 0 : plane_path<T>(K)) {
        const size_t lds = plane_lds_bytes<T>(K);
        const dim3 grid((unsigned)(nmesh(b) * K));
        if (sizeof(T) == 4 && K == 128) {
            // two-stage 16 x 8 transforms (FP32: 16 complex registers)
            allow_lds<T>((const void*)fft_plane_inv_kernel<T, 5, 8, 16, 8>, lds);
            hipLaunchKernelGGL((fft_plane_inv_kernel<T, 5, 8, 16, 8>), grid, dim3(kPlaneThreads), lds, s,
                               fa, d_spec.get(), d_real.get(), d_tw.get());
        } else if (K == 64) {
            allow_lds<T>((const void*)fft_plane_inv_kernel<T, 5, 8, 8, 8>, lds);
            hipLaunchKernelGGL((fft_plane_inv_kernel<T, 5, 8, 8, 8>), grid, dim3(kPlaneThreads), lds, s,
                               fa, d_spec.get(), d_real.get(), d_tw.get());
        } else if (pv == 1) {
            allow_lds<T>((const void*)fft_plane_inv_kernel<T, 5, 8, 0, 0>, lds);
            hipLaunchKernelGGL((fft_plane_inv_kernel<T, 5, 8, 0, 0>), grid, dim3(kPlaneThreads), lds, s,
                               fa, d_spec.get(), d_real.get(), d_tw.get());
        } else if constexpr (sizeof(T) == 4) {
            allow_lds<T>((const void*)fft_plane_inv_kernel<T, 8, 16, 0, 0>, lds);
            hipLaunchKernelGGL((fft_plane_inv_kernel<T, 8, 16, 0, 0>), grid, dim3(kPlaneThreads), lds, s,
                               fa, d_spec.get(), d_real.get(), d_tw.get());
        }
        SDT_HIP_CHECK(hipGetLastError());
        return;
    }
    const int lnp = pick_lnp<T>(K, 1, 48 * 1024);
    hipLaunchKernelGGL((fft_y_kernel<T, +1>), dim3(ceil_div(Kh, 1 << lnp) * nmesh(b) * K),
                       dim3(256), fft_lds_bytes<T>(K, 1 << lnp), s, fa, d_spec.get(), d_tw.get(),
                       lnp);
    SDT_HIP_CHECK(hipGetLastError());
    const int NR = 2 << pick_lnp<T>(K, 1, 48 * 1024);
    hipLaunchKernelGGL((fft_x_c2r_kernel<T>), dim3(ceil_div(nrows, NR)), dim3(256),
                       fft_lds_bytes<T>(K, NR / 2), s, fa, nrows, d_spec.get(), d_real.get(),
                       d_tw.get(), NR);
    SDT_HIP_CHECK(hipGetLastError());
}

template <typename T>
void SpmeGpu<T>::interp(int b, double alpha, double beta, double* Y, hipStream_t s) {
    const int NC = poly ? 6 : 3;
    // alpha * K^3 / L^3 with the unnormalised inverse transform: alpha / L^3
    const double a_eff = alpha / (box * box * box);
    const int RB = 4;
    const int R = porder - 1;
    // strip width (tiles along x) and particle chunk (measured at K = 128,
    // p = 6): FP32 4-tile strips, FP64 2-tile strips
    // interp_form > 0 overrides the meshes per pass (FP32 1, 2 or 4; FP64 1
    // or 2); one is fastest at K = 128, p = 6 in both precisions (the kernel
    // is bound by LDS reads, which wider vectors do not reduce)
    const int MV = interp_form > 0 ? interp_form : 1;
    SDT_REQUIRE(MV == 1 || MV == 2 || (MV == 4 && sizeof(T) == 4), SDT_EINVAL,
                "interp: meshes per pass must be 1, 2 or (FP32) 4");
    const int rlim = MV == 4 ? 16 : 26;
    int sxt = sizeof(T) == 8 ? 2 : 4, ch = sizeof(T) == 8 ? 64 : 128;
    const SpmeTuning& tn = tune[width_class(b)];
    if (tn.interp_sxt > 0) sxt = tn.interp_sxt;
    if (tn.interp_ch > 0) ch = tn.interp_ch;
    auto lds_of = [&](int sx, int c) {
        const size_t nreg = (size_t)(sx * ts + R) * (ts + R) * (ts + R);
        return nreg * MV * sizeof(T) + (size_t)c * porder * MV * sizeof(T) +
               (size_t)c * 3 * porder * sizeof(T) + (size_t)c * 3 * sizeof(int) + 8 +
               (size_t)c * RB * 3 * sizeof(double);
    };
    // narrower strips for wide stencils: the region must fit the prefetch
    // registers (26 points per thread)
    auto rreg_of = [&](int sx) {
        return (int)ceil_div((size_t)(sx * ts + R) * (ts + R) * (ts + R), 256);
    };
    while (rreg_of(sxt) > rlim && sxt > 1) sxt /= 2;
    while (lds_of(sxt, ch) > 80 * 1024 && (sxt > 1 || ch > 32)) {
        if (sxt > 1) sxt /= 2;
        else ch /= 2;
    }
    const size_t lds = lds_of(sxt, ch);
    SDT_REQUIRE(lds <= 80 * 1024, SDT_EINVAL, "interp: LDS region too large");
    const int nsx = (ntile + sxt - 1) / sxt;
    // region points per thread (registers of the prefetch)
    const int rreg = rreg_of(sxt);
    SDT_REQUIRE(rreg <= rlim, SDT_EINVAL, "interp: region too large for the prefetch registers");
    dim3 grid(nsx * ntile * ntile, ceil_div(b, RB));
#define SDT_INTERP_R(PP, RR)                                                                  \
    if (MV == 1) SDT_INTERP_M(PP, RR, 1);                                                     \
    else if (MV == 2) SDT_INTERP_M(PP, RR, 2);                                                \
    else if constexpr (sizeof(T) == 4) SDT_INTERP_M(PP, 16, 4);
#define SDT_INTERP_M(PP, RR, MM)                                                              \
    do {                                                                                      \
    allow_lds<T>((const void*)interp_tile_kernel<T, PP, RR, MM>, lds);                            \
    hipLaunchKernelGGL((interp_tile_kernel<T, PP, RR, MM>), grid, dim3(256), lds, s, K, ts, ntile, \
                       nsx, sxt, ch, d_tile_start.get(), d_frac.get(), d_base.get(),              \
                       poly ? d_a2.get() : (double*)nullptr, d_real.get(), b, RB, NC, a_eff, beta, \
                       Y);                                                                \
    } while (0)
#define SDT_INTERP(PP)                                                                        \
    case PP:                                                                                  \
        if (rreg <= 16) { SDT_INTERP_R(PP, 16); }                                             \
        else { SDT_INTERP_R(PP, 26); }                                                        \
        break;
    switch (porder) { SDT_INTERP(4) SDT_INTERP(5) SDT_INTERP(6) SDT_INTERP(7) SDT_INTERP(8) }
#undef SDT_INTERP_R
#undef SDT_INTERP_M
#undef SDT_INTERP
    SDT_HIP_CHECK(hipGetLastError());
}

template <typename T>
void SpmeGpu<T>::apply(int b, double alpha, const double* X, double beta, double* Y,
                       hipStream_t s, T* xb, bool fill_xb) {
    reserve_rhs(b);
    {
        StageScope sc(timer, 0, s, b);
        if (fill_xb) real_block_copy<T>(npos, b, X, xb, s);
        spread(b, X, s, xb);
    }
    { StageScope sc(timer, 1, s, b); fft_forward(b, s); }
    { StageScope sc(timer, 2, s, b); influence(b, s); }
    { StageScope sc(timer, 3, s, b); fft_backward(b, s); }
    { StageScope sc(timer, 4, s, b); interp(b, alpha, beta, Y, s); }
}

template struct SpmeGpu<float>;
template struct SpmeGpu<double>;

}  // namespace sdt
