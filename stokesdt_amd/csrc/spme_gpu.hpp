// Reciprocal-space SPME engine on the GPU (host-side object).
//
// Mesh data in HBM:
//   real meshes   [m][z][y][x]      K^3 reals per mesh (m = rhs*3+comp,
//                                    plus 3*b more for the polydisperse a^2 f)
//   half spectra  [m][z][y][kx]     K*K*(K/2+1) complex per mesh
// Particles are kept sorted by the spread tile that holds their B-spline
// base point; vectors inside the library use the packed layout
//   X[(3*i_sorted + c) * b + r]     (row-major 3N x b, FP64)
#pragma once
#include "sdt_internal.hpp"

namespace sdt {

struct SortedSystem;   // particle order shared by SPME and real space

// |b(m)|^2 = 1 / |sum_{i<P-1} M_P(i+1) e^{2 pi i m i / K}|^2 for m = 0..K-1
// (src/mob/spme.cc:80-128), with the neighbour mean at a vanishing
// denominator (odd P at the Nyquist index)
void bspline_moduli(int K, int P, std::vector<double>& out);

// Launch choices of one SPME engine that leave the operator unchanged (up to
// the rounding of the FFT's stage order): picked by the autotuner, 0 = the
// built-in heuristic.
struct SpmeTuning {
    int fft_order = 0;        // generic radix chain: 0 radix 8 first, then 4, 2, 3, 5;
                              // 1 radix 4 first (no 8); 2 radices 3 and 5 first
    int interp_sxt = 0;       // interp strip width in tiles (1, 2, 4)
    int interp_ch = 0;        // interp particle chunk per pass (32 .. 128)
    int spread_stage_kb = 0;  // spread candidate stage in KB (32, 48, 64)
    bool operator==(const SpmeTuning& o) const {
        return fft_order == o.fft_order && interp_sxt == o.interp_sxt &&
               interp_ch == o.interp_ch && spread_stage_kb == o.spread_stage_kb;
    }
};

template <typename T>
struct SpmeGpu {
    int npos = 0, K = 0, porder = 0;
    double box = 0, xi = 0;
    bool poly = false;
    int ts = 8;            // spread/interp tile side (mesh points)
    int ntile = 0;         // tiles per dimension
    int bcap = 0;          // RHS capacity of the mesh buffers
    FFTPlan plan;          // the default radix plan
    // launch choices and FFT plans of narrow (width class 0: b < 8) and wide
    // (1: b >= 8) launches
    SpmeTuning tune[2];
    FFTPlan plans[2];
    static int width_class(int b) { return b >= 8 ? 1 : 0; }
    // interp meshes per LDS pass: 0 the built-in choice (1), else 1, 2 or
    // (FP32) 4
    int interp_form = 0;
    // z transform launch shape: 0 the built-in choice (FP32 2, FP64 3),
    // 1 .. 5 (threads, pencils) = (1024, 16), (1024, 16), (512, 16),
    // (512, 8), (256, 8), (256, 4)
    int z_form = 0;
    // spread launch form: bits 0-1, force reads of wide FP32 launches (0
    // built-in = scalar loads, 1 staged in LDS, 2 scalar loads); bit 2, mesh
    // stores as whole rows through LDS instead of per lane
    int spread_form = 0;
    // x / y transforms: 0 the fused plane kernels where a plane fits LDS,
    // 1 the separate x and y passes
    int fft_form = 0;
    // replaces the launch choices of a width class (re-plans the FFT order)
    void set_tuning(const SpmeTuning& t, int width);
    DevBuf<double> d_tw;   // K roots of unity (cos, sin) of -2*pi*j/K
    DevBuf<double> d_bsp2; // |b(m)|^2 per wavenumber index, K entries
    DevBuf<double> d_frac; // 3N fractional offsets (sorted order)
    DevBuf<int> d_base;    // 3N B-spline base mesh indices (sorted order)
    DevBuf<double> d_a2;   // N radii squared (sorted order), poly only
    DevBuf<int> d_tile_start;  // ntile^3 + 1
    DevBuf<T> d_real;      // nmesh * K^3
    DevBuf<cplx<T>> d_spec;    // nmesh * K*K*Kh

    void create(int npos, double box, double xi, int K, int porder, bool poly);
    void reserve_rhs(int b);
    // positions (original order, FP64 device) -> particle tiles; returns the
    // sort key of every particle in key_out (original order)
    void prepare(const double* d_pos, int* d_key_out, hipStream_t s);
    // after the shared sort: gather per-particle data into sorted order
    void gather_sorted(const int* d_perm, const double* d_pos, const double* d_rdi,
                       const int* d_tile_start_in, hipStream_t s);
    // Y = alpha * M_recip X + beta * Y, packed FP64 vectors, b RHS; xb:
    // optional column-blocked copy of X in T with blocked_cols(b) columns
    // per block (Xb[cb][j][c][r], zero-padded), which the spread then reads
    // (written first, within the spread stage, when fill_xb is set)
    void apply(int b, double alpha, const double* X, double beta, double* Y,
               hipStream_t s, T* xb = nullptr, bool fill_xb = false);
    // columns per block of the copy the spread can read for b (0: none)
    int blocked_cols(int b) const;
    bool point_owner() const;

    // stages (exposed for tests/benchmarks)
    void spread(int b, const double* X, hipStream_t s, const T* xb = nullptr);
    void fft_forward(int b, hipStream_t s);
    void influence(int b, hipStream_t s);
    void fft_backward(int b, hipStream_t s);
    void interp(int b, double alpha, double beta, double* Y, hipStream_t s);
    int nmesh(int b) const { return (poly ? 6 : 3) * b; }
    StageTimer* timer = nullptr;

    // scratch from prepare()
    DevBuf<double> d_frac_u;
    DevBuf<int> d_base_u;
};

}  // namespace sdt
