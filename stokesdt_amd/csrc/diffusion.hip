// Diffusion coefficients from a stored trajectory (tools/calc_diff/
// calc_diff.cc:100-140): for lag i the mean square displacement over the
// first nframes/2 origins and all particles, D(i) = MSD / (6 i dt).
//
// The trajectory is frame-major on the device (frame f at f * 3N, AoS
// within a frame). One thread per (particle, lag) sums over the origins;
// a workgroup reduces its particles by a fixed tree and a second kernel sums
// the workgroups of each lag in order, so the result is deterministic.
#include <mutex>
#include "diffusion.hpp"

namespace sdt {

constexpr int kMsdThreads = 256;

__global__ void __launch_bounds__(kMsdThreads)
msd_partial_kernel(const double* __restrict__ traj, int npos, int norig, long long fstride,
                   double* __restrict__ part) {
    const int lag = blockIdx.y + 1;
    const int k = blockIdx.x * kMsdThreads + threadIdx.x;
    double acc = 0.0;
    if (k < npos) {
        const double* a = traj + 3 * (size_t)k;
        const double* b = a + (size_t)lag * fstride;
        for (int j = 0; j < norig; j++) {
            const double dx = b[0] - a[0], dy = b[1] - a[1], dz = b[2] - a[2];
            acc += dx * dx + dy * dy + dz * dz;
            a += fstride;
            b += fstride;
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o);
    __shared__ double s_w[kMsdThreads / 64];
    if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < kMsdThreads / 64; w++) t += s_w[w];
        part[(size_t)blockIdx.y * gridDim.x + blockIdx.x] = t;
    }
}

__global__ void msd_final_kernel(int nblk, int nlags, const double* __restrict__ part,
                                 double scale, double* __restrict__ D) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= nlags) return;
    double t = 0.0;
    for (int b = 0; b < nblk; b++) t += part[(size_t)l * nblk + b];
    D[l] = t * scale / (double)(l + 1);
}

int diffusion_lags(int nframes) { return nframes / 2 - 1 > 0 ? nframes / 2 - 1 : 0; }

void diffusion(const double* h_traj, int nframes, int npos, double dt, double* h_D,
               hipStream_t s) {
    SDT_REQUIRE(npos > 0 && nframes > 2, SDT_EINVAL,
                "diffusion: need npos > 0 and more than two frames");
    SDT_REQUIRE(dt != 0.0, SDT_EINVAL, "diffusion: the frame interval must be non-zero");
    const int nlags = diffusion_lags(nframes);
    if (nlags == 0) return;
    SDT_REQUIRE(nlags <= 65535, SDT_EINVAL, "diffusion: at most 131072 frames");
    const int norig = nframes / 2;
    const size_t fr = (size_t)3 * npos;
    const int nblk = (npos + kMsdThreads - 1) / kMsdThreads;
    DevBuf<double> d_traj, d_part, d_D;
    d_traj.reserve(fr * nframes);
    d_part.reserve((size_t)nblk * nlags);
    d_D.reserve(nlags);
    SDT_HIP_CHECK(hipMemcpyAsync(d_traj.get(), h_traj, sizeof(double) * fr * nframes,
                                 hipMemcpyHostToDevice, s));
    msd_partial_kernel<<<dim3(nblk, nlags), kMsdThreads, 0, s>>>(d_traj.get(), npos, norig,
                                                                 (long long)fr, d_part.get());
    SDT_HIP_CHECK(hipGetLastError());
    // calc_diff.cc:130-131: divided by 6 * particles * origins * lag * dt
    const double scale = 1.0 / (6.0 * npos * (double)norig * dt);
    msd_final_kernel<<<(nlags + 255) / 256, 256, 0, s>>>(nblk, nlags, d_part.get(), scale,
                                                         d_D.get());
    SDT_HIP_CHECK(hipGetLastError());
    SDT_HIP_CHECK(hipMemcpyAsync(h_D, d_D.get(), sizeof(double) * nlags, hipMemcpyDeviceToHost, s));
    SDT_HIP_CHECK(hipStreamSynchronize(s));
}

// ------------------------------------------------------- no allocation ----
// The allocation-free variant keeps the per-workgroup partial sums in a
// file-scope device array (one copy per GPU, held until the process ends).
// Host code never passes the array's name to a kernel: the name is a
// host-side stand-in, and the device address of each GPU's copy comes from
// hipGetSymbolAddress on the first call with that device current.
static __device__ __align__(256) double scratch[kDiffusionScratchDoubles];

namespace {
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {nullptr};
}  // namespace

static double* scratch_ptr() {
    int dev = 0;
    SDT_HIP_CHECK(hipGetDevice(&dev));
    SDT_REQUIRE(dev >= 0 && dev < kMaxDevices, SDT_EINVAL, "diffusion: device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        SDT_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        SDT_REQUIRE(p != nullptr, SDT_EHIP, "diffusion: no device address for the scratch array");
        g_scratch_ptr[dev] = static_cast<double*>(p);
    }
    return g_scratch_ptr[dev];
}

void diffusion_noalloc(const double* d_traj, int nframes, int npos, double dt, double* d_D,
                       hipStream_t s) {
    SDT_REQUIRE(d_traj != nullptr && d_D != nullptr, SDT_EINVAL, "diffusion: null device pointer");
    SDT_REQUIRE(npos > 0 && nframes > 2, SDT_EINVAL,
                "diffusion: need npos > 0 and more than two frames");
    SDT_REQUIRE(dt != 0.0, SDT_EINVAL, "diffusion: the frame interval must be non-zero");
    const int nlags = diffusion_lags(nframes);
    if (nlags == 0) return;
    SDT_REQUIRE(nlags <= 65535, SDT_EINVAL, "diffusion: at most 131072 frames");
    const int norig = nframes / 2;
    const size_t fr = (size_t)3 * npos;
    const int nblk = (npos + kMsdThreads - 1) / kMsdThreads;
    SDT_REQUIRE((size_t)nblk * nlags <= kDiffusionScratchDoubles, SDT_ECAP,
                "diffusion (no allocation): ceil(npos / 256) * (nframes / 2 - 1) = " +
                    std::to_string((size_t)nblk * nlags) + " partial sums exceed the " +
                    std::to_string(kDiffusionScratchDoubles) + " of the static scratch array");
    double* part = scratch_ptr();
    msd_partial_kernel<<<dim3(nblk, nlags), kMsdThreads, 0, s>>>(d_traj, npos, norig,
                                                                 (long long)fr, part);
    SDT_HIP_CHECK(hipGetLastError());
    const double scale = 1.0 / (6.0 * npos * (double)norig * dt);
    msd_final_kernel<<<(nlags + 255) / 256, 256, 0, s>>>(nblk, nlags, part, scale, d_D);
    SDT_HIP_CHECK(hipGetLastError());
}

}  // namespace sdt
