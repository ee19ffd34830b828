// Diffusion coefficients from a trajectory (tools/calc_diff/calc_diff.cc).
#pragma once
#include "sdt_internal.hpp"

namespace sdt {

// number of lags the trajectory gives: nframes / 2 - 1 (0 for fewer than 4)
int diffusion_lags(int nframes);
// traj: nframes frames of npos AoS positions (host, frame-major); D receives
// diffusion_lags(nframes) values, D[i - 1] for lag i (dt: frame interval).
// Synchronises.
void diffusion(const double* traj, int nframes, int npos, double dt, double* D, hipStream_t s);

// Partial sums the allocation-free variant holds in its static device array
// (8 MiB per GPU): ceil(npos / 256) * (nframes / 2 - 1) must not exceed it.
constexpr size_t kDiffusionScratchDoubles = (size_t)1 << 20;
constexpr int kMaxDevices = 64;
// The same on device pointers with no device allocation: d_traj and d_D on
// the current device; enqueued on s without waiting. Calls that overlap on
// one GPU share the scratch array and race.
void diffusion_noalloc(const double* d_traj, int nframes, int npos, double dt, double* d_D,
                       hipStream_t s);

}  // namespace sdt
