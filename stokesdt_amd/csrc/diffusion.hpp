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

}  // namespace sdt
