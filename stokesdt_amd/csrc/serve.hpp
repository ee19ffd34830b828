// Serving loop for BD jobs: continuous batching on one GPU.
#pragma once
#include <atomic>
#include <condition_variable>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>
#include "mob_gpu.hpp"

namespace sdt {

// kinds of job step: the optimised path in FP64 or FP32-mixed, or the plain
// FP64 path
enum ServeKind : int { kServeFP64 = 0, kServeFP32 = 1, kServePlain = 2 };

struct ServeParams {
    int kind = kServeFP32;
    double capacity = 0;      // bytes of device memory jobs may reserve
    int max_active = 64;      // jobs on the device at once
    int porder = 6, max_iters = 100, max_nrhs = 8, mob_interval = 100;
    double tol = 1e-6, r0 = 2.0, k0 = 125.0, dt = 1e-3;
};

struct ServeJob {
    // request
    int npos = 0, dim = 0, nsteps = 0;
    double box = 0, xi = 0, rmax = 0, arrival = 0;   // arrival: seconds after run() starts
    unsigned long long seed = 0;
    std::vector<double> pos, rdi;
    // outcome
    enum State : int { kQueued = 0, kRunning = 1, kDone = 2, kFailed = 3, kStopped = 4 };
    std::atomic<int> state{kQueued};
    double admit = -1, finish = -1;   // seconds after run() starts
    size_t bytes = 0;                 // predicted footprint reserved at admission
    int steps_done = 0;
    std::vector<double> step_ms;      // host wall time of every step
    std::string error;
    std::thread worker;
};

// Predicted device footprint (bytes) of one job: particle arrays, the
// Gaussian and displacement blocks of an interval, the Lanczos basis of one
// lockstep batch, its meshes and the real-space pair list (optimised path),
// or the stored tables, blocks and one group's basis (plain path).
size_t serve_footprint(const ServeParams& p, int npos, int dim, double box, double rmax);

class ServeGpu {
  public:
    explicit ServeGpu(const ServeParams& p);
    ~ServeGpu();
    // returns the job id (submission order)
    int submit(int npos, const double* pos, const double* rdi, double box, double xi, double rmax,
               int dim, int nsteps, unsigned long long seed, double arrival);
    // Runs the loop: admits queued jobs whose arrival time has passed, first
    // come first served, while a slot is free and their predicted footprint
    // fits the capacity left; each admitted job steps on its own stream and
    // host thread, its state resident on the device, and releases its
    // reservation when it ends. Returns when every job has ended, or at the
    // deadline (seconds; <= 0: none), when running jobs stop after their
    // current step and queued ones stay queued.
    void run(double deadline);
    int njobs() const { return (int)jobs_.size(); }
    const ServeJob& job(int i) const { return *jobs_[i]; }
    double wall() const { return wall_; }
    size_t peak_bytes() const { return peak_bytes_; }
    int peak_active() const { return peak_active_; }

  private:
    void work(ServeJob* j);
    ServeParams p_;
    std::vector<std::unique_ptr<ServeJob>> jobs_;
    std::mutex mu_;
    std::condition_variable cv_;
    std::atomic<bool> stop_{false};
    double wall_ = 0;
    size_t used_ = 0, peak_bytes_ = 0;
    int active_ = 0, peak_active_ = 0;
    std::chrono::steady_clock::time_point t0_;
    double now() const;
};

}  // namespace sdt
