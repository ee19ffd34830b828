// Batched ensemble of independent BD systems (uneven particle counts, boxes
// and meshes) advanced together on one GPU.
#pragma once
#include <memory>
#include <string>
#include <vector>
#include "mob_gpu.hpp"
#include "tune.hpp"

namespace sdt {

struct EnsembleSystem {
    int npos = 0;
    double box = 0, xi = 0, rmax = 0;
    int dim = 0;
};

class EnsembleGpu {
  public:
    // pos / rdi: every system's AoS positions and radii, one after another;
    // system s draws its Gaussian blocks from seed + s
    EnsembleGpu(const std::vector<EnsembleSystem>& sys, const double* pos, const double* rdi,
                int porder, int precision, int max_iters, int max_nrhs, double tol, double r0,
                double k0, double dt, int mob_interval, unsigned long long seed);
    ~EnsembleGpu();
    // allocates every system; with tune on, each system's SPME launch
    // choices are picked by the autotuner for the single-column step and
    // for the interval's Lanczos batch (cache_path: earlier results, may be
    // empty)
    void init(bool tune, const std::string& cache_path);
    // nsteps steps of every system; step_ms (optional) receives the host
    // wall time of each step (the slowest system's work, all in flight)
    void advance(int nsteps, double* step_ms);
    void get_positions(int s, double* pos);
    int size() const { return (int)bd_.size(); }
    BDGpu* system(int s) { return bd_[s].get(); }
    const std::vector<TuneResult>& tuning() const { return tune_; }
    void set_graphs(bool on) { use_graph_ = on; }
    long long graph_launches() const { return graph_launches_; }

  private:
    void tail_all();
    void release_graph();
    std::vector<std::unique_ptr<BDGpu>> bd_;
    std::vector<TuneResult> tune_;   // per system: narrow (2 s) and wide (2 s + 1)
    // The step graph is split into lanes: each lane's systems are one graph
    // on the lane's own origin stream, and the lanes' graphs are launched
    // side by side (the branches of one graph are dispatched one after
    // another, and the small systems' tails are bound by that dispatch).
    static constexpr int kGraphLanes = 4;
    std::vector<std::vector<int>> lane_;   // the systems of each lane
    std::vector<hipStream_t> s0_;          // origin stream of each lane's graph
    std::vector<hipEvent_t> join_;         // per system
    std::vector<hipEvent_t> fork_;         // per lane
    std::vector<hipGraph_t> graph_;
    std::vector<hipGraphExec_t> exec_;
    long long graph_block_ = -1;
    long long graph_launches_ = 0;
    bool use_graph_ = true, inited_ = false;
    int mob_interval_;
};

}  // namespace sdt
