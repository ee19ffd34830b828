// Ensemble step: every mobility interval, each system's update and Lanczos
// solve run concurrently (host threads, one HIP stream per system); every
// other step, the tails of all systems (forces, M f, position update) are
// one HIP graph launch: the systems' tails are captured as parallel branches
// forked from and joined to one origin stream, and the graph is replayed for
// the rest of the interval.
#include <algorithm>
#include <chrono>
#include <thread>
#include "ensemble.hpp"

namespace sdt {

EnsembleGpu::EnsembleGpu(const std::vector<EnsembleSystem>& sys, const double* pos,
                         const double* rdi, int porder, int precision, int max_iters, int max_nrhs,
                         double tol, double r0, double k0, double dt, int mob_interval,
                         unsigned long long seed)
    : mob_interval_(mob_interval) {
    SDT_REQUIRE(!sys.empty(), SDT_EINVAL, "ensemble: no systems");
    SDT_REQUIRE(pos && rdi, SDT_EINVAL, "ensemble: null particle arrays");
    size_t off = 0;
    for (size_t s = 0; s < sys.size(); s++) {
        const EnsembleSystem& e = sys[s];
        SDT_REQUIRE(e.npos > 0, SDT_EINVAL, "ensemble: every system needs particles");
        bd_.emplace_back(new BDGpu(e.npos, pos + 3 * off, rdi + off, e.box, e.xi, e.rmax, e.dim,
                                   porder, precision, max_iters, max_nrhs, tol, r0, k0, dt,
                                   mob_interval, seed + s));
        off += (size_t)e.npos;
    }
    SDT_HIP_CHECK(stream_create(&s0_));
    SDT_HIP_CHECK(hipEventCreateWithFlags(&fork_, hipEventDisableTiming));
    join_.resize(bd_.size(), nullptr);
    for (hipEvent_t& e : join_) SDT_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
}

EnsembleGpu::~EnsembleGpu() {
    release_graph();
    for (hipEvent_t e : join_)
        if (e) (void)hipEventDestroy(e);
    if (fork_) (void)hipEventDestroy(fork_);
    if (s0_) (void)stream_destroy(s0_);
}

void EnsembleGpu::release_graph() {
    CaptureLock lk(capture_mutex());
    if (exec_) (void)hipGraphExecDestroy(exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    exec_ = nullptr;
    graph_ = nullptr;
    graph_block_ = -1;
}

// f(s) for every system on up to 16 host threads
template <typename F>
static void for_systems(int n, F f) {
    const int nt = std::max(1, std::min(n, 16));
    std::vector<std::thread> th;
    std::vector<Error> err(nt, Error{SDT_OK, ""});
    for (int t = 0; t < nt; t++)
        th.emplace_back([&, t] {
            for (int s = t; s < n; s += nt) {
                try {
                    f(s);
                } catch (const Error& e) {
                    err[t] = e;
                    return;
                } catch (const std::exception& e) {
                    err[t] = Error{SDT_EINVAL, e.what()};
                    return;
                }
            }
        });
    for (auto& x : th) x.join();
    for (const Error& e : err)
        if (e.code != SDT_OK) throw e;
}

void EnsembleGpu::init(bool tune, const std::string& cache_path) {
    SDT_REQUIRE(!inited_, SDT_ESTATE, "ensemble: init called twice");
    for (auto& b : bd_) {
        b->init();
        b->lanczos()->eval_threads = 1;   // the systems' tests run side by side
    }
    tune_.assign(2 * bd_.size(), TuneResult());
    if (tune) {
        // one system after another, on its start positions: timings are
        // not disturbed by other systems' work, and equal keys reuse the cache
        const int wide = std::min(mob_interval_, BrwnLanczosGpu::kMaxBatchCols);
        for (size_t s = 0; s < bd_.size(); s++) {
            MobSpmeGpu* m = bd_[s]->mob();
            std::vector<double> pos(3 * (size_t)bd_[s]->npos());
            bd_[s]->get_positions(pos.data());
            m->update(pos.data(), nullptr, false);
            tune_[2 * s] = autotune_spme(m, 1, cache_path);
            tune_[2 * s + 1] = autotune_spme(m, wide, cache_path);
        }
    }
    inited_ = true;
}

void EnsembleGpu::tail_all() {
    const int n = (int)bd_.size();
    bool warm = true;
    for (auto& b : bd_) warm = warm && b->warm_block() == b->block();
    if (!use_graph_ || !warm) {
        // eager, each system on its own stream
        for (auto& b : bd_) b->enqueue_tail(b->stream());
        return;
    }
    const long long blk = bd_[0]->block();
    if (graph_block_ != blk) {
        // no other thread's allocation, free or synchronous copy meanwhile
        CaptureLock lk(capture_mutex());
        release_graph();
        SDT_HIP_CHECK(hipStreamBeginCapture(s0_, hipStreamCaptureModeThreadLocal));
        try {
            SDT_HIP_CHECK(hipEventRecord(fork_, s0_));
            for (int s = 0; s < n; s++) {
                hipStream_t ss = bd_[s]->stream();
                SDT_HIP_CHECK(hipStreamWaitEvent(ss, fork_, 0));
                bd_[s]->enqueue_tail(ss);
                SDT_HIP_CHECK(hipEventRecord(join_[s], ss));
                SDT_HIP_CHECK(hipStreamWaitEvent(s0_, join_[s], 0));
            }
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(s0_, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        SDT_HIP_CHECK(hipStreamEndCapture(s0_, &graph_));
        SDT_HIP_CHECK(hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0));
        graph_block_ = blk;
    }
    SDT_HIP_CHECK(hipGraphLaunch(exec_, s0_));
    graph_launches_++;
}

void EnsembleGpu::advance(int nsteps, double* step_ms) {
    SDT_REQUIRE(inited_, SDT_ESTATE, "ensemble: advance called before init");
    const int n = (int)bd_.size();
    for (int k = 0; k < nsteps; k++) {
        const auto t0 = std::chrono::steady_clock::now();
        // interval starts: all systems' updates and solves side by side
        for_systems(n, [&](int s) {
            if (bd_[s]->begin_step()) SDT_HIP_CHECK(hipStreamSynchronize(bd_[s]->stream()));
        });
        tail_all();
        for (auto& b : bd_) b->end_step();
        SDT_HIP_CHECK(hipStreamSynchronize(s0_));
        for (auto& b : bd_) SDT_HIP_CHECK(hipStreamSynchronize(b->stream()));
        if (step_ms)
            step_ms[k] = std::chrono::duration<double, std::milli>(
                             std::chrono::steady_clock::now() - t0).count();
    }
}

void EnsembleGpu::get_positions(int s, double* pos) {
    SDT_REQUIRE(s >= 0 && s < (int)bd_.size(), SDT_EINVAL, "ensemble: system index out of range");
    bd_[s]->get_positions(pos);
}

}  // namespace sdt
