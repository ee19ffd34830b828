// Ensemble step: every mobility interval, each system's update and Lanczos
// solve run concurrently (host threads, one HIP stream per system); every
// other step, the tails of all systems (forces, M f, position update) are
// HIP graph launches: the systems' tails are captured as parallel branches
// forked from and joined to an origin stream, one graph per lane of systems,
// and the lanes' graphs are replayed side by side for the rest of the
// interval.
#include <algorithm>
#include <chrono>
#include <numeric>
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
    // systems dealt to the lanes largest first, in snake order, so that the
    // lanes hold similar numbers of systems and of particles
    const int nl = std::min<int>(kGraphLanes, (int)sys.size());
    std::vector<int> order(sys.size());
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return sys[a].npos > sys[b].npos; });
    lane_.assign(nl, std::vector<int>());
    for (size_t k = 0; k < order.size(); k++) {
        const int r = (int)(k % (size_t)(2 * nl));
        lane_[r < nl ? r : 2 * nl - 1 - r].push_back(order[k]);
    }
    s0_.assign(nl, nullptr);
    fork_.assign(nl, nullptr);
    graph_.assign(nl, nullptr);
    exec_.assign(nl, nullptr);
    for (int l = 0; l < nl; l++) {
        SDT_HIP_CHECK(stream_create(&s0_[l]));
        SDT_HIP_CHECK(hipEventCreateWithFlags(&fork_[l], hipEventDisableTiming));
    }
    join_.resize(bd_.size(), nullptr);
    for (hipEvent_t& e : join_) SDT_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
}

EnsembleGpu::~EnsembleGpu() {
    release_graph();
    for (hipEvent_t e : join_)
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : fork_)
        if (e) (void)hipEventDestroy(e);
    for (hipStream_t st : s0_)
        if (st) (void)stream_destroy(st);
}

void EnsembleGpu::release_graph() {
    CaptureLock lk(capture_mutex());
    for (hipGraphExec_t& x : exec_) {
        if (x) (void)hipGraphExecDestroy(x);
        x = nullptr;
    }
    for (hipGraph_t& g : graph_) {
        if (g) (void)hipGraphDestroy(g);
        g = nullptr;
    }
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
        for (size_t l = 0; l < lane_.size(); l++) {
            SDT_HIP_CHECK(hipStreamBeginCapture(s0_[l], hipStreamCaptureModeThreadLocal));
            try {
                SDT_HIP_CHECK(hipEventRecord(fork_[l], s0_[l]));
                for (int s : lane_[l]) {
                    hipStream_t ss = bd_[s]->stream();
                    SDT_HIP_CHECK(hipStreamWaitEvent(ss, fork_[l], 0));
                    bd_[s]->enqueue_tail(ss);
                    SDT_HIP_CHECK(hipEventRecord(join_[s], ss));
                    SDT_HIP_CHECK(hipStreamWaitEvent(s0_[l], join_[s], 0));
                }
            } catch (...) {
                hipGraph_t g = nullptr;
                (void)hipStreamEndCapture(s0_[l], &g);
                if (g) (void)hipGraphDestroy(g);
                release_graph();
                throw;
            }
            SDT_HIP_CHECK(hipStreamEndCapture(s0_[l], &graph_[l]));
            SDT_HIP_CHECK(hipGraphInstantiate(&exec_[l], graph_[l], nullptr, nullptr, 0));
        }
        graph_block_ = blk;
    }
    // the lanes' graphs are submitted from one host thread each
    const int nl = (int)lane_.size();
    std::vector<hipError_t> st(nl, hipSuccess);
    std::vector<std::thread> th;
    for (int l = 1; l < nl; l++) th.emplace_back([&, l] { st[l] = hipGraphLaunch(exec_[l], s0_[l]); });
    st[0] = hipGraphLaunch(exec_[0], s0_[0]);
    for (auto& x : th) x.join();
    for (int l = 0; l < nl; l++) SDT_HIP_CHECK(st[l]);
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
        for (hipStream_t st : s0_) SDT_HIP_CHECK(hipStreamSynchronize(st));
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
