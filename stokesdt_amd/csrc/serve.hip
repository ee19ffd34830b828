// Continuous batching of BD jobs: a scheduler admits jobs as slots and
// predicted device memory free up; every running job advances step by step
// on its own HIP stream from its own host thread, so the steps of different
// jobs overlap on the device, and its positions, meshes and Krylov state
// stay resident until it ends.
#include <algorithm>
#include <chrono>
#include <cmath>
#include "plain_gpu.hpp"
#include "serve.hpp"

namespace sdt {

size_t serve_footprint(const ServeParams& p, int npos, int dim, double box, double rmax) {
    const double n3 = 3.0 * npos, K3 = (double)dim * dim * dim;
    const double spec = (double)dim * dim * (dim / 2 + 1) * 2.0;   // reals per half spectrum
    const int mi = p.mob_interval;
    // pair-list entries: self plus the expected neighbours within rmax at
    // the system's number density, with a 50 % margin
    const double dens = npos / (box * box * box);
    const double nnz = npos * (1.0 + 1.5 * dens * 4.0 / 3.0 * M_PI * rmax * rmax * rmax);
    double bytes = 0;
    bytes += n3 * 8.0 * 8;              // positions, forces, step vectors, copies
    bytes += 2.0 * n3 * mi * 8.0;       // Gaussian block and displacements of an interval
    bytes += nnz * 4.0 * 2 + npos * 64.0;   // CSR and cell lists
    if (p.kind == kServePlain) {
        const int p3 = p.porder * p.porder * p.porder, ldp = (p3 + 7) / 8 * 8;
        bytes += (double)npos * ldp * 12.0;             // weight and index tables
        bytes += 3.0 * K3 * 8.0 + 3.0 * spec * 8.0 + spec / 2 * 8.0;   // meshes, lm2
        bytes += nnz * 72.0;                            // stored 3x3 blocks
        bytes += n3 * p.max_nrhs * 8.0 * (p.max_iters + 1 + 2);   // one group's basis
    } else {
        const int e = p.kind == kServeFP32 ? 4 : 8;
        const int cols = std::min(mi, BrwnLanczosGpu::kMaxBatchCols);
        bytes += n3 * cols * 8.0 * (p.max_iters + 1 + 4);   // basis, Z, Y, active subsets
        bytes += 3.0 * cols * (K3 + spec) * e;             // meshes of a batch
        bytes += n3 * cols * e;                            // blocked input copy
    }
    return (size_t)(bytes * 1.1) + ((size_t)64 << 20);
}

ServeGpu::ServeGpu(const ServeParams& p) : p_(p) {
    SDT_REQUIRE(p.kind >= 0 && p.kind <= 2, SDT_EINVAL, "serve: kind must be 0, 1 or 2");
    SDT_REQUIRE(p.capacity > 0 && p.max_active > 0, SDT_EINVAL,
                "serve: capacity and slots must be positive");
}

ServeGpu::~ServeGpu() {
    stop_ = true;
    for (auto& j : jobs_)
        if (j->worker.joinable()) j->worker.join();
}

double ServeGpu::now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count();
}

int ServeGpu::submit(int npos, const double* pos, const double* rdi, double box, double xi,
                     double rmax, int dim, int nsteps, unsigned long long seed, double arrival) {
    SDT_REQUIRE(npos > 0 && pos && rdi && nsteps > 0, SDT_EINVAL,
                "serve: a job needs particles and steps");
    std::unique_ptr<ServeJob> j(new ServeJob());
    j->npos = npos;
    j->dim = dim;
    j->nsteps = nsteps;
    j->box = box;
    j->xi = xi;
    j->rmax = rmax;
    j->arrival = arrival;
    j->seed = seed;
    j->pos.assign(pos, pos + 3 * (size_t)npos);
    j->rdi.assign(rdi, rdi + npos);
    j->bytes = serve_footprint(p_, npos, dim, box, rmax);
    jobs_.push_back(std::move(j));
    return (int)jobs_.size() - 1;
}

// one job from admission to its end, on this thread and its own streams
void ServeGpu::work(ServeJob* j) {
    try {
        if (p_.kind == kServePlain) {
            PlainBDGpu bd(j->npos, j->pos.data(), j->rdi.data(), j->box, j->xi, j->rmax, j->dim,
                          p_.porder, p_.max_iters, p_.max_nrhs, p_.tol, p_.r0, p_.k0, p_.dt,
                          p_.mob_interval, j->seed);
            bd.init();
            bd.set_cancel(&stop_);
            for (int k = 0; k < j->nsteps && !stop_; k++) {
                const auto t = std::chrono::steady_clock::now();
                bd.advance(1);
                j->step_ms.push_back(std::chrono::duration<double, std::milli>(
                                         std::chrono::steady_clock::now() - t).count());
                j->steps_done++;
            }
            bd.get_positions(j->pos.data());
        } else {
            BDGpu bd(j->npos, j->pos.data(), j->rdi.data(), j->box, j->xi, j->rmax, j->dim,
                     p_.porder, p_.kind == kServeFP32 ? kFP32 : kFP64, p_.max_iters, p_.max_nrhs,
                     p_.tol, p_.r0, p_.k0, p_.dt, p_.mob_interval, j->seed);
            bd.init();
            // step tails run eagerly: a graph capture in this thread would
            // have to exclude every other job's allocations, frees and
            // synchronous copies (they can synchronise with the capturing
            // stream), and a graph saves about 2 % of a single system's tail
            bd.set_graphs(false);
            bd.lanczos()->eval_threads = 2;   // jobs run side by side
            bd.set_cancel(&stop_);
            for (int k = 0; k < j->nsteps && !stop_; k++) {
                const auto t = std::chrono::steady_clock::now();
                bd.advance(1);
                j->step_ms.push_back(std::chrono::duration<double, std::milli>(
                                         std::chrono::steady_clock::now() - t).count());
                j->steps_done++;
            }
            bd.get_positions(j->pos.data());
        }
        j->state = j->steps_done == j->nsteps ? ServeJob::kDone : ServeJob::kStopped;
    } catch (const Error& e) {
        // a solve cancelled at the deadline ends the job as stopped
        const bool cancelled = stop_ && e.code == SDT_ESTATE && e.msg == "cancelled";
        j->error = cancelled ? "" : e.msg;
        j->state = cancelled ? ServeJob::kStopped : ServeJob::kFailed;
    } catch (const std::exception& e) {
        j->error = e.what();
        j->state = ServeJob::kFailed;
    }
    {
        std::lock_guard<std::mutex> lk(mu_);
        j->finish = now();
        used_ -= j->bytes;
        active_--;
    }
    cv_.notify_all();
}

void ServeGpu::run(double deadline) {
    t0_ = std::chrono::steady_clock::now();
    stop_ = false;
    std::vector<ServeJob*> order;
    for (auto& j : jobs_)
        if (j->state == ServeJob::kQueued) order.push_back(j.get());
    std::stable_sort(order.begin(), order.end(),
                     [](const ServeJob* a, const ServeJob* b) { return a->arrival < b->arrival; });
    size_t next = 0;   // first job not yet admitted (first come, first served)
    std::unique_lock<std::mutex> lk(mu_);
    for (;;) {
        const double t = now();
        if (deadline > 0 && t >= deadline) break;
        // admit in arrival order while the head of the queue fits
        while (next < order.size()) {
            ServeJob* j = order[next];
            if (j->arrival > t) break;
            if (j->bytes > p_.capacity) {
                // can never be admitted: refused, not left blocking the queue
                j->error = "the predicted footprint exceeds the whole capacity";
                j->state = ServeJob::kFailed;
                j->admit = j->finish = t;
                next++;
                continue;
            }
            if (active_ >= p_.max_active || used_ + j->bytes > p_.capacity) break;
            used_ += j->bytes;
            active_++;
            peak_bytes_ = std::max(peak_bytes_, used_);
            peak_active_ = std::max(peak_active_, active_);
            j->admit = t;
            j->state = ServeJob::kRunning;
            j->worker = std::thread(&ServeGpu::work, this, j);
            next++;
        }
        if (next == order.size() && active_ == 0) break;
        // wake at the next arrival, a job's end or the deadline
        double wake = deadline > 0 ? deadline : t + 1.0;
        if (next < order.size() && order[next]->arrival > t)
            wake = std::min(wake, order[next]->arrival);
        else
            wake = std::min(wake, t + 1.0);
        cv_.wait_for(lk, std::chrono::duration<double>(std::max(0.0, wake - t)));
    }
    stop_ = true;
    lk.unlock();
    for (auto& j : jobs_)
        if (j->worker.joinable()) j->worker.join();
    wall_ = now();
}

}  // namespace sdt
