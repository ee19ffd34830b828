// Host orchestration of the device mobility, Krylov displacements, steric
// force and BD step.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include "mob_gpu.hpp"
#include "brwn_random.hpp"
#include "ewald_gpu.hpp"

namespace sdt {

// ------------------------------------------------------------ helpers ------
__global__ void gather_sorted_kernel(int npos, const int* __restrict__ perm,
                                     const double* __restrict__ pos, const double* __restrict__ rdi,
                                     double* __restrict__ pos_s, double* __restrict__ rdi_s) {
    const int is = blockIdx.x * blockDim.x + threadIdx.x;
    if (is >= npos) return;
    const int i = perm[is];
    pos_s[3 * is] = pos[3 * i];
    pos_s[3 * is + 1] = pos[3 * i + 1];
    pos_s[3 * is + 2] = pos[3 * i + 2];
    rdi_s[is] = rdi[i];
}

static size_t span(int rows, int ncols, int ld) {
    return (size_t)ld * (ncols - 1) + rows;
}

// ------------------------------------------------------------ MobSpme ------
MobSpmeGpu::MobSpmeGpu(int npos, const double* rdi, double box, double xi, double rmax, int dim,
                       int porder, int precision)
    : npos_(npos), K_(dim), porder_(porder), prec_(precision), box_(box), xi_(xi), rmax_(rmax) {
    SDT_REQUIRE(npos > 0 && rdi != nullptr, SDT_EINVAL, "npos must be positive and rdi non-null");
    rdi_.assign(rdi, rdi + npos);
    SDT_REQUIRE(precision == kFP64 || precision == kFP32, SDT_EINVAL, "precision must be 0 or 1");
    SDT_HIP_CHECK(stream_create(&stream_));
    SDT_HIP_CHECK(stream_create(&side_));
    SDT_HIP_CHECK(hipEventCreateWithFlags(&ev_fork_, hipEventDisableTiming));
    SDT_HIP_CHECK(hipEventCreateWithFlags(&ev_join_, hipEventDisableTiming));
}

MobSpmeGpu::~MobSpmeGpu() {
    if (ev_fork_) (void)hipEventDestroy(ev_fork_);
    if (ev_join_) (void)hipEventDestroy(ev_join_);
    if (side_) (void)stream_destroy(side_);
    if (stream_) (void)stream_destroy(stream_);
}

void MobSpmeGpu::init() {
    // the checks of src/mob/mob_spme.cc:55-93
    SDT_REQUIRE(box_ > 0.0, SDT_EINVAL, "box size must be positive");
    SDT_REQUIRE(rmax_ > 0.0, SDT_EINVAL, "real-space cutoff must be positive");
    SDT_REQUIRE(xi_ > 0.0, SDT_EINVAL, "Ewald parameter must be positive");
    SDT_REQUIRE(K_ > 0, SDT_EINVAL, "mesh dimension must be positive");
    SDT_REQUIRE(porder_ > 1, SDT_EINVAL, "interpolation order must be greater than 1");
    SDT_REQUIRE(2.0 * rmax_ < box_, SDT_EINVAL, "rmax must be less than box_size/2");
    SDT_REQUIRE(4 * porder_ <= K_, SDT_EINVAL, "porder must be smaller than dim/4");
    const double mx = *std::max_element(rdi_.begin(), rdi_.end());
    const double mn = *std::min_element(rdi_.begin(), rdi_.end());
    SDT_REQUIRE(mn > 0.0, SDT_EINVAL, "particle radii must be positive");
    // mono path only when every radius is exactly 1 (src/mob/mob_spme.cc:106)
    mono_ = (mx == 1.0 && mn == 1.0);
    if (prec_ == kFP32) {
        spme32_.reset(new SpmeGpu<float>());
        spme32_->create(npos_, box_, xi_, K_, porder_, !mono_);
    } else {
        spme64_.reset(new SpmeGpu<double>());
        spme64_->create(npos_, box_, xi_, K_, porder_, !mono_);
    }
    SDT_REQUIRE(pairs_.init(npos_, rdi_.data(), box_, rmax_, 'R') > 0, SDT_EINVAL,
                "failed to initialise the real-space pair list");
    d_pos_.reserve(3 * (size_t)npos_);
    d_rdi_.reserve(npos_);
    d_pos_s_.reserve(3 * (size_t)npos_);
    d_rdi_s_.reserve(npos_);
    d_key_.reserve(npos_);
    d_perm_.reserve(npos_);
    d_iperm_.reserve(npos_);
    const int nt = prec_ == kFP32 ? spme32_->ntile : spme64_->ntile;
    d_tstart_.reserve((size_t)nt * nt * nt + 1);
    inited_ = true;
}

void MobSpmeGpu::set_tuning(const SpmeTuning& t, int b) {
    SDT_REQUIRE(inited_, SDT_ESTATE, "set_tuning called before init");
    const int w = SpmeGpu<float>::width_class(b);
    if (spme32_) spme32_->set_tuning(t, w);
    else spme64_->set_tuning(t, w);
}

SpmeTuning MobSpmeGpu::tuning(int b) const {
    const int w = SpmeGpu<float>::width_class(b);
    if (spme32_) return spme32_->tune[w];
    if (spme64_) return spme64_->tune[w];
    return SpmeTuning();
}

void MobSpmeGpu::set_timer(StageTimer* t) {
    timer_ = t;
    if (spme32_) spme32_->timer = t;
    if (spme64_) spme64_->timer = t;
}

void MobSpmeGpu::update(const double* pos, const double* rdi, bool dev) {
    SDT_REQUIRE(inited_, SDT_ESTATE, "update called before init");
    SDT_REQUIRE(pos != nullptr, SDT_EINVAL, "pos is null");
    hipStream_t s = stream_;
    StageScope scope(timer_, 8, s);
    const hipMemcpyKind kind = dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    SDT_HIP_CHECK(hipMemcpyAsync(d_pos_.get(), pos, sizeof(double) * 3 * npos_, kind, s));
    const double* r = rdi ? rdi : rdi_.data();
    const hipMemcpyKind rkind = (rdi && dev) ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    SDT_HIP_CHECK(hipMemcpyAsync(d_rdi_.get(), r, sizeof(double) * npos_, rkind, s));
    int nt;
    if (prec_ == kFP32) {
        spme32_->prepare(d_pos_.get(), d_key_.get(), s);
        nt = spme32_->ntile;
    } else {
        spme64_->prepare(d_pos_.get(), d_key_.get(), s);
        nt = spme64_->ntile;
    }
    bucket_sort(d_key_.get(), npos_, nt * nt * nt, d_perm_.get(), d_tstart_.get(), d_iperm_.get(),
                sc_, s);
    if (prec_ == kFP32)
        spme32_->gather_sorted(d_perm_.get(), d_pos_.get(), d_rdi_.get(), d_tstart_.get(), s);
    else
        spme64_->gather_sorted(d_perm_.get(), d_pos_.get(), d_rdi_.get(), d_tstart_.get(), s);
    hipLaunchKernelGGL(gather_sorted_kernel, dim3(ceil_div(npos_, 256)), dim3(256), 0, s, npos_,
                       d_perm_.get(), d_pos_.get(), d_rdi_.get(), d_pos_s_.get(), d_rdi_s_.get());
    SDT_HIP_CHECK(hipGetLastError());
    // real-space pairs in sorted order: rows by perm, columns through iperm
    real_nnz_ = pairs_.build(d_pos_.get(), d_perm_.get(), d_iperm_.get(), false, s);
    ovl_.build(npos_, pairs_.d_rowptr.get(), pairs_.d_colidx.get(), d_pos_s_.get(),
               d_rdi_s_.get(), box_, s);
    updated_ = true;
}

template <typename T>
void MobSpmeGpu::apply_as(SpmeGpu<T>* spme, int mode, int b, double alpha, const double* X,
                          double beta, double* Y) {
    hipStream_t s = stream_;
    const int* rp = pairs_.d_rowptr.get();
    const int* ci = pairs_.d_colidx.get();
    // the column-blocked copy of X the real space gathers; the spread reads
    // the same copy when its blocks are as wide
    const int rc = real_block_cols<T>(b);
    if (rc > 0) d_xblk_.reserve(ceil_div(real_scratch_bytes<T>(npos_, b), sizeof(double)));
    T* xb = reinterpret_cast<T*>(d_xblk_.get());
    bool copied = false;
    // wide full products (the Lanczos batches): the real space runs on the
    // side stream while the z and inverse transforms run here (it is bound
    // by gather latency; started after the forward transform, which streams
    // HBM at its fastest and slowed most beside it: K = 128, 100 columns,
    // 7.5 against 7.7 ms when it started after the spread); it writes
    // Y = beta Y + alpha M_real X first and the interpolation then adds the
    // reciprocal part. Not under graph capture (single-column tails stay on
    // one stream).
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    SDT_HIP_CHECK(hipStreamIsCapturing(s, &cap));
    const bool two_streams = mode == 0 && rc > 0 && b >= 8 && cap == hipStreamCaptureStatusNone;
    if (fill_y && !two_streams) {
        StageScope scope(timer_, 6, s);
        fill_y(s);
    }
    if (two_streams) {
        const bool share = spme->blocked_cols(b) == rc;
        {
            StageScope scope(timer_, 9, s, b);
            real_block_copy<T>(npos_, b, X, xb, s);
        }
        // the reciprocal pipeline of SpmeGpu::apply, stage by stage, with the
        // fork after the forward transform and the join before the
        // interpolation
        StageTimer* tm = spme->timer;
        spme->reserve_rhs(b);
        { StageScope sc(tm, 0, s, b); spme->spread(b, X, s, share ? xb : nullptr); }
        { StageScope sc(tm, 1, s, b); spme->fft_forward(b, s); }
        SDT_HIP_CHECK(hipEventRecord(ev_fork_, s));
        SDT_HIP_CHECK(hipStreamWaitEvent(side_, ev_fork_, 0));
        if (fill_y) {
            StageScope scope(timer_, 6, side_);
            fill_y(side_);
        }
        {
            StageScope scope(timer_, 5, side_, b);
            real_matvec<T>(npos_, rp, ci, d_pos_s_.get(), d_rdi_s_.get(), box_, xi_, b, alpha, X,
                           beta, Y, d_xblk_.get(), true, side_);
            real_overlap(ovl_, d_pos_s_.get(), d_rdi_s_.get(), box_, xi_, b, alpha, X, Y, side_);
        }
        SDT_HIP_CHECK(hipEventRecord(ev_join_, side_));
        { StageScope sc(tm, 2, s, b); spme->influence(b, s); }
        { StageScope sc(tm, 3, s, b); spme->fft_backward(b, s); }
        SDT_HIP_CHECK(hipStreamWaitEvent(s, ev_join_, 0));
        { StageScope sc(tm, 4, s, b); spme->interp(b, alpha, 1.0, Y, s); }
        return;
    }
    if (mode == 0 || mode == 2) {
        const bool share = rc > 0 && spme->blocked_cols(b) == rc;
        spme->apply(b, alpha, X, beta, Y, s, share ? xb : nullptr, share);
        copied = share;
    }
    if (mode == 0 || mode == 1) {
        StageScope scope(timer_, 5, s, b);
        const double rb = (mode == 0) ? 1.0 : beta;
        real_matvec<T>(npos_, rp, ci, d_pos_s_.get(), d_rdi_s_.get(), box_, xi_, b, alpha, X, rb, Y,
                       d_xblk_.get(), copied, s);
        real_overlap(ovl_, d_pos_s_.get(), d_rdi_s_.get(), box_, xi_, b, alpha, X, Y, s);
    }
}

void MobSpmeGpu::apply_packed(int mode, int b, double alpha, const double* X, double beta,
                              double* Y) {
    SDT_REQUIRE(updated_, SDT_ESTATE, "mulvector called before update");
    SDT_REQUIRE(b >= 1, SDT_EINVAL, "at least one right-hand side");
    if (prec_ == kFP32) apply_as<float>(spme32_.get(), mode, b, alpha, X, beta, Y);
    else apply_as<double>(spme64_.get(), mode, b, alpha, X, beta, Y);
}

void MobSpmeGpu::mul_vector(int mode, int nrhs, double alpha, int ldf, const double* f,
                            double beta, int ldv, double* v, bool dev) {
    mul_vector_impl(mode, nrhs, alpha, ldf, f, beta, ldv, v, dev, true);
}

void MobSpmeGpu::enqueue_mul_vector(int mode, int nrhs, double alpha, int ldf, const double* d_f,
                                    double beta, int ldv, double* d_v) {
    mul_vector_impl(mode, nrhs, alpha, ldf, d_f, beta, ldv, d_v, true, false);
}

void MobSpmeGpu::mul_vector_impl(int mode, int nrhs, double alpha, int ldf, const double* f,
                                 double beta, int ldv, double* v, bool dev, bool sync) {
    SDT_REQUIRE(updated_, SDT_ESTATE, "mulvector called before update");
    SDT_REQUIRE(mode >= 0 && mode <= 2, SDT_EINVAL, "mode must be full, real or recip");
    SDT_REQUIRE(nrhs >= 1, SDT_EINVAL, "num_rhs must be positive");
    SDT_REQUIRE(ldf >= 3 * npos_ && ldv >= 3 * npos_, SDT_EINVAL,
                "leading dimensions must be at least 3*npos");
    SDT_REQUIRE(f != nullptr && v != nullptr, SDT_EINVAL, "null vector");
    hipStream_t s = stream_;
    const int chunk = launch_width_;
    const int bmax = std::min(nrhs, chunk);
    d_X_.reserve((size_t)3 * npos_ * bmax);
    d_Y_.reserve((size_t)3 * npos_ * bmax);
    const double* df = f;
    double* dv = v;
    if (!dev) {
        d_io_f_.reserve(span(3 * npos_, nrhs, ldf));
        d_io_v_.reserve(span(3 * npos_, nrhs, ldv));
        SDT_HIP_CHECK(hipMemcpyAsync(d_io_f_.get(), f, sizeof(double) * span(3 * npos_, nrhs, ldf),
                                     hipMemcpyHostToDevice, s));
        if (beta != 0.0)
            SDT_HIP_CHECK(hipMemcpyAsync(d_io_v_.get(), v,
                                         sizeof(double) * span(3 * npos_, nrhs, ldv),
                                         hipMemcpyHostToDevice, s));
        df = d_io_f_.get();
        dv = d_io_v_.get();
    }
    for (int r0 = 0; r0 < nrhs; r0 += chunk) {
        const int b = std::min(chunk, nrhs - r0);
        pack(npos_, b, d_perm_.get(), ldf, df + (size_t)r0 * ldf, d_X_.get(), s);
        if (beta != 0.0) pack(npos_, b, d_perm_.get(), ldv, dv + (size_t)r0 * ldv, d_Y_.get(), s);
        apply_packed(mode, b, alpha, d_X_.get(), beta, d_Y_.get());
        unpack(npos_, b, d_perm_.get(), ldv, d_Y_.get(), dv + (size_t)r0 * ldv, s);
    }
    if (!dev) {
        // copy back only the columns (the gaps between columns are untouched)
        SDT_HIP_CHECK(hipMemcpy2DAsync(v, sizeof(double) * ldv, d_io_v_.get(), sizeof(double) * ldv,
                                       sizeof(double) * 3 * npos_, nrhs, hipMemcpyDeviceToHost, s));
    }
    if (sync || !dev) SDT_HIP_CHECK(hipStreamSynchronize(s));
}

// ----------------------------------------------------------- Lanczos -------
// BrwnLanczos::Compute (src/brwn/brwn_lanczos.cc:279-304) splits the
// right-hand sides into groups of max_nrhs and runs BlockLanczos
// (brwn_lanczos.cc:36-221) on each group in turn. Here the groups of a batch
// run in lockstep: one mobility application per iteration covers every
// column of the batch, the small per-group products (Gram, projection,
// CholQR) are batched kernels over column groups, and the per-group
// convergence test T^{1/2} E_1 runs on host threads while the device computes
// the next iteration.
BrwnLanczosGpu::BrwnLanczosGpu(int dim, int max_iters, int max_nrhs, double tol)
    : dim_(dim), max_iters_(max_iters), max_nrhs_(max_nrhs), tol_(tol) {}

void BrwnLanczosGpu::init() {
    // src/brwn/brwn_lanczos.cc:224-275
    SDT_REQUIRE(dim_ > 0, SDT_EINVAL, "dimension must be positive");
    SDT_REQUIRE(max_iters_ > 0, SDT_EINVAL, "max_iters must be positive");
    SDT_REQUIRE(max_nrhs_ > 0, SDT_EINVAL, "max_nrhs must be positive");
    SDT_REQUIRE(tol_ > 0.0, SDT_EINVAL, "tolerance must be positive");
    if (max_nrhs_ > dim_ / 3) max_nrhs_ = dim_ / 3;
    // groups are at most 16 columns wide on the device (one 16 x 16 Gram
    // tile per group); wider groups would change the reference's grouping
    SDT_REQUIRE(max_nrhs_ <= 16, SDT_EINVAL,
                "max_nrhs must be at most 16 on the device (groups of max_nrhs columns)");
    d_S_.reserve(8192);
    inited_ = true;
}

namespace {

// host state of one column group
struct LzGroup {
    int c0 = 0, b = 0;
    bool done = false;
    int it = -1;                  // iteration of convergence, -1 if none
    double normz = 0.0;           // ||Z_g||_F
    std::vector<double> R, H, B;  // b x b row-major: Z = V_0 R, H_i, B_i
    std::vector<double> Tb;       // block tridiagonal T in band storage
    std::vector<double> S, D, Dprev, Dfin;
    int d_iter = -1;              // iteration that D belongs to
    bool met = false;             // the last evaluation met the tolerance
    BandSqrt bsq;
};

// D_i = T_i^{1/2} E_1 R and the reference's criterion
// ||Y_i - Y_{i-1}||_F / ||Z||_F < tol (src/brwn/brwn_lanczos.cc:179-207),
// where Y_i - Y_{i-1} = V (D_i - [D_{i-1}; 0]) and V has orthonormal columns.
void lz_eval(LzGroup& g, int iter, double tol) {
    const int b = g.b, m = (iter + 1) * b;
    g.S.resize((size_t)m * b);
    g.bsq.compute_band(m, b, g.Tb.data(), 2 * b + 1, b, g.S.data());
    std::swap(g.Dprev, g.D);
    const int prev_iter = g.d_iter;
    g.D.assign((size_t)m * b, 0.0);
    for (int i = 0; i < m; i++)
        for (int c = 0; c < b; c++) {
            double acc = 0.0;
            for (int k = 0; k <= c; k++) acc += g.S[(size_t)i * b + k] * g.R[k * b + c];
            g.D[(size_t)i * b + c] = acc;
        }
    g.d_iter = iter;
    g.met = false;
    if (prev_iter == iter - 1 && iter >= 1) {
        double dn = 0.0;
        for (size_t t = 0; t < g.D.size(); t++) {
            const double prev = t < g.Dprev.size() ? g.Dprev[t] : 0.0;
            const double d = g.D[t] - prev;
            dn += d * d;
        }
        g.met = sqrt(dn) / g.normz < tol;
    }
}

// runs f(0 .. n-1) on up to max_threads host threads
template <typename F>
void parallel_for(int n, int max_threads, F f) {
    const int nt = std::max(1, std::min({n, max_threads, (int)std::thread::hardware_concurrency()}));
    if (nt == 1) {
        for (int i = 0; i < n; i++) f(i);
        return;
    }
    std::vector<std::thread> th;
    for (int t = 0; t < nt; t++)
        th.emplace_back([&, t] {
            for (int i = t; i < n; i += nt) f(i);
        });
    for (auto& x : th) x.join();
}

}  // namespace

// CholQR2 of every column group of the packed n x btot block W in place:
// W_g = Q_g R_g, R_g = R2 R1 stored into (g.*dst). Done groups are zeroed.
// Returns false in ok[g] when W_g is numerically rank deficient (the group's
// columns are then zeroed). trace (optional) receives tr(W_g^T W_g).
static void cholqr2_groups(long long n, int btot, int bs, double* W,
                           std::vector<LzGroup>& grp, std::vector<double> LzGroup::*dst,
                           std::vector<char>& ok, std::vector<double>* trace, DevBuf<double>& dG,
                           DevBuf<double>& dS, DevBuf<double>& part, hipStream_t s,
                           StageTimer* timer) {
    const int ng = (int)grp.size();
    std::vector<double> G((size_t)ng * 256), hS((size_t)ng * bs * bs);
    std::vector<std::vector<double>> R1(ng);
    ok.assign(ng, 1);
    for (int pass = 0; pass < 2; pass++) {
        { StageScope sc(timer, 6, s); gram_groups(n, btot, bs, W, W, dG.get(), part, s); }
        SDT_HIP_CHECK(hipMemcpyAsync(G.data(), dG.get(), sizeof(double) * G.size(),
                                     hipMemcpyDeviceToHost, s));
        SDT_HIP_CHECK(hipStreamSynchronize(s));
        std::fill(hS.begin(), hS.end(), 0.0);
        for (int gi = 0; gi < ng; gi++) {
            LzGroup& g = grp[gi];
            const int b = g.b;
            if (g.done || !ok[gi]) continue;
            std::vector<double> Gg((size_t)b * b), Rp((size_t)b * b), Ri((size_t)b * b);
            for (int i = 0; i < b; i++)
                for (int j = 0; j < b; j++) {
                    // symmetrise
                    Gg[i * b + j] = 0.5 * (G[(size_t)gi * 256 + i * 16 + j] +
                                           G[(size_t)gi * 256 + j * 16 + i]);
                }
            if (pass == 0 && trace) {
                double tr = 0.0;
                for (int i = 0; i < b; i++) tr += Gg[i * b + i];
                (*trace)[gi] = tr;
            }
            if (!chol_upper(b, Gg.data(), Rp.data())) {
                ok[gi] = 0;
                continue;
            }
            inv_upper(b, Rp.data(), Ri.data());
            for (int i = 0; i < b; i++)
                for (int j = 0; j < b; j++) hS[((size_t)gi * bs + i) * bs + j] = Ri[i * b + j];
            if (pass == 0) R1[gi] = Rp;
            else mul_upper(b, Rp.data(), R1[gi].data(), (g.*dst).data());
        }
        SDT_HIP_CHECK(hipMemcpyAsync(dS.get(), hS.data(), sizeof(double) * hS.size(),
                                     hipMemcpyHostToDevice, s));
        StageScope sc(timer, 6, s);
        tsmm_groups1(n, btot, bs, W, dS.get(), 1.0, 0.0, W, s);   // in place
    }
}

void BrwnLanczosGpu::lanczos_batch(MobGpu* mob, int btot, int gbase, const double* Z,
                                   double* Y, std::vector<int>& iters) {
    const long long n = dim_;
    hipStream_t s = mob->stream();
    const int bs = max_nrhs_;
    const int ng = (btot + bs - 1) / bs;
    const size_t blk = (size_t)n * btot;
    d_V_.reserve((size_t)(max_iters_ + 1) * blk);
    d_G_.reserve((size_t)ng * 256);
    double* V = d_V_.get();
    const size_t sset = (size_t)ng * bs * bs;   // one coefficient block per group
    std::vector<LzGroup> grp(ng);
    for (int gi = 0; gi < ng; gi++) {
        LzGroup& g = grp[gi];
        g.c0 = gi * bs;
        g.b = std::min(bs, btot - g.c0);
        g.R.assign((size_t)g.b * g.b, 0.0);
        g.H.assign((size_t)g.b * g.b, 0.0);
        g.B.assign((size_t)g.b * g.b, 0.0);
        g.Tb.assign((size_t)(max_iters_ + 1) * g.b * (2 * g.b + 1), 0.0);
    }
    SDT_HIP_CHECK(hipMemcpyAsync(V, Z, sizeof(double) * blk, hipMemcpyDeviceToDevice, s));
    std::vector<char> ok;
    std::vector<double> tr(ng, 0.0);
    cholqr2_groups(n, btot, bs, V, grp, &LzGroup::R, ok, &tr, d_G_, d_S_, d_part_, s,
                   timer);
    for (int gi = 0; gi < ng; gi++) {
        if (!ok[gi])
            throw Error{SDT_ECONVERGE, "block Lanczos: the right-hand sides are rank deficient"};
        grp[gi].normz = sqrt(tr[gi]);
    }
    auto upload_S = [&](const std::vector<double>& h) {
        SDT_HIP_CHECK(hipMemcpyAsync(d_S_.get(), h.data(), sizeof(double) * h.size(),
                                     hipMemcpyHostToDevice, s));
    };
    // The coefficient blocks of an iteration (H_i, the CholQR factors, B_i)
    // are formed on the device; their records are copied to page-locked
    // memory and read one iteration later, while the device runs the next
    // iteration: the evaluation of iteration i - 1 decides which groups
    // iterate in i + 1, as with a host round trip per iteration.
    const size_t rsz = (size_t)ng * kLzRec;
    d_rec_.reserve(2 * rsz);
    d_R1_.reserve((size_t)ng * 256);
    d_mask_.reserve(ng);
    h_rec_.reserve((size_t)max_iters_ * rsz);
    h_mask_.reserve((size_t)max_iters_ * ng);
    struct Events {
        hipEvent_t e[2] = {nullptr, nullptr};
        Events() {
            for (hipEvent_t& x : e) SDT_HIP_CHECK(hipEventCreateWithFlags(&x, hipEventDisableTiming));
        }
        ~Events() {
            for (hipEvent_t x : e)
                if (x) (void)hipEventDestroy(x);
        }
    } ev;
    // reads the records of iteration pi into the tridiagonal blocks and
    // evaluates the groups; true when every group is done
    auto process = [&](int pi) {
        SDT_HIP_CHECK(hipEventSynchronize(ev.e[pi & 1]));
        const double* hr = h_rec_.get() + (size_t)pi * rsz;
        for (int gi = 0; gi < ng; gi++) {
            LzGroup& g = grp[gi];
            if (g.done) continue;
            const double* r = hr + (size_t)gi * kLzRec;
            const int b = g.b, o = pi * b, lt = 2 * b + 1;
            ok[gi] = r[512] != 0.0;
            for (int i = 0; i < b; i++)
                for (int j = 0; j < b; j++) {
                    g.H[i * b + j] = r[i * 16 + j];
                    g.B[i * b + j] = r[256 + i * 16 + j];
                }
            // T: H_i on the diagonal, B_i below it (and B_i^T above)
            auto T = [&](int rr, int cc) -> double& { return g.Tb[(size_t)rr * lt + (cc - rr + b)]; };
            for (int c = 0; c < b; c++)
                for (int a = c; a < b; a++) {
                    // H(a, c) for a >= c from the row-major Gram (V_i^T W)[a][c]
                    const double hv = g.H[a * b + c];
                    T(o + a, o + c) = hv;
                    T(o + c, o + a) = hv;
                }
            if (ok[gi])
                for (int a = 0; a < b; a++)
                    for (int c = a; c < b; c++) {
                        const double bv = g.B[a * b + c];   // B(a, c), upper
                        T(o + b + a, o + c) = bv;
                        T(o + c, o + b + a) = bv;
                    }
        }
        // an invariant subspace was found: the iterate of pi is exact
        for (int gi = 0; gi < ng; gi++) {
            LzGroup& g = grp[gi];
            if (g.done || ok[gi]) continue;
            lz_eval(g, pi, tol_);
            g.done = true;
            g.it = pi;
            g.Dfin = g.D;
        }
        std::vector<int> todo;
        for (int gi = 0; gi < ng; gi++)
            if (!grp[gi].done) todo.push_back(gi);
        const double tol = tol_;
        parallel_for((int)todo.size(), eval_threads, [&](int t) { lz_eval(grp[todo[t]], pi, tol); });
        bool all = true;
        for (LzGroup& g : grp) {
            if (!g.done && g.d_iter == pi && g.met) {
                g.done = true;
                g.it = pi;
                g.Dfin = g.D;
            }
            all = all && g.done;
        }
        return all;
    };
    int iter;
    bool all = false;
    for (iter = 0; iter < max_iters_ && !all; iter++) {
        if (cancel && cancel->load()) {
            SDT_HIP_CHECK(hipStreamSynchronize(s));
            throw Error{SDT_ESTATE, "cancelled"};
        }
        double* Vi = V + (size_t)iter * blk;
        double* W = V + (size_t)(iter + 1) * blk;
        double* rec = d_rec_.get() + (size_t)(iter & 1) * rsz;
        const double* prev = d_rec_.get() + (size_t)((iter + 1) & 1) * rsz;
        int* hm = h_mask_.get() + (size_t)iter * ng;
        for (int gi = 0; gi < ng; gi++) hm[gi] = grp[gi].done ? 0 : 1;
        SDT_HIP_CHECK(hipMemcpyAsync(d_mask_.get(), hm, sizeof(int) * ng, hipMemcpyHostToDevice, s));
        double beta = 0.0;
        // W = V_{i-1} B_{i-1}^T: its coefficients here, the block product
        // inside the matvec's fill of W (beside the transforms of a wide SPME
        // product) or here before a column subset is taken
        auto fill_w = [&](hipStream_t hs) {
            tsmm_groups1(n, btot, bs, V + (size_t)(iter - 1) * blk, d_S_.get(), 1.0, 0.0, W, hs);
        };
        if (iter > 0) {
            StageScope sc(timer, 6, s);
            lz_coef(ng, bs, btot, d_mask_.get(), nullptr, prev, nullptr, d_S_.get(), 1, s);
            beta = -1.0;
        }
        // W = M V_i - V_{i-1} B^T, one application for the columns of
        // the groups still iterating (the columns of finished groups stay
        // zero: their coefficient blocks are zero from here on)
        std::vector<int> act;
        for (const LzGroup& g : grp)
            if (!g.done)
                for (int c = 0; c < g.b; c++) act.push_back(g.c0 + c);
        if ((int)act.size() == btot || 4 * (int)act.size() > 3 * btot) {
            struct FillReset {
                MobGpu* m;
                ~FillReset() { m->fill_y = nullptr; }
            } reset{mob};
            if (iter > 0) mob->fill_y = fill_w;
            mob->apply_packed(0, btot, 1.0, Vi, beta, W);
        } else {
            if (iter > 0) {
                StageScope sc(timer, 6, s);
                fill_w(s);
            }
            const int na = (int)act.size();
            d_cols_.reserve(na);
            d_Xa_.reserve((size_t)n * na);
            d_Wa_.reserve((size_t)n * na);
            SDT_HIP_CHECK(hipMemcpyAsync(d_cols_.get(), act.data(), sizeof(int) * na,
                                         hipMemcpyHostToDevice, s));
            column_subset(n, btot, na, d_cols_.get(), Vi, d_Xa_.get(), true, s);
            if (beta != 0.0) column_subset(n, btot, na, d_cols_.get(), W, d_Wa_.get(), true, s);
            mob->apply_packed(0, na, 1.0, d_Xa_.get(), beta, d_Wa_.get());
            column_subset(n, btot, na, d_cols_.get(), d_Wa_.get(), W, false, s);
        }
        // H = V_i^T W ; W -= V_i H
        {
            StageScope sc(timer, 6, s);
            gram_groups(n, btot, bs, Vi, W, d_G_.get(), d_part_, s);
            lz_coef(ng, bs, btot, d_mask_.get(), d_G_.get(), nullptr, rec, d_S_.get(), 0, s);
            tsmm_groups1(n, btot, bs, Vi, d_S_.get(), -1.0, 1.0, W, s);
        }
        if (reorth) {
            // full reorthogonalisation against V_0 .. V_i (classical
            // Gram-Schmidt, one pass)
            std::vector<double> C((size_t)(iter + 1) * ng * 256);
            for (int k = 0; k <= iter; k++) {
                gram_groups(n, btot, bs, V + (size_t)k * blk, W, d_G_.get(), d_part_, s);
                SDT_HIP_CHECK(hipMemcpyAsync(&C[(size_t)k * ng * 256], d_G_.get(),
                                             sizeof(double) * ng * 256, hipMemcpyDeviceToHost,
                                             s));
            }
            SDT_HIP_CHECK(hipStreamSynchronize(s));
            const int per = std::max(1, 8192 / (int)sset);
            for (int k0 = 0; k0 <= iter; k0 += per) {
                const int nk = std::min(per, iter + 1 - k0);
                std::vector<double> hC((size_t)nk * sset, 0.0);
                for (int gi = 0; gi < ng; gi++) {
                    if (grp[gi].done) continue;
                    for (int k = 0; k < nk; k++)
                        for (int i = 0; i < bs; i++)
                            for (int j = 0; j < bs; j++)
                                hC[(((size_t)gi * nk + k) * bs + i) * bs + j] =
                                    C[((size_t)(k0 + k) * ng + gi) * 256 + i * 16 + j];
                }
                upload_S(hC);
                tsmm_groups(n, btot, bs, nk, V + (size_t)k0 * blk, (long long)blk, d_S_.get(),
                            -1.0, 1.0, W, s);
            }
        }
        // [V_{i+1}, B_i] = QR(W), CholQR2
        {
            StageScope sc(timer, 6, s);
            for (int pass = 0; pass < 2; pass++) {
                gram_groups(n, btot, bs, W, W, d_G_.get(), d_part_, s);
                lz_chol(ng, bs, btot, d_mask_.get(), d_G_.get(), d_R1_.get(), rec, d_S_.get(), pass,
                        s);
                tsmm_groups1(n, btot, bs, W, d_S_.get(), 1.0, 0.0, W, s);   // in place
            }
        }
        SDT_HIP_CHECK(hipMemcpyAsync(h_rec_.get() + (size_t)iter * rsz, rec, sizeof(double) * rsz,
                                     hipMemcpyDeviceToHost, s));
        SDT_HIP_CHECK(hipEventRecord(ev.e[iter & 1], s));
        if (iter > 0) all = process(iter - 1);
    }
    // the last iteration's records when the loop ran out of iterations
    if (!all) process(iter - 1);
    // groups that did not converge keep the last iterate (brwn_lanczos.cc:208-221)
    for (LzGroup& g : grp) {
        if (g.done) continue;
        if (g.d_iter != max_iters_ - 1) lz_eval(g, max_iters_ - 1, tol_);
        g.Dfin = g.D;
        g.it = -1;
    }
    int used = 0;
    for (const LzGroup& g : grp)
        used = std::max(used, g.it >= 0 ? g.it + 1 : max_iters_);
    // Y = [V_0 .. V_{used-1}] D, in chunks that fit the LDS coefficient tile
    const int per = std::max(1, 8192 / (int)sset);
    for (int k0 = 0; k0 < used; k0 += per) {
        const int nk = std::min(per, used - k0);
        std::vector<double> hD((size_t)nk * sset, 0.0);
        for (int gi = 0; gi < ng; gi++) {
            const LzGroup& g = grp[gi];
            const int rows = (int)(g.Dfin.size() / g.b);
            for (int k = 0; k < nk; k++)
                for (int i = 0; i < g.b; i++) {
                    const int r = (k0 + k) * g.b + i;
                    if (r >= rows) continue;
                    for (int j = 0; j < g.b; j++)
                        hD[(((size_t)gi * nk + k) * bs + i) * bs + j] = g.Dfin[(size_t)r * g.b + j];
                }
        }
        upload_S(hD);
        StageScope sc(timer, 6, s);
        tsmm_groups(n, btot, bs, nk, V + (size_t)k0 * blk, (long long)blk, d_S_.get(), 1.0,
                    k0 == 0 ? 0.0 : 1.0, Y, s);
    }
    SDT_HIP_CHECK(hipStreamSynchronize(s));
    for (int gi = 0; gi < ng; gi++) iters.push_back(grp[gi].it);
}

void BrwnLanczosGpu::compute(MobGpu* mob, int num_rhs, int ldz, const double* z, int ldy,
                             double* y, bool dev) {
    SDT_REQUIRE(inited_, SDT_ESTATE, "compute called before init");
    SDT_REQUIRE(mob && mob->updated(), SDT_ESTATE, "mobility not updated");
    SDT_REQUIRE(mob->dim() == dim_, SDT_EINVAL, "mobility dimension mismatch");
    SDT_REQUIRE(num_rhs >= 1 && ldz >= dim_ && ldy >= dim_, SDT_EINVAL, "bad shapes");
    hipStream_t s = mob->stream();
    const int npos = dim_ / 3;
    // columns per lockstep batch: whole groups, at most kMaxBatchCols, and a
    // Krylov basis within half of the free device memory
    size_t free_b = 0, total_b = 0;
    SDT_HIP_CHECK(hipMemGetInfo(&free_b, &total_b));
    const size_t col_bytes = sizeof(double) * (size_t)dim_ * (max_iters_ + 4);
    int cols = (int)std::min<size_t>(kMaxBatchCols, (free_b / 2 + d_V_.bytes()) / col_bytes);
    cols = std::max(max_nrhs_, cols / max_nrhs_ * max_nrhs_);
    const int bcols = std::min(cols, (num_rhs + max_nrhs_ - 1) / max_nrhs_ * max_nrhs_);
    d_Zp_.reserve((size_t)dim_ * bcols);
    d_Yp_.reserve((size_t)dim_ * bcols);
    const double* dz = z;
    double* dy = y;
    if (!dev) {
        d_io_z_.reserve(span(dim_, num_rhs, ldz));
        d_io_y_.reserve(span(dim_, num_rhs, ldy));
        SDT_HIP_CHECK(hipMemcpyAsync(d_io_z_.get(), z, sizeof(double) * span(dim_, num_rhs, ldz),
                                     hipMemcpyHostToDevice, s));
        dz = d_io_z_.get();
        dy = d_io_y_.get();
    }
    iters_.clear();
    for (int irhs = 0; irhs < num_rhs; irhs += bcols) {
        const int btot = std::min(bcols, num_rhs - irhs);
        pack(npos, btot, mob->perm(), ldz, dz + (size_t)irhs * ldz, d_Zp_.get(), s);
        lanczos_batch(mob, btot, irhs / max_nrhs_, d_Zp_.get(), d_Yp_.get(), iters_);
        unpack(npos, btot, mob->perm(), ldy, d_Yp_.get(), dy + (size_t)irhs * ldy, s);
    }
    if (!dev)
        SDT_HIP_CHECK(hipMemcpy2DAsync(y, sizeof(double) * ldy, d_io_y_.get(), sizeof(double) * ldy,
                                       sizeof(double) * dim_, num_rhs, hipMemcpyDeviceToHost, s));
    SDT_HIP_CHECK(hipStreamSynchronize(s));
}

// ------------------------------------------------------------- steric ------
StericGpu::StericGpu(int npos, const double* rdi, double box, double r0, double k0)
    : npos_(npos), box_(box), r0_(r0), k0_(k0) {
    SDT_REQUIRE(npos > 0 && rdi != nullptr, SDT_EINVAL, "npos must be positive and rdi non-null");
    rdi_.assign(rdi, rdi + npos);
}

void StericGpu::init() {
    // src/force/force_steric.cc:77-96
    SDT_REQUIRE(k0_ > 0.0, SDT_EINVAL, "the steric coefficient must be positive");
    SDT_REQUIRE(pairs_.init(npos_, rdi_.data(), box_, r0_, 'S') > 0, SDT_EINVAL,
                "failed to initialise the steric pair list");
    d_pos_.reserve(3 * (size_t)npos_);
    d_f_.reserve(3 * (size_t)npos_);
    inited_ = true;
}

void StericGpu::accumulate(const double* pos, double* f, bool dev, hipStream_t s) {
    SDT_REQUIRE(inited_, SDT_ESTATE, "steric force used before init");
    const double* dp = pos;
    double* df = f;
    if (!dev) {
        SDT_HIP_CHECK(hipMemcpyAsync(d_pos_.get(), pos, sizeof(double) * 3 * npos_,
                                     hipMemcpyHostToDevice, s));
        SDT_HIP_CHECK(hipMemcpyAsync(d_f_.get(), f, sizeof(double) * 3 * npos_,
                                     hipMemcpyHostToDevice, s));
        dp = d_pos_.get();
        df = d_f_.get();
    }
    StageScope scope(timer, 7, s);
    pairs_.steric_direct(dp, r0_, k0_, df, s);
    if (!dev) {
        SDT_HIP_CHECK(hipMemcpyAsync(f, d_f_.get(), sizeof(double) * 3 * npos_,
                                     hipMemcpyDeviceToHost, s));
        SDT_HIP_CHECK(hipStreamSynchronize(s));
    }
}

// ----------------------------------------------------------------- RNG -----
__device__ __forceinline__ void philox_round(unsigned& c0, unsigned& c1, unsigned& c2, unsigned& c3,
                                             unsigned k0, unsigned k1) {
    const unsigned long long p0 = (unsigned long long)0xD2511F53u * c0;
    const unsigned long long p1 = (unsigned long long)0xCD9E8D57u * c2;
    const unsigned hi0 = (unsigned)(p0 >> 32), lo0 = (unsigned)p0;
    const unsigned hi1 = (unsigned)(p1 >> 32), lo1 = (unsigned)p1;
    const unsigned n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
    c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
}

__global__ void gaussian_kernel(unsigned long long seed, unsigned stream, long long n,
                                double* __restrict__ out) {
    const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (2 * q >= n) return;
    unsigned c0 = (unsigned)q, c1 = (unsigned)((unsigned long long)q >> 32), c2 = stream, c3 = 0u;
    unsigned k0 = (unsigned)seed, k1 = (unsigned)(seed >> 32);
#pragma unroll
    for (int r = 0; r < 10; r++) {
        philox_round(c0, c1, c2, c3, k0, k1);
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    const unsigned long long a = (((unsigned long long)c0 << 32) | c1) >> 11;
    const unsigned long long b = (((unsigned long long)c2 << 32) | c3) >> 11;
    const double inv53 = 1.0 / 9007199254740992.0;
    const double u1 = ((double)a + 0.5) * inv53;
    const double u2 = ((double)b + 0.5) * inv53;
    const double r = sqrt(-2.0 * log(u1));
    const double t = 2.0 * M_PI * u2;
    out[2 * q] = r * cos(t);
    if (2 * q + 1 < n) out[2 * q + 1] = r * sin(t);
}

void gaussian_fill_device(unsigned long long seed, unsigned stream, long long n, double* d_out,
                          hipStream_t s) {
    const long long npairs = (n + 1) / 2;
    hipLaunchKernelGGL(gaussian_kernel, dim3(ceil_div(npairs, 256)), dim3(256), 0, s, seed, stream,
                       n, d_out);
    SDT_HIP_CHECK(hipGetLastError());
}

// ----------------------------------------------------------------- BD ------
BDGpu::BDGpu(int npos, const double* pos, const double* rdi, double box, double xi, double rmax,
             int dim, int porder, int precision, int max_iters, int max_nrhs, double tol,
             double r0, double k0, double dt, int mob_interval, unsigned long long seed)
    : npos_(npos), box_(box), dt_(dt), mob_interval_(mob_interval), seed_(seed) {
    SDT_REQUIRE(npos > 0 && pos && rdi, SDT_EINVAL, "bad particle arrays");
    SDT_REQUIRE(mob_interval >= 1, SDT_EINVAL, "mob_interval must be positive");
    SDT_REQUIRE(dt > 0.0, SDT_EINVAL, "dt must be positive");
    h_pos0_.assign(pos, pos + 3 * (size_t)npos);
    rdi_.assign(rdi, rdi + npos);
    steric_on_ = r0 > 0.0 && k0 != 0.0;
    spme_ = new MobSpmeGpu(npos, rdi, box, xi, rmax, dim, porder, precision);
    mob_.reset(spme_);
    brwn_.reset(new BrwnLanczosGpu(3 * npos, max_iters, max_nrhs, tol));
    steric_.reset(new StericGpu(npos, rdi, box, r0, k0));
}

BDGpu::BDGpu(int npos, const double* pos, const double* rdi, double box, double ewald_tol,
             bool chol, int max_iters, int max_nrhs, double tol, double r0, double k0, double dt,
             int mob_interval, unsigned long long seed)
    : npos_(npos), box_(box), dt_(dt), mob_interval_(mob_interval), seed_(seed) {
    SDT_REQUIRE(npos > 0 && pos && rdi, SDT_EINVAL, "bad particle arrays");
    SDT_REQUIRE(mob_interval >= 1, SDT_EINVAL, "mob_interval must be positive");
    SDT_REQUIRE(dt > 0.0, SDT_EINVAL, "dt must be positive");
    h_pos0_.assign(pos, pos + 3 * (size_t)npos);
    rdi_.assign(rdi, rdi + npos);
    steric_on_ = r0 > 0.0 && k0 != 0.0;
    mob_.reset(new MobEwaldGpu(npos, rdi, box, ewald_tol, -1.0));
    brwn_.reset(new BrwnLanczosGpu(3 * npos, max_iters, max_nrhs, tol));
    if (chol) chol_.reset(new BrwnCholGpu(3 * npos));
    steric_.reset(new StericGpu(npos, rdi, box, r0, k0));
    use_graph_ = false;   // the tail's GEMM is a library call: launched eagerly
}

BDGpu::~BDGpu() { release_graph(); }

void BDGpu::release_graph() {
    CaptureLock lk(capture_mutex());
    if (tail_exec_) (void)hipGraphExecDestroy(tail_exec_);
    if (tail_graph_) (void)hipGraphDestroy(tail_graph_);
    tail_exec_ = nullptr;
    tail_graph_ = nullptr;
    graph_block_ = -1;
}

// out = D[:, *col] (the displacement column of the step), then the tail
// advances *col after the position update
__global__ void select_column_kernel(long long n, const double* __restrict__ D, long long ld,
                                     const int* __restrict__ col, double* __restrict__ out) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = D[(size_t)(*col) * ld + i];
}

__global__ void set_counter_kernel(int* c, int v) { *c = v; }
__global__ void next_counter_kernel(int* c) { *c += 1; }

void BDGpu::use_random(int num_vecs, int max_nrhs) {
    SDT_REQUIRE(!d_pos_.get(), SDT_ESTATE, "use_random after init");
    SDT_REQUIRE(!chol_, SDT_EINVAL, "use_random with Cholesky displacements");
    // its own Gaussian stream, as the reference's BrwnRandom has its own
    // RndStream (src/brwn/brwn_random.cc:109)
    rnd_.reset(new BrwnRandomGpu(3 * npos_, num_vecs, max_nrhs, seed_ ^ 0x5678ull));
}

void BDGpu::init() {
    mob_->init();
    if (chol_) chol_->init();
    else if (rnd_) rnd_->init();
    else brwn_->init();
    if (steric_on_) steric_->init();
    if (spme_) spme_->set_timer(&timer_);
    brwn_->timer = &timer_;
    steric_->timer = &timer_;
    ldm_ = 3 * npos_;
    d_pos_.reserve((size_t)ldm_);
    d_Z_.reserve((size_t)ldm_ * mob_interval_);
    d_D_.reserve((size_t)ldm_ * mob_interval_);
    d_f_.reserve((size_t)ldm_);
    d_cur_.reserve((size_t)ldm_);
    d_irhs_.reserve(1);
    SDT_HIP_CHECK(sync_memcpy(d_pos_.get(), h_pos0_.data(), sizeof(double) * ldm_,
                            hipMemcpyHostToDevice));
}

bool BDGpu::begin_step() {
    hipStream_t s = mob_->stream();
    const int irhs = (int)(istep_ % mob_interval_);
    const long long blk = istep_ / mob_interval_;
    // a new interval, or the first step after a start inside an interval
    // (then D of this interval is computed from the current positions)
    if (irhs != 0 && blk == d_block_) return false;
    gaussian_fill_device(seed_, (unsigned)blk, (long long)mob_interval_ * ldm_, d_Z_.get(), s);
    mob_->update(d_pos_.get(), nullptr, true);
    if (chol_) chol_->compute(mob_.get(), mob_interval_, ldm_, d_Z_.get(), ldm_, d_D_.get(), true);
    else if (rnd_) rnd_->compute(mob_.get(), mob_interval_, ldm_, d_Z_.get(), ldm_, d_D_.get(), true);
    else brwn_->compute(mob_.get(), mob_interval_, ldm_, d_Z_.get(), ldm_, d_D_.get(), true);
    d_block_ = blk;
    hipLaunchKernelGGL(set_counter_kernel, dim3(1), dim3(1), 0, s, d_irhs_.get(), irhs);
    SDT_HIP_CHECK(hipGetLastError());
    return true;
}

// disp = dt M f + sqrt(2 dt) D[:, irhs]; pos += disp (simbox_bd.cc:323-339)
void BDGpu::enqueue_tail(hipStream_t s) {
    hipLaunchKernelGGL(select_column_kernel, dim3(ceil_div(ldm_, 256)), dim3(256), 0, s,
                       (long long)ldm_, d_D_.get(), (long long)ldm_, d_irhs_.get(), d_cur_.get());
    SDT_HIP_CHECK(hipGetLastError());
    fill(ldm_, 0.0, d_f_.get(), s);
    if (steric_on_) steric_->accumulate(d_pos_.get(), d_f_.get(), true, s);
    mob_->enqueue_mul_vector(0, 1, dt_, ldm_, d_f_.get(), sqrt(2.0 * dt_), ldm_, d_cur_.get());
    axpby(ldm_, 1.0, d_cur_.get(), 1.0, d_pos_.get(), s);
    hipLaunchKernelGGL(next_counter_kernel, dim3(1), dim3(1), 0, s, d_irhs_.get());
    SDT_HIP_CHECK(hipGetLastError());
    warm_block_ = d_block_;
}

// The first tail of an interval runs eagerly (its buffers then exist at
// their final sizes); the second captures the tail into a graph, which the
// rest of the interval replays. A mobility update can move buffers and
// changes launch sizes, so each interval captures its own graph.
void BDGpu::tail_step() {
    hipStream_t s = mob_->stream();
    if (!use_graph_ || timer_.on || warm_block_ != d_block_) {
        enqueue_tail(s);
        return;
    }
    if (graph_block_ != d_block_) {
        // no other thread's allocation, free or synchronous copy meanwhile
        CaptureLock lk(capture_mutex());
        release_graph();
        SDT_HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        try {
            enqueue_tail(s);
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(s, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        SDT_HIP_CHECK(hipStreamEndCapture(s, &tail_graph_));
        SDT_HIP_CHECK(hipGraphInstantiate(&tail_exec_, tail_graph_, nullptr, nullptr, 0));
        graph_block_ = d_block_;
    }
    SDT_HIP_CHECK(hipGraphLaunch(tail_exec_, s));
}

void BDGpu::advance(int nsteps) {
    hipStream_t s = mob_->stream();
    for (int k = 0; k < nsteps; k++) {
        begin_step();
        tail_step();
        end_step();
    }
    SDT_HIP_CHECK(hipStreamSynchronize(s));
    timer_.collect();
}

void BDGpu::set_step(long long istep) {
    SDT_REQUIRE(istep >= 0, SDT_EINVAL, "step must be non-negative");
    istep_ = istep;
    d_block_ = -1;
    warm_block_ = -1;
    release_graph();
}

void BDGpu::get_positions(double* pos) {
    SDT_HIP_CHECK(sync_memcpy(pos, d_pos_.get(), sizeof(double) * ldm_, hipMemcpyDeviceToHost));
}

}  // namespace sdt
