// GPU mobility (MobSpme), block Lanczos (BrwnLanczos), steric force and the
// BD step driver (BDSimBox::Advance) behind the C ABI.
#pragma once
#include <atomic>
#include <functional>
#include <memory>
#include "block_gpu.hpp"
#include "host_linalg.hpp"
#include "pair_gpu.hpp"
#include "real_gpu.hpp"
#include "sort_gpu.hpp"
#include "spme_gpu.hpp"

namespace sdt {

enum Precision : int { kFP64 = 0, kFP32 = 1 };

// A device mobility (MobBase, src/mob/mob_base.h:19-80) as the Brownian
// solvers see it. Vectors inside the library use the packed layout of the
// mobility's own particle order: X[(3 i + c) b + r] for the particle perm()[i]
// of the caller.
class MobGpu {
  public:
    virtual ~MobGpu() {}
    virtual void init() = 0;
    // positions and radii: host (dev = false) or device pointers
    virtual void update(const double* pos, const double* rdi, bool dev) = 0;
    // v := alpha M f + beta v, reference layout; mode 0 full, 1 real, 2 recip
    virtual void mul_vector(int mode, int nrhs, double alpha, int ldf, const double* f,
                            double beta, int ldv, double* v, bool dev) = 0;
    // the same on device vectors, enqueued on stream() without waiting
    virtual void enqueue_mul_vector(int mode, int nrhs, double alpha, int ldf, const double* d_f,
                                    double beta, int ldv, double* d_v) = 0;
    // packed FP64 vectors (row-major 3N x b) in the order of perm()
    virtual void apply_packed(int mode, int b, double alpha, const double* X, double beta,
                              double* Y) = 0;
    virtual int npos() const = 0;
    int dim() const { return 3 * npos(); }
    virtual hipStream_t stream() const = 0;
    virtual const int* perm() const = 0;
    virtual bool updated() const = 0;
    // Optional work that writes Y's initial contents (the beta Y of the next
    // apply_packed), enqueued by apply_packed on the stream it is given
    // before Y is read: a wide SPME product runs it on its second stream
    // beside the transforms, other products first on stream(). Set by the
    // caller around one apply_packed.
    std::function<void(hipStream_t)> fill_y;
};

// MobSpme (src/mob/mob_spme.h:21-125) with device-resident state.
class MobSpmeGpu : public MobGpu {
  public:
    MobSpmeGpu(int npos, const double* rdi, double box, double xi, double rmax, int dim,
               int porder, int precision);
    ~MobSpmeGpu() override;
    // validation and allocation (src/mob/mob_spme.cc:53-132)
    void init() override;
    void update(const double* pos, const double* rdi, bool dev) override;
    void mul_vector(int mode, int nrhs, double alpha, int ldf, const double* f, double beta,
                    int ldv, double* v, bool dev) override;
    // the same on device vectors, enqueued on stream() without waiting (no
    // host synchronisation: capturable into a graph once its buffers exist)
    void enqueue_mul_vector(int mode, int nrhs, double alpha, int ldf, const double* d_f,
                            double beta, int ldv, double* d_v) override;
    // packed, sorted-order FP64 vectors (row-major 3N x b)
    void apply_packed(int mode, int b, double alpha, const double* X, double beta,
                      double* Y) override;
    int npos() const override { return npos_; }
    hipStream_t stream() const override { return stream_; }
    const int* perm() const override { return d_perm_.get(); }
    bool updated() const override { return updated_; }
    int precision() const { return prec_; }
    long long real_nnz() const { return real_nnz_; }
    int K() const { return K_; }
    int porder() const { return porder_; }
    double box() const { return box_; }
    // timing hooks for the benchmark: last recip/real split is not tracked
    SpmeGpu<float>* spme32() { return spme32_.get(); }
    SpmeGpu<double>* spme64() { return spme64_.get(); }
    void set_timer(StageTimer* t);
    StageTimer* timer() const { return timer_; }
    // launch choices of the reciprocal-space engine for launches of b
    // columns (the narrow class b < 8 or the wide class b >= 8; after init)
    void set_tuning(const SpmeTuning& t, int b);
    SpmeTuning tuning(int b) const;
    double xi() const { return xi_; }
    bool mono() const { return mono_; }
    // right-hand sides per launch of mul_vector (bounds the mesh memory)
    void set_launch_width(int w) { launch_width_ = w; }

  private:
    int launch_width_ = 16;
    void mul_vector_impl(int mode, int nrhs, double alpha, int ldf, const double* f, double beta,
                         int ldv, double* v, bool dev, bool sync);
    template <typename T>
    void apply_as(SpmeGpu<T>* spme, int mode, int b, double alpha, const double* X, double beta,
                  double* Y);
    int npos_, K_, porder_, prec_;
    double box_, xi_, rmax_;
    std::vector<double> rdi_;
    bool mono_ = true, inited_ = false, updated_ = false;
    hipStream_t stream_ = nullptr;
    // the real space of wide launches runs on side_ beside the reciprocal
    // pipeline (fork after the input copy, join before the interpolation)
    hipStream_t side_ = nullptr;
    hipEvent_t ev_fork_ = nullptr, ev_join_ = nullptr;
    std::unique_ptr<SpmeGpu<float>> spme32_;
    std::unique_ptr<SpmeGpu<double>> spme64_;
    PairListGpu pairs_;
    RealOverlap ovl_;
    DevBuf<double> d_xblk_;   // column-blocked copy of the packed input (real space)
    SortScratch sc_;
    DevBuf<double> d_pos_, d_rdi_, d_pos_s_, d_rdi_s_;
    DevBuf<int> d_key_, d_perm_, d_iperm_, d_tstart_;
    DevBuf<double> d_io_f_, d_io_v_, d_X_, d_Y_;
    long long real_nnz_ = 0;
    StageTimer* timer_ = nullptr;
};

// BrwnLanczos (src/brwn/brwn_lanczos.h:45-46) on a device mobility.
class BrwnLanczosGpu {
  public:
    BrwnLanczosGpu(int dim, int max_iters, int max_nrhs, double tol);
    void init();
    // reference layout z/y (host or device); returns the iteration count of
    // every group of max_nrhs columns in iters()
    void compute(MobGpu* mob, int num_rhs, int ldz, const double* z, int ldy, double* y,
                 bool dev);
    const std::vector<int>& iters() const { return iters_; }
    int max_nrhs() const { return max_nrhs_; }
    int reorth = 0;        // 0: none (reference); 1: full FP64 reorthogonalisation
    int eval_threads = 16; // host threads of the per-group convergence tests
    // when set, a solve stops between iterations with SDT_ESTATE "cancelled"
    const std::atomic<bool>* cancel = nullptr;
    StageTimer* timer = nullptr;
    // columns of one lockstep batch
    static constexpr int kMaxBatchCols = 128;

  private:
    // lockstep block Lanczos of the groups of a packed 3N x btot block (row
    // major, sorted particle order); gbase is the index of its first group
    void lanczos_batch(MobGpu* mob, int btot, int gbase, const double* Z, double* Y,
                       std::vector<int>& iters);
    int dim_, max_iters_, max_nrhs_;
    double tol_;
    bool inited_ = false;
    std::vector<int> iters_;
    DevBuf<double> d_V_, d_G_, d_S_, d_part_, d_io_z_, d_io_y_, d_Zp_, d_Yp_;
    DevBuf<double> d_Xa_, d_Wa_;   // active columns of a partly finished batch
    DevBuf<int> d_cols_;
    // device-side iteration steps: per-group records of the last two
    // iterations, the first CholQR factor, the group mask; host copies of
    // every iteration's records and masks
    DevBuf<double> d_rec_, d_R1_;
    DevBuf<int> d_mask_;
    PinnedBuf<double> h_rec_;
    PinnedBuf<int> h_mask_;
};

// Steric force (src/force/force_steric.cc) on the S pair list.
class StericGpu {
  public:
    StericGpu(int npos, const double* rdi, double box, double r0, double k0);
    void init();
    // f := 0 then accumulate (ForceBase::Compute); accumulate adds
    void accumulate(const double* pos, double* f, bool dev, hipStream_t s);
    int npos() const { return npos_; }
    StageTimer* timer = nullptr;

  private:
    int npos_;
    double box_, r0_, k0_;
    std::vector<double> rdi_;
    PairListGpu pairs_;
    DevBuf<double> d_pos_, d_f_;
    bool inited_ = false;
};

// BDSimBox::Advance (src/simbox/simbox_bd.cc:296-366) with device-resident
// positions, Gaussian block, displacements and forces.
class BrwnRandomGpu;
class BrwnCholGpu;
class BDGpu {
  public:
    // SPME mobility (InitMob's spme branch, simbox_bd.cc:103-111) with block
    // Lanczos displacements
    BDGpu(int npos, const double* pos, const double* rdi, double box, double xi, double rmax,
          int dim, int porder, int precision, int max_iters, int max_nrhs, double tol, double r0,
          double k0, double dt, int mob_interval, unsigned long long seed);
    // the dense Ewald mobility (the ewald branch, simbox_bd.cc:99-102, tolerance
    // ewald_tol) with Cholesky displacements (chol = true, BrwnChol) or block
    // Lanczos; FP64
    BDGpu(int npos, const double* pos, const double* rdi, double box, double ewald_tol, bool chol,
          int max_iters, int max_nrhs, double tol, double r0, double k0, double dt,
          int mob_interval, unsigned long long seed);
    ~BDGpu();
    // Brownian displacements by randomised sampling (BrwnRandom) instead of
    // block Lanczos; before init
    void use_random(int num_vecs, int max_nrhs);
    void init();
    void advance(int nsteps);
    void get_positions(double* pos);
    const std::vector<int>& last_iters() const { return brwn_->iters(); }
    // the SPME mobility (null for the dense Ewald mobility)
    MobSpmeGpu* mob() { return spme_; }
    MobGpu* mobility() { return mob_.get(); }
    BrwnLanczosGpu* lanczos() { return brwn_.get(); }
    long long istep() const { return istep_; }
    // continue from step istep (a restart); a start inside a mobility
    // interval computes that interval's D from the positions at the start
    void set_step(long long istep);
    int npos() const { return npos_; }
    int mob_interval() const { return mob_interval_; }
    StageTimer& timer() { return timer_; }
    hipStream_t stream() const { return mob_->stream(); }

    // The parts of one step. begin_step: when the step starts a mobility
    // interval (or follows a restart), the Gaussian block, the mobility
    // update and the interval's displacements (synchronous); returns true if
    // it ran. enqueue_tail: the step's forces, M f and position update on s,
    // with no host synchronisation (the displacement column comes from a
    // device counter that the tail advances), so it can be captured into a
    // graph once its buffers exist. end_step: the step counter.
    bool begin_step();
    void enqueue_tail(hipStream_t s);
    void end_step() { istep_++; }
    // interval whose buffers an eager tail has touched (-1: none yet)
    long long warm_block() const { return warm_block_; }
    long long block() const { return d_block_; }
    // tails as one replayed HIP graph per interval (default on; profiling
    // runs them eagerly, with per-stage events)
    void set_graphs(bool on) { use_graph_ = on; }
    // a set flag stops an interval's displacement solve between iterations
    void set_cancel(const std::atomic<bool>* c) { brwn_->cancel = c; }

  private:
    void tail_step();
    void release_graph();
    int npos_;
    double box_, dt_;
    int mob_interval_;
    unsigned long long seed_;
    std::vector<double> h_pos0_, rdi_;
    std::unique_ptr<MobGpu> mob_;
    MobSpmeGpu* spme_ = nullptr;   // mob_ when it is the SPME mobility
    std::unique_ptr<BrwnLanczosGpu> brwn_;
    std::unique_ptr<BrwnRandomGpu> rnd_;
    std::unique_ptr<BrwnCholGpu> chol_;
    std::unique_ptr<StericGpu> steric_;
    DevBuf<double> d_pos_, d_Z_, d_D_, d_f_, d_cur_;
    DevBuf<int> d_irhs_;       // displacement column of the next tail
    long long istep_ = 0;
    long long d_block_ = -1;   // interval whose displacements d_D_ holds
    long long warm_block_ = -1;
    bool steric_on_ = true;    // r0 > 0 and k0 != 0
    int ldm_ = 0;
    StageTimer timer_;
    bool use_graph_ = true;
    hipGraph_t tail_graph_ = nullptr;
    hipGraphExec_t tail_exec_ = nullptr;
    long long graph_block_ = -1;   // interval the instantiated tail graph belongs to
};

// Philox-4x32-10 Gaussian block on the device (same stream as the oracle):
// element e of stream `stream` -> out[e], n elements.
void gaussian_fill_device(unsigned long long seed, unsigned stream, long long n, double* d_out,
                          hipStream_t s);

}  // namespace sdt
