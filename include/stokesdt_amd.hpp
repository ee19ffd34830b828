/*
 * stokesdt_amd.hpp -- C++ drop-in classes over the C ABI (stokesdt_amd.h).
 *
 * Each class subclasses the reference's own base class, so code written
 * against MobBase / BrwnBase / ForceBase / PairListBase takes these in place
 * of MobSpme, MobEwald, BrwnLanczos, BrwnChol, StericForce and PairListR/S
 * with no other change. Compile with the reference's headers on the include
 * path (src/util, src/mob, src/brwn, src/force, src/pairlist) and link
 * libstokesdt_amd.so and the reference's base-class sources (mob_base.cc,
 * brwn_base.cc, force_base.cc, pair_base.cc). Header-only, C++98.
 *
 * Error behaviour follows the reference: Init() returns false (or 0) with a
 * message on stderr; Update / MulVector / Compute are void, and a failing
 * device call prints sdt_last_error() to stderr (the reference has no error
 * channel there either). last_status() returns the status of the last call.
 */
#ifndef STOKESDT_AMD_HPP_
#define STOKESDT_AMD_HPP_

#include <cstdio>
#include <vector>
#include "brwn_base.h"
#include "force_base.h"
#include "mob_base.h"
#include "pair_base.h"
#include "stokesdt_amd.h"

namespace stokesdt_amd {

inline int report(int rc, const char* where) {
    if (rc != SDT_OK) std::fprintf(stderr, "ERROR: %s: %s\n", where, sdt_last_error());
    return rc;
}

/* A mobility backed by a library handle: what the Brownian solvers below
 * need from the MobBase they are given. */
class MobHip : public stokesdt::MobBase {
  public:
    MobHip() : h_(NULL), status_(SDT_OK) {}
    virtual ~MobHip() {
        if (h_) sdt_mob_destroy(h_);
    }
    sdt_mob_t handle() const { return h_; }
    int last_status() const { return status_; }
    /* MobBase::Update (src/mob/mob_base.h:41-47) */
    virtual void Update(const double* pos, const double* rdi) {
        status_ = report(sdt_mob_update(h_, pos, rdi), "Update");
    }
    /* MobBase::MulVector (src/mob/mob_base.h:52-69): v = alpha M f + beta v */
    virtual void MulVector(const int num_rhs, const double alpha, const int ldf, const double* f,
                           const double beta, const int ldv, double* v) {
        status_ = report(sdt_mob_mulvector(h_, num_rhs, alpha, ldf, f, beta, ldv, v), "MulVector");
    }
    virtual int dim() {
        int d = 0;
        status_ = report(sdt_mob_dim(h_, &d), "dim");
        return d;
    }

  protected:
    bool init_handle() {
        if (!h_) return false;
        status_ = report(sdt_mob_init(h_), "Init");
        return status_ == SDT_OK;
    }
    sdt_mob_t h_;
    int status_;

  private:
    MobHip(const MobHip&);
    void operator=(const MobHip&);
};

/* MobSpme(npos, rdi, box_size, xi, rmax, dim, porder)
 * (src/mob/mob_spme.h:27-91, src/mob/mob_spme.cc:19-262); precision
 * SDT_FP64 (the reference's arithmetic) or SDT_FP32. */
class MobSpmeHip : public MobHip {
  public:
    MobSpmeHip(const int npos, const double* rdi, const double box_size, const double xi,
               const double rmax, const int dim, const int porder, const int precision = SDT_FP64) {
        status_ = report(sdt_mob_spme_create(npos, rdi, box_size, xi, rmax, dim, porder, precision,
                                             &h_),
                         "MobSpmeHip");
    }
    virtual bool Init() { return init_handle(); }
    /* MobSpme::RealMulVector / RecipMulVector (src/mob/mob_spme.h:65-91) */
    void RealMulVector(const int num_rhs, const double alpha, const int ldf, const double* f,
                       const double beta, const int ldv, double* v) {
        status_ = report(sdt_mob_spme_mulvector_real(h_, num_rhs, alpha, ldf, f, beta, ldv, v),
                         "RealMulVector");
    }
    void RecipMulVector(const int num_rhs, const double alpha, const int ldf, const double* f,
                        const double beta, const int ldv, double* v) {
        status_ = report(sdt_mob_spme_mulvector_recip(h_, num_rhs, alpha, ldf, f, beta, ldv, v),
                         "RecipMulVector");
    }
};

/* MobEwald(npos, rdi, box_size, tol[, xi]) (src/mob/mob_ewald.h) with
 * MobMatrix::GetMatrix. */
class MobEwaldHip : public MobHip {
  public:
    MobEwaldHip(const int npos, const double* rdi, const double box_size, const double tol,
                const double xi = -1.0) {
        status_ = report(sdt_mob_ewald_create(npos, rdi, box_size, tol, xi, &h_), "MobEwaldHip");
    }
    virtual bool Init() { return init_handle(); }
    void GetMatrix(const int ldm, double* mat) {
        status_ = report(sdt_mob_ewald_get_matrix(h_, ldm, mat), "GetMatrix");
    }
};

/* BrwnLanczos(dim, max_iters, max_nrhs, tol) (src/brwn/brwn_lanczos.h:45-46):
 * Compute(mob, ...) (src/brwn/brwn_base.h:44-46) takes a MobHip. */
class BrwnLanczosHip : public stokesdt::BrwnBase {
  public:
    BrwnLanczosHip(const int dim, const int max_iters, const int max_nrhs, const double tol)
        : h_(NULL), status_(SDT_OK) {
        status_ = report(sdt_brwn_lanczos_create(dim, max_iters, max_nrhs, tol, &h_),
                         "BrwnLanczosHip");
    }
    virtual ~BrwnLanczosHip() {
        if (h_) sdt_brwn_lanczos_destroy(h_);
    }
    virtual bool Init() {
        if (!h_) return false;
        status_ = report(sdt_brwn_lanczos_init(h_), "Init");
        return status_ == SDT_OK;
    }
    virtual void Compute(stokesdt::MobBase* mob, const int num_rhs, const int ldz, const double* z,
                         const int ldy, double* y) {
        MobHip* m = dynamic_cast<MobHip*>(mob);
        if (!m) {
            std::fprintf(stderr, "ERROR: BrwnLanczosHip::Compute needs a MobHip mobility\n");
            status_ = SDT_EINVAL;
            return;
        }
        const int rc = sdt_brwn_lanczos_compute(h_, m->handle(), num_rhs, ldz, z, ldy, y);
        // non-convergence keeps the result, as the reference does
        status_ = rc == SDT_ECONVERGE ? rc : report(rc, "Compute");
    }
    int last_status() const { return status_; }
    /* iteration of convergence of each group of the last Compute (-1: none) */
    std::vector<int> iters() {
        std::vector<int> out(4096);
        int ng = 0;
        report(sdt_brwn_lanczos_iters(h_, &out[0], (int)out.size(), &ng), "iters");
        out.resize(ng < (int)out.size() ? ng : out.size());
        return out;
    }

  private:
    BrwnLanczosHip(const BrwnLanczosHip&);
    void operator=(const BrwnLanczosHip&);
    sdt_brwn_t h_;
    int status_;
};

/* BrwnChol(dim) (src/brwn/brwn_chol.cc:29-83): needs a MobEwaldHip. */
class BrwnCholHip : public stokesdt::BrwnBase {
  public:
    explicit BrwnCholHip(const int dim) : h_(NULL), status_(SDT_OK) {
        status_ = report(sdt_brwn_chol_create(dim, &h_), "BrwnCholHip");
    }
    virtual ~BrwnCholHip() {
        if (h_) sdt_brwn_chol_destroy(h_);
    }
    virtual bool Init() {
        if (!h_) return false;
        status_ = report(sdt_brwn_chol_init(h_), "Init");
        return status_ == SDT_OK;
    }
    virtual void Compute(stokesdt::MobBase* mob, const int num_rhs, const int ldz, const double* z,
                         const int ldy, double* y) {
        MobHip* m = dynamic_cast<MobHip*>(mob);
        if (!m) {
            std::fprintf(stderr, "ERROR: BrwnCholHip::Compute needs a MobHip mobility\n");
            status_ = SDT_EINVAL;
            return;
        }
        status_ = report(sdt_brwn_chol_compute(h_, m->handle(), num_rhs, ldz, z, ldy, y),
                         "Compute");
    }
    int last_status() const { return status_; }

  private:
    BrwnCholHip(const BrwnCholHip&);
    void operator=(const BrwnCholHip&);
    sdt_brwn_chol_t h_;
    int status_;
};

/* StericForce(npos, rdi, box_size, r0, k0) (src/force/force_steric.cc:56-112);
 * ForceBase::Compute (zero, then Accumulate) comes from the base class. */
class StericForceHip : public stokesdt::ForceBase {
  public:
    StericForceHip(const int npos, const double* rdi, const double box_size, const double r0,
                   const double k0)
        : stokesdt::ForceBase(npos), h_(NULL), status_(SDT_OK) {
        status_ = report(sdt_steric_create(npos, rdi, box_size, r0, k0, &h_), "StericForceHip");
    }
    virtual ~StericForceHip() {
        if (h_) sdt_steric_destroy(h_);
    }
    virtual bool Init() {
        if (!h_) return false;
        status_ = report(sdt_steric_init(h_), "Init");
        return status_ == SDT_OK;
    }
    virtual void Accumulate(const double* pos, const double* rdi, double* f) {
        status_ = report(sdt_steric_accumulate(h_, pos, rdi, f), "Accumulate");
    }
    int last_status() const { return status_; }

  private:
    StericForceHip(const StericForceHip&);
    void operator=(const StericForceHip&);
    sdt_force_t h_;
    int status_;
};

/* PairListR / PairListS (src/pairlist/pair_r.h, pair_s.h) on the device:
 * mode 'R' or 'S'. PairListBase::Init / Build / GetPairs are the base
 * class's; FindPairs fills the base's pair table from the device list (the
 * base's own host cell list is reduced to one cell, which FindPairs does not
 * read). The sets are those of the reference's lists. */
class PairListHip : public stokesdt::PairListBase {
  public:
    PairListHip(const int npos, const double* rdi, const double box_size, const double cutoff,
                const char mode)
        : stokesdt::PairListBase(npos, rdi, box_size, cutoff), h_(NULL), status_(SDT_OK),
          npos_(npos) {
        status_ = report(sdt_pairlist_create(npos, rdi, box_size, cutoff, mode, &h_),
                         "PairListHip");
    }
    virtual ~PairListHip() {
        if (h_) sdt_pairlist_destroy(h_);
    }
    int last_status() const { return status_; }

  private:
    PairListHip(const PairListHip&);
    void operator=(const PairListHip&);
    virtual int NumCells() { return 1; }
    virtual void InitSearchIdx() {}
    virtual int InitPairs(const double*, std::vector<std::vector<int> >* pairs) {
        pairs->assign(npos_, std::vector<int>());
        int64_t est = 0;
        if (!h_) return 0;
        status_ = report(sdt_pairlist_init(h_, &est), "Init");
        return status_ == SDT_OK ? (int)est : 0;
    }
    virtual int FindPairs(const double* pos, const double*, std::vector<std::vector<int> >* pairs) {
        int64_t nnz = 0;
        status_ = report(sdt_pairlist_build(h_, pos, &nnz), "Build");
        if (status_ != SDT_OK) return 0;
        std::vector<int32_t> rowptr(npos_ + 1), colidx(nnz > 0 ? (size_t)nnz : 1);
        status_ = report(sdt_pairlist_get_pairs(h_, &rowptr[0], &colidx[0]), "GetPairs");
        if (status_ != SDT_OK) return 0;
        pairs->assign(npos_, std::vector<int>());
        for (int i = 0; i < npos_; i++)
            (*pairs)[i].assign(colidx.begin() + rowptr[i], colidx.begin() + rowptr[i + 1]);
        return (int)nnz;
    }
    sdt_pairlist_t h_;
    int status_;
    int npos_;
};

}  // namespace stokesdt_amd

#endif /* STOKESDT_AMD_HPP_ */
