// extern "C" boundary of libstokesdt_amd (declared in include/stokesdt_amd.h
// and include/stokesdt_amd_debug.h).
#include <cstring>
#include <string>
#include "../../include/stokesdt_amd.h"
#include "../../include/stokesdt_amd_debug.h"
#include "brwn_random.hpp"
#include "diffusion.hpp"
#include "ensemble.hpp"
#include "ewald_gpu.hpp"
#include "mob_gpu.hpp"
#include "plain_gpu.hpp"
#include "serve.hpp"
#include "tune.hpp"
#include "simbox.hpp"

namespace sdt {
long long rsa_positions(int n, double box, double a, uint64_t seed, long long max_attempts,
                        double* pos);
static thread_local std::string g_err;
void set_error(const std::string& m) { g_err = m; }
const char* last_error() { return g_err.c_str(); }
}  // namespace sdt

using namespace sdt;

struct sdt_mob_s { MobGpu* m; };
struct sdt_brwn_chol_s { BrwnCholGpu* b; };
struct sdt_ensemble_s { EnsembleGpu* e; };
struct sdt_serve_s { ServeGpu* s; ServeParams p; };

// the SPME mobility of a handle (entry points that exist only for MobSpme)
static MobSpmeGpu* spme_of(sdt_mob_t mob) {
    MobSpmeGpu* m = dynamic_cast<MobSpmeGpu*>(mob->m);
    if (!m) throw Error{SDT_EINVAL, "this entry point needs an SPME mobility handle"};
    return m;
}
struct sdt_brwn_s { BrwnLanczosGpu* b; };
struct sdt_brwn_random_s { BrwnRandomGpu* b; };
struct sdt_pairlist_s {
    PairListGpu pl;
    std::vector<double> rdi;
    DevBuf<double> d_pos;
    hipStream_t s = nullptr;
    bool inited = false, built = false;
};
struct sdt_force_s { StericGpu* f; hipStream_t s; };
struct sdt_bd_s { BDGpu* bd; };
struct sdt_simbox_s { SimBoxGpu* sb; };
struct sdt_plain_mob_s {
    PlainMobGpu* m;
    DevBuf<double> d_pos, d_f, d_v;
};
struct sdt_plain_bd_s { PlainBDGpu* bd; };

template <typename F>
static int guard(F&& fn) {
    try {
        fn();
        return SDT_OK;
    } catch (const Error& e) {
        set_error(e.msg);
        return e.code;
    } catch (const std::bad_alloc&) {
        set_error("host allocation failed");
        return SDT_ENOMEM;
    } catch (const std::exception& e) {
        set_error(e.what());
        return SDT_EINVAL;
    }
}

#define SDT_CHECK_HANDLE(h)                                              \
    do {                                                                 \
        if (!(h)) {                                                      \
            set_error("null handle");                                    \
            return SDT_EINVAL;                                           \
        }                                                                \
    } while (0)

extern "C" {

const char* sdt_last_error(void) { return last_error(); }
const char* sdt_version(void) { return "stokesdt_amd 0.1.0 (gfx950)"; }

// ---------------------------------------------------------------- MobSpme --
int sdt_mob_spme_create(int npos, const double* rdi, double box_size, double xi, double rmax,
                        int dim, int porder, int precision, sdt_mob_t* out) {
    if (!out) { set_error("null output handle"); return SDT_EINVAL; }
    *out = nullptr;
    return guard([&] {
        MobSpmeGpu* m = new MobSpmeGpu(npos, rdi, box_size, xi, rmax, dim, porder, precision);
        *out = new sdt_mob_s{m};
    });
}
int sdt_mob_spme_init(sdt_mob_t mob) {
    SDT_CHECK_HANDLE(mob);
    return guard([&] { mob->m->init(); });
}
int sdt_mob_spme_update(sdt_mob_t mob, const double* pos, const double* rdi) {
    SDT_CHECK_HANDLE(mob);
    return guard([&] {
        mob->m->update(pos, rdi, false);
        SDT_HIP_CHECK(hipStreamSynchronize(mob->m->stream()));
    });
}
int sdt_mob_spme_update_dev(sdt_mob_t mob, const double* d_pos, const double* d_rdi) {
    SDT_CHECK_HANDLE(mob);
    return guard([&] {
        mob->m->update(d_pos, d_rdi, true);
        SDT_HIP_CHECK(hipStreamSynchronize(mob->m->stream()));
    });
}
int sdt_mob_spme_mulvector(sdt_mob_t mob, int num_rhs, double alpha, int ldf, const double* f,
                           double beta, int ldv, double* v) {
    SDT_CHECK_HANDLE(mob);
    return guard([&] { mob->m->mul_vector(0, num_rhs, alpha, ldf, f, beta, ldv, v, false); });
}
int sdt_mob_spme_mulvector_real(sdt_mob_t mob, int num_rhs, double alpha, int ldf,
                                const double* f, double beta, int ldv, double* v) {
    SDT_CHECK_HANDLE(mob);
    return guard([&] { mob->m->mul_vector(1, num_rhs, alpha, ldf, f, beta, ldv, v, false); });
}
int sdt_mob_spme_mulvector_recip(sdt_mob_t mob, int num_rhs, double alpha, int ldf,
                                 const double* f, double beta, int ldv, double* v) {
    SDT_CHECK_HANDLE(mob);
    return guard([&] { mob->m->mul_vector(2, num_rhs, alpha, ldf, f, beta, ldv, v, false); });
}
int sdt_mob_spme_mulvector_dev(sdt_mob_t mob, int mode, int num_rhs, double alpha, int ldf,
                               const double* d_f, double beta, int ldv, double* d_v) {
    SDT_CHECK_HANDLE(mob);
    return guard([&] { mob->m->mul_vector(mode, num_rhs, alpha, ldf, d_f, beta, ldv, d_v, true); });
}
int sdt_mob_spme_dim(sdt_mob_t mob, int* dim) {
    SDT_CHECK_HANDLE(mob);
    if (!dim) { set_error("null output"); return SDT_EINVAL; }
    *dim = mob->m->dim();
    return SDT_OK;
}
int sdt_mob_spme_real_nnz(sdt_mob_t mob, int64_t* nnz) {
    SDT_CHECK_HANDLE(mob);
    if (!nnz) { set_error("null output"); return SDT_EINVAL; }
    return guard([&] { *nnz = spme_of(mob)->real_nnz(); });
}
int sdt_mob_spme_destroy(sdt_mob_t mob) {
    if (!mob) return SDT_OK;
    return guard([&] {
        delete mob->m;
        delete mob;
    });
}

// ------------------------------------------------ MobEwald / generic mob --
int sdt_mob_ewald_create(int npos, const double* rdi, double box_size, double tol, double xi,
                         sdt_mob_t* out) {
    if (!out) { set_error("null output handle"); return SDT_EINVAL; }
    *out = nullptr;
    return guard([&] {
        MobEwaldGpu* m = new MobEwaldGpu(npos, rdi, box_size, tol, xi);
        *out = new sdt_mob_s{m};
    });
}
int sdt_mob_ewald_get_matrix(sdt_mob_t mob, int ldm, double* mat) {
    SDT_CHECK_HANDLE(mob);
    return guard([&] {
        MobEwaldGpu* m = dynamic_cast<MobEwaldGpu*>(mob->m);
        SDT_REQUIRE(m != nullptr, SDT_EINVAL, "this entry point needs a dense Ewald mobility handle");
        m->get_matrix(ldm, mat, false);
    });
}
int sdt_mob_init(sdt_mob_t mob) { return sdt_mob_spme_init(mob); }
int sdt_mob_update(sdt_mob_t mob, const double* pos, const double* rdi) {
    return sdt_mob_spme_update(mob, pos, rdi);
}
int sdt_mob_mulvector(sdt_mob_t mob, int num_rhs, double alpha, int ldf, const double* f,
                      double beta, int ldv, double* v) {
    return sdt_mob_spme_mulvector(mob, num_rhs, alpha, ldf, f, beta, ldv, v);
}
int sdt_mob_dim(sdt_mob_t mob, int* dim) { return sdt_mob_spme_dim(mob, dim); }
int sdt_mob_destroy(sdt_mob_t mob) { return sdt_mob_spme_destroy(mob); }

// ------------------------------------------------------------- BrwnChol ---
int sdt_brwn_chol_create(int dim, sdt_brwn_chol_t* out) {
    if (!out) { set_error("null output handle"); return SDT_EINVAL; }
    *out = nullptr;
    return guard([&] { *out = new sdt_brwn_chol_s{new BrwnCholGpu(dim)}; });
}
int sdt_brwn_chol_init(sdt_brwn_chol_t br) {
    SDT_CHECK_HANDLE(br);
    return guard([&] { br->b->init(); });
}
int sdt_brwn_chol_compute(sdt_brwn_chol_t br, sdt_mob_t mob, int num_rhs, int ldz,
                          const double* z, int ldy, double* y) {
    SDT_CHECK_HANDLE(br);
    SDT_CHECK_HANDLE(mob);
    return guard([&] { br->b->compute(mob->m, num_rhs, ldz, z, ldy, y, false); });
}
int sdt_brwn_chol_compute_dev(sdt_brwn_chol_t br, sdt_mob_t mob, int num_rhs, int ldz,
                              const double* d_z, int ldy, double* d_y) {
    SDT_CHECK_HANDLE(br);
    SDT_CHECK_HANDLE(mob);
    return guard([&] { br->b->compute(mob->m, num_rhs, ldz, d_z, ldy, d_y, true); });
}
int sdt_brwn_chol_destroy(sdt_brwn_chol_t br) {
    if (!br) return SDT_OK;
    return guard([&] {
        delete br->b;
        delete br;
    });
}

int sdt_rpy_spme_ex(const double* pos, const double* rdi, int npos, double L, const double* f,
                    int nrhs, int ldf, const char* mode, double xi, double rmax, int dim,
                    int porder, int precision, double* v, int ldv) {
    // argument checks of mex/rpy_spme_mex.cc:12-51
    if (!pos || !rdi || !f || !v || npos <= 0) { set_error("invalid array arguments"); return SDT_EINVAL; }
    if (L <= 0. || xi <= 0. || rmax <= 0 || dim < 1) {
        set_error("Scalar arguments must be positive.");
        return SDT_EINVAL;
    }
    if (porder <= 1) { set_error("porder must be greater than 1."); return SDT_EINVAL; }
    if (2. * rmax >= L) { set_error("rmax must be less than L/2."); return SDT_EINVAL; }
    if (!mode) { set_error("invalid argument: mode."); return SDT_EINVAL; }
    int md;
    if (strcmp(mode, "full") == 0) md = 0;
    else if (strcmp(mode, "real") == 0) md = 1;
    else if (strcmp(mode, "recip") == 0) md = 2;
    else { set_error("mode must be full, real, or recip."); return SDT_EINVAL; }
    return guard([&] {
        MobSpmeGpu m(npos, rdi, L, xi, rmax, dim, porder, precision);
        m.init();
        m.update(pos, rdi, false);
        m.mul_vector(md, nrhs, 1.0, ldf, f, 0.0, ldv, v, false);
    });
}
int sdt_rpy_spme(const double* pos, const double* rdi, int npos, double L, const double* f,
                 int nrhs, int ldf, const char* mode, double xi, double rmax, int dim, int porder,
                 double* v, int ldv) {
    return sdt_rpy_spme_ex(pos, rdi, npos, L, f, nrhs, ldf, mode, xi, rmax, dim, porder, SDT_FP64,
                           v, ldv);
}

int sdt_rpy_ewald_matrix(const double* pos, const double* rdi, int npos, double L, double tol,
                         const char* mode, double xi, double* mat, int ldm) {
    // argument checks of mex/rpy_ewald_matrix_mex.cc:12-50 and MobDebug::Init
    // (src/mob/mob_debug.cc)
    if (!pos || !rdi || !mat || npos <= 0) { set_error("invalid array arguments"); return SDT_EINVAL; }
    if (L <= 0. || tol <= 0.) { set_error("Scalar arguments must be positive."); return SDT_EINVAL; }
    if (!mode) { set_error("invalid argument: mode."); return SDT_EINVAL; }
    int md;
    if (strcmp(mode, "full") == 0) md = 0;
    else if (strcmp(mode, "real") == 0) md = 1;
    else if (strcmp(mode, "recip") == 0) md = 2;
    else { set_error("mode must be full, real, or recip."); return SDT_EINVAL; }
    if (ldm < 3 * npos) { set_error("ldm must be at least 3*npos"); return SDT_EINVAL; }
    // the reference's default splitting when xi is not given
    if (xi <= 0.) xi = pow(10.0, 1.0 / 6.0) * sqrt(M_PI) / L;
    return guard([&] {
        hipStream_t s;
        SDT_HIP_CHECK(stream_create(&s));
        try {
            ewald_matrix(npos, pos, rdi, L, xi, tol, md, ldm, mat, s);
        } catch (...) {
            (void)stream_destroy(s);
            throw;
        }
        (void)stream_destroy(s);
    });
}

int sdt_rpy_ewald_rows(const double* pos, const double* rdi, int npos, double L, double tol,
                       double xi, int nrows, const int* rows, const double* f, double* v,
                       double* xi_used) {
    if (!pos || !rdi || !rows || !f || !v || npos <= 0 || nrows <= 0) {
        set_error("invalid array arguments");
        return SDT_EINVAL;
    }
    if (L <= 0. || tol <= 0.) { set_error("Scalar arguments must be positive."); return SDT_EINVAL; }
    return guard([&] {
        const double x = xi > 0. ? xi : ewald_rows_xi(npos, nrows, L, tol);
        if (xi_used) *xi_used = x;
        hipStream_t s;
        SDT_HIP_CHECK(stream_create(&s));
        try {
            ewald_rows(npos, pos, rdi, L, x, tol, nrows, rows, f, v, s);
        } catch (...) {
            (void)stream_destroy(s);
            throw;
        }
        (void)stream_destroy(s);
    });
}

int sdt_diffusion(const double* pos, int nframes, int npos, double dt, double* D, int* nlags) {
    if (!pos || !D || !nlags || npos <= 0) { set_error("invalid array arguments"); return SDT_EINVAL; }
    // calc_diff.cc:47 requires more than two frames
    if (nframes <= 2) { set_error("need more than two frames"); return SDT_EINVAL; }
    if (dt == 0.) { set_error("the frame interval must be non-zero"); return SDT_EINVAL; }
    *nlags = diffusion_lags(nframes);
    return guard([&] {
        hipStream_t s;
        SDT_HIP_CHECK(stream_create(&s));
        try {
            diffusion(pos, nframes, npos, dt, D, s);
        } catch (...) {
            (void)stream_destroy(s);
            throw;
        }
        (void)stream_destroy(s);
    });
}

int sdt_diffusion_dev_noalloc(const double* d_pos, int nframes, int npos, double dt, double* d_D,
                              int* nlags, void* stream) {
    if (!d_pos || !d_D || !nlags || npos <= 0) { set_error("invalid array arguments"); return SDT_EINVAL; }
    if (nframes <= 2) { set_error("need more than two frames"); return SDT_EINVAL; }
    if (dt == 0.) { set_error("the frame interval must be non-zero"); return SDT_EINVAL; }
    *nlags = diffusion_lags(nframes);
    return guard([&] {
        diffusion_noalloc(d_pos, nframes, npos, dt, d_D, static_cast<hipStream_t>(stream));
    });
}

// ------------------------------------------------------------- BrwnRandom --
int sdt_brwn_random_create(int dim, int num_vecs, int max_nrhs, uint64_t seed,
                           sdt_brwn_random_t* out) {
    if (!out) { set_error("null output handle"); return SDT_EINVAL; }
    *out = nullptr;
    return guard([&] { *out = new sdt_brwn_random_s{new BrwnRandomGpu(dim, num_vecs, max_nrhs, seed)}; });
}
int sdt_brwn_random_init(sdt_brwn_random_t br) {
    SDT_CHECK_HANDLE(br);
    return guard([&] { br->b->init(); });
}
int sdt_brwn_random_compute(sdt_brwn_random_t br, sdt_mob_t mob, int num_rhs, int ldz,
                            const double* z, int ldy, double* y) {
    SDT_CHECK_HANDLE(br);
    return guard([&] { br->b->compute(mob ? mob->m : nullptr, num_rhs, ldz, z, ldy, y, false); });
}
int sdt_brwn_random_compute_dev(sdt_brwn_random_t br, sdt_mob_t mob, int num_rhs, int ldz,
                                const double* d_z, int ldy, double* d_y) {
    SDT_CHECK_HANDLE(br);
    return guard([&] { br->b->compute(mob ? mob->m : nullptr, num_rhs, ldz, d_z, ldy, d_y, true); });
}
int sdt_brwn_random_destroy(sdt_brwn_random_t br) {
    if (!br) return SDT_OK;
    return guard([&] {
        delete br->b;
        delete br;
    });
}

// --------------------------------------------------------------- PairList --
int sdt_pairlist_create(int npos, const double* rdi, double box_size, double cutoff, char mode,
                        sdt_pairlist_t* out) {
    if (!out) { set_error("null output handle"); return SDT_EINVAL; }
    *out = nullptr;
    if (npos <= 0 || !rdi) { set_error("npos must be positive and rdi non-null"); return SDT_EINVAL; }
    return guard([&] {
        sdt_pairlist_s* p = new sdt_pairlist_s();
        p->rdi.assign(rdi, rdi + npos);
        p->pl.npos = npos;
        p->pl.box = box_size;
        p->pl.cutoff = cutoff;
        p->pl.mode = mode;
        hipError_t e = stream_create(&p->s);
        if (e != hipSuccess) {
            delete p;
            throw Error{SDT_EHIP, "hipStreamCreate failed"};
        }
        *out = p;
    });
}
int sdt_pairlist_init(sdt_pairlist_t pl, int64_t* est_nnz) {
    SDT_CHECK_HANDLE(pl);
    return guard([&] {
        int est = pl->pl.init(pl->pl.npos, pl->rdi.data(), pl->pl.box, pl->pl.cutoff, pl->pl.mode);
        if (est_nnz) *est_nnz = est;
        SDT_REQUIRE(est > 0, SDT_EINVAL, "pair list init failed (invalid parameters)");
        pl->d_pos.reserve(3 * (size_t)pl->pl.npos);
        pl->inited = true;
    });
}
int sdt_pairlist_build(sdt_pairlist_t pl, const double* pos, int64_t* nnz) {
    SDT_CHECK_HANDLE(pl);
    return guard([&] {
        SDT_REQUIRE(pl->inited, SDT_ESTATE, "build called before init");
        SDT_REQUIRE(pos != nullptr, SDT_EINVAL, "pos is null");
        SDT_HIP_CHECK(hipMemcpyAsync(pl->d_pos.get(), pos, sizeof(double) * 3 * pl->pl.npos,
                                     hipMemcpyHostToDevice, pl->s));
        int n = pl->pl.build(pl->d_pos.get(), nullptr, nullptr, true, pl->s);
        SDT_HIP_CHECK(hipStreamSynchronize(pl->s));
        if (nnz) *nnz = n;
        pl->built = true;
    });
}
int sdt_pairlist_get_pairs(sdt_pairlist_t pl, int32_t* rowptr, int32_t* colidx) {
    SDT_CHECK_HANDLE(pl);
    return guard([&] {
        SDT_REQUIRE(pl->built, SDT_ESTATE, "get_pairs called before build");
        SDT_REQUIRE(rowptr && (colidx || pl->pl.nnz == 0), SDT_EINVAL, "null output arrays");
        SDT_HIP_CHECK(sync_memcpy(rowptr, pl->pl.d_rowptr.get(), sizeof(int) * (pl->pl.npos + 1),
                                hipMemcpyDeviceToHost));
        if (pl->pl.nnz > 0)
            SDT_HIP_CHECK(sync_memcpy(colidx, pl->pl.d_colidx.get(), sizeof(int) * pl->pl.nnz,
                                    hipMemcpyDeviceToHost));
    });
}
int sdt_pairlist_destroy(sdt_pairlist_t pl) {
    if (!pl) return SDT_OK;
    return guard([&] {
        if (pl->s) (void)stream_destroy(pl->s);
        delete pl;
    });
}
// The list of a size query is kept (per thread) and handed out by the next
// call on the same inputs, so the two-call protocol builds the list once.
namespace {
struct PairlistQuery {
    sdt_pairlist_t h = nullptr;
    std::vector<double> pos, rdi;
    double L = 0.0, cutoff = 0.0;
    char mode = 0;
    int64_t nnz = -1;
    ~PairlistQuery() { drop(); }
    void drop() {
        if (h) sdt_pairlist_destroy(h);
        h = nullptr;
        nnz = -1;
    }
    bool same(const double* p, const double* r, int npos, double L_, double c, char m) const {
        return h && (int)rdi.size() == npos && L == L_ && cutoff == c && mode == m &&
               std::memcmp(pos.data(), p, sizeof(double) * 3 * (size_t)npos) == 0 &&
               std::memcmp(rdi.data(), r, sizeof(double) * (size_t)npos) == 0;
    }
};
thread_local PairlistQuery g_pl_query;
}  // namespace

int sdt_pairlist(const double* pos, const double* rdi, int npos, double L, double cutoff,
                 char mode, int32_t* ia, int32_t* ja, int64_t ja_cap, int64_t* nnz) {
    // argument checks of mex/pairlist_mex.cc:13-53
    if (!pos || !rdi || npos <= 0) { set_error("invalid array arguments"); return SDT_EINVAL; }
    if (L <= 0. || cutoff <= 0.) { set_error("Scalar arguments must be positive."); return SDT_EINVAL; }
    if (2. * cutoff >= L) { set_error("cutoff must be less than L/2."); return SDT_EINVAL; }
    if (mode != 'R' && mode != 'S') { set_error("mode must be R or S."); return SDT_EINVAL; }
    PairlistQuery& q = g_pl_query;
    int rc = SDT_OK;
    int64_t n = 0;
    if (q.same(pos, rdi, npos, L, cutoff, mode)) {
        n = q.nnz;   // built by the size query before
    } else {
        q.drop();
        rc = sdt_pairlist_create(npos, rdi, L, cutoff, mode, &q.h);
        if (rc) return rc;
        int64_t est = 0;
        rc = sdt_pairlist_init(q.h, &est);
        if (!rc) rc = sdt_pairlist_build(q.h, pos, &n);
        if (rc) {
            q.drop();
            return rc;
        }
    }
    if (nnz) *nnz = n;
    if (n > ja_cap || !ia || (n > 0 && !ja)) {
        // a size query (null outputs) or too small a ja: report nnz only and
        // keep the list for the call with buffers
        q.pos.assign(pos, pos + 3 * (size_t)npos);
        q.rdi.assign(rdi, rdi + npos);
        q.L = L;
        q.cutoff = cutoff;
        q.mode = mode;
        q.nnz = n;
        set_error("ja capacity is smaller than nnz");
        return SDT_ECAP;
    }
    rc = sdt_pairlist_get_pairs(q.h, ia, ja);
    q.drop();
    return rc;
}

// ------------------------------------------------------------ BrwnLanczos --
int sdt_brwn_lanczos_create(int dim, int max_iters, int max_nrhs, double tol, sdt_brwn_t* out) {
    if (!out) { set_error("null output handle"); return SDT_EINVAL; }
    *out = nullptr;
    return guard([&] { *out = new sdt_brwn_s{new BrwnLanczosGpu(dim, max_iters, max_nrhs, tol)}; });
}
int sdt_brwn_lanczos_init(sdt_brwn_t br) {
    SDT_CHECK_HANDLE(br);
    return guard([&] { br->b->init(); });
}
static int lanczos_status(BrwnLanczosGpu* b) {
    for (int it : b->iters())
        if (it < 0) {
            set_error("Lanczos did not converge for at least one group (result written)");
            return SDT_ECONVERGE;
        }
    return SDT_OK;
}
int sdt_brwn_lanczos_compute(sdt_brwn_t br, sdt_mob_t mob, int num_rhs, int ldz, const double* z,
                             int ldy, double* y) {
    SDT_CHECK_HANDLE(br);
    SDT_CHECK_HANDLE(mob);
    int rc = guard([&] { br->b->compute(mob->m, num_rhs, ldz, z, ldy, y, false); });
    return rc ? rc : lanczos_status(br->b);
}
int sdt_brwn_lanczos_compute_dev(sdt_brwn_t br, sdt_mob_t mob, int num_rhs, int ldz,
                                 const double* d_z, int ldy, double* d_y) {
    SDT_CHECK_HANDLE(br);
    SDT_CHECK_HANDLE(mob);
    int rc = guard([&] { br->b->compute(mob->m, num_rhs, ldz, d_z, ldy, d_y, true); });
    return rc ? rc : lanczos_status(br->b);
}
int sdt_brwn_lanczos_iters(sdt_brwn_t br, int* iters, int cap, int* ngroups) {
    SDT_CHECK_HANDLE(br);
    const std::vector<int>& it = br->b->iters();
    if (ngroups) *ngroups = (int)it.size();
    for (int i = 0; i < (int)it.size() && i < cap && iters; i++) iters[i] = it[i];
    return SDT_OK;
}
int sdt_brwn_lanczos_set_reorth(sdt_brwn_t br, int reorth) {
    SDT_CHECK_HANDLE(br);
    br->b->reorth = reorth ? 1 : 0;
    return SDT_OK;
}
int sdt_brwn_lanczos_destroy(sdt_brwn_t br) {
    if (!br) return SDT_OK;
    return guard([&] {
        delete br->b;
        delete br;
    });
}

// ---------------------------------------------------------------- Steric ---
int sdt_steric_create(int npos, const double* rdi, double box_size, double r0, double k0,
                      sdt_force_t* out) {
    if (!out) { set_error("null output handle"); return SDT_EINVAL; }
    *out = nullptr;
    return guard([&] {
        StericGpu* f = new StericGpu(npos, rdi, box_size, r0, k0);
        hipStream_t s;
        hipError_t e = stream_create(&s);
        if (e != hipSuccess) {
            delete f;
            throw Error{SDT_EHIP, "hipStreamCreate failed"};
        }
        *out = new sdt_force_s{f, s};
    });
}
int sdt_steric_init(sdt_force_t fo) {
    SDT_CHECK_HANDLE(fo);
    return guard([&] { fo->f->init(); });
}
int sdt_steric_accumulate(sdt_force_t fo, const double* pos, const double* rdi, double* f) {
    (void)rdi;
    SDT_CHECK_HANDLE(fo);
    return guard([&] {
        SDT_REQUIRE(pos && f, SDT_EINVAL, "null array");
        fo->f->accumulate(pos, f, false, fo->s);
    });
}
int sdt_steric_compute(sdt_force_t fo, const double* pos, const double* rdi, double* f) {
    (void)rdi;
    SDT_CHECK_HANDLE(fo);
    // ForceBase::Compute zeroes f before accumulating
    // (src/force/force_base.cc:24-28)
    return guard([&] {
        SDT_REQUIRE(pos && f, SDT_EINVAL, "null array");
        memset(f, 0, sizeof(double) * 3 * (size_t)fo->f->npos());
        fo->f->accumulate(pos, f, false, fo->s);
    });
}
int sdt_force_steric(const double* pos, const double* rdi, int npos, double L, double r0,
                     double k0, double* f) {
    // argument checks of mex/force_steric_mex.cc:13-37
    if (!pos || !rdi || !f || npos <= 0) { set_error("invalid array arguments"); return SDT_EINVAL; }
    if (L <= 0. || r0 <= 0. || k0 <= 0.) {
        set_error("Scalar arguments must be positive.");
        return SDT_EINVAL;
    }
    sdt_force_t h = nullptr;
    int rc = sdt_steric_create(npos, rdi, L, r0, k0, &h);
    if (!rc) rc = sdt_steric_init(h);
    if (!rc) rc = sdt_steric_compute(h, pos, rdi, f);
    sdt_steric_destroy(h);
    return rc;
}
int sdt_steric_destroy(sdt_force_t fo) {
    if (!fo) return SDT_OK;
    return guard([&] {
        (void)stream_destroy(fo->s);
        delete fo->f;
        delete fo;
    });
}

// -------------------------------------------------------------------- BD ---
int sdt_bd_create(int npos, const double* pos, const double* rdi, double box_size, double xi,
                  double rmax, int dim, int porder, int precision, int max_iters, int max_nrhs,
                  double tol, double steric_r0, double steric_k0, double dt, int mob_interval,
                  uint64_t seed, sdt_bd_t* out) {
    if (!out) { set_error("null output handle"); return SDT_EINVAL; }
    *out = nullptr;
    return guard([&] {
        BDGpu* bd = new BDGpu(npos, pos, rdi, box_size, xi, rmax, dim, porder, precision, max_iters,
                              max_nrhs, tol, steric_r0, steric_k0, dt, mob_interval, seed);
        *out = new sdt_bd_s{bd};
    });
}
int sdt_bd_init(sdt_bd_t bd) {
    SDT_CHECK_HANDLE(bd);
    return guard([&] { bd->bd->init(); });
}
int sdt_bd_advance(sdt_bd_t bd, int nsteps) {
    SDT_CHECK_HANDLE(bd);
    return guard([&] { bd->bd->advance(nsteps); });
}
int sdt_bd_get_positions(sdt_bd_t bd, double* pos) {
    SDT_CHECK_HANDLE(bd);
    return guard([&] { bd->bd->get_positions(pos); });
}
int sdt_bd_last_iters(sdt_bd_t bd, int* iters, int cap, int* ngroups) {
    SDT_CHECK_HANDLE(bd);
    const std::vector<int>& it = bd->bd->last_iters();
    if (ngroups) *ngroups = (int)it.size();
    for (int i = 0; i < (int)it.size() && i < cap && iters; i++) iters[i] = it[i];
    return SDT_OK;
}
int sdt_bd_destroy(sdt_bd_t bd) {
    if (!bd) return SDT_OK;
    return guard([&] {
        delete bd->bd;
        delete bd;
    });
}

// ------------------------------------------------------------------ misc ---
int sdt_gaussian(uint64_t seed, uint32_t stream, int64_t n, double* out) {
    if (!out || n < 0) { set_error("invalid arguments"); return SDT_EINVAL; }
    return guard([&] {
        DevBuf<double> d;
        d.reserve((size_t)n);
        gaussian_fill_device(seed, stream, n, d.get(), 0);
        SDT_HIP_CHECK(sync_memcpy(out, d.get(), sizeof(double) * n, hipMemcpyDeviceToHost));
    });
}
int sdt_device_synchronize(void) {
    return guard([&] { SDT_HIP_CHECK(hipDeviceSynchronize()); });
}

// ---------------------------------------------------------------- SimBox --
int sdt_simbox_create(const char* config_file, const char* model_file, const char* xyz_file,
                      sdt_simbox_t* out) {
    if (!out) { set_error("null output handle"); return SDT_EINVAL; }
    *out = nullptr;
    return guard([&] { *out = new sdt_simbox_s{new SimBoxGpu(config_file, model_file, xyz_file)}; });
}
int sdt_simbox_init(sdt_simbox_t sb) {
    SDT_CHECK_HANDLE(sb);
    return guard([&] { sb->sb->init(); });
}
int sdt_simbox_advance(sdt_simbox_t sb, int nsteps) {
    SDT_CHECK_HANDLE(sb);
    return guard([&] { sb->sb->advance(nsteps); });
}
int sdt_simbox_npos(sdt_simbox_t sb, int* npos) {
    SDT_CHECK_HANDLE(sb);
    if (!npos) { set_error("null output"); return SDT_EINVAL; }
    *npos = sb->sb->npos();
    return SDT_OK;
}
int sdt_simbox_state(sdt_simbox_t sb, int64_t* step, double* time) {
    SDT_CHECK_HANDLE(sb);
    if (step) *step = sb->sb->step();
    if (time) *time = sb->sb->time();
    return SDT_OK;
}
int sdt_simbox_params(sdt_simbox_t sb, double* dparams, int64_t* iparams) {
    SDT_CHECK_HANDLE(sb);
    const SimBoxParams& p = sb->sb->params();
    if (dparams) {
        const double d[7] = {p.spme_xi, p.spme_rmax, p.lanczos_tol, p.steric_r0, p.steric_k0,
                             p.delta_t, p.ewald_tol};
        memcpy(dparams, d, sizeof(d));
    }
    if (iparams) {
        const int64_t v[11] = {p.spme_dim, p.spme_porder, p.mob_interval, p.lanczos_maxiters,
                               p.lanczos_maxnrhs, p.traj_interval, p.traj_start, p.precision,
                               (int64_t)p.seed, p.ewald ? 1 : 0,
                               p.chol ? 2 : (p.random ? 1 : 0)};
        memcpy(iparams, v, sizeof(v));
    }
    return SDT_OK;
}
int sdt_simbox_get_positions(sdt_simbox_t sb, double* pos) {
    SDT_CHECK_HANDLE(sb);
    if (!pos) { set_error("null output"); return SDT_EINVAL; }
    return guard([&] { sb->sb->positions(pos); });
}
int sdt_simbox_get_radii(sdt_simbox_t sb, double* rdi) {
    SDT_CHECK_HANDLE(sb);
    if (!rdi) { set_error("null output"); return SDT_EINVAL; }
    return guard([&] { sb->sb->radii(rdi); });
}
int sdt_simbox_write_xyz(sdt_simbox_t sb, const char* path) {
    SDT_CHECK_HANDLE(sb);
    if (!path) { set_error("null path"); return SDT_EINVAL; }
    return guard([&] { sb->sb->write_xyz(path); });
}
int sdt_simbox_destroy(sdt_simbox_t sb) {
    if (!sb) return SDT_OK;
    delete sb->sb;
    delete sb;
    return SDT_OK;
}

// ----------------------------------------------------------- debug hooks ---
int sdt_debug_gram_groups(int64_t n, int btot, int bs, const double* A, const double* B,
                          double* G) {
    return guard([&] {
        SDT_REQUIRE(n > 0 && btot >= 1 && bs >= 1 && bs <= 16 && A && B && G, SDT_EINVAL,
                    "bad arguments");
        const int ng = (btot + bs - 1) / bs;
        DevBuf<double> dA, dB, dG, part;
        dA.reserve((size_t)n * btot);
        dB.reserve((size_t)n * btot);
        dG.reserve((size_t)ng * 256);
        SDT_HIP_CHECK(sync_memcpy(dA.get(), A, sizeof(double) * n * btot, hipMemcpyHostToDevice));
        SDT_HIP_CHECK(sync_memcpy(dB.get(), B, sizeof(double) * n * btot, hipMemcpyHostToDevice));
        gram_groups(n, btot, bs, dA.get(), dB.get(), dG.get(), part, 0);
        std::vector<double> h((size_t)ng * 256);
        SDT_HIP_CHECK(sync_memcpy(h.data(), dG.get(), sizeof(double) * h.size(),
                                hipMemcpyDeviceToHost));
        for (int g = 0; g < ng; g++)
            for (int i = 0; i < bs; i++)
                for (int j = 0; j < bs; j++)
                    G[((size_t)g * bs + i) * bs + j] = h[(size_t)g * 256 + i * 16 + j];
    });
}
int sdt_debug_gram(int64_t n, int b, const double* A, const double* B, double* G) {
    return sdt_debug_gram_groups(n, b, b, A, B, G);
}
int sdt_debug_band_sqrt(int m, int kd, const double* T, int kb, double* S) {
    return guard([&] {
        SDT_REQUIRE(m > 0 && kd >= 0 && kb > 0 && kb <= m && T && S, SDT_EINVAL, "bad arguments");
        BandSqrt bs;
        bs.compute(m, kd, T, kb, S);
    });
}
int sdt_debug_spme_spread(sdt_mob_t mob, const double* f, double* mesh) {
    SDT_CHECK_HANDLE(mob);
    return guard([&] {
        MobSpmeGpu* m = spme_of(mob);
        SDT_REQUIRE(m->updated(), SDT_ESTATE, "mobility not updated");
        const int n3 = m->dim();
        const size_t K3 = (size_t)m->K() * m->K() * m->K();
        DevBuf<double> df, dX;
        df.reserve(n3);
        dX.reserve(n3);
        hipStream_t s = m->stream();
        SDT_HIP_CHECK(hipMemcpyAsync(df.get(), f, sizeof(double) * n3, hipMemcpyHostToDevice, s));
        pack(m->npos(), 1, m->perm(), n3, df.get(), dX.get(), s);
        std::vector<float> tmp;
        if (m->precision() == kFP32) {
            SpmeGpu<float>* e = m->spme32();
            e->reserve_rhs(1);
            e->spread(1, dX.get(), s);
            tmp.resize(3 * K3);
            SDT_HIP_CHECK(hipMemcpyAsync(tmp.data(), e->d_real.get(), sizeof(float) * 3 * K3,
                                         hipMemcpyDeviceToHost, s));
            SDT_HIP_CHECK(hipStreamSynchronize(s));
            for (size_t i = 0; i < 3 * K3; i++) mesh[i] = tmp[i];
        } else {
            SpmeGpu<double>* e = m->spme64();
            e->reserve_rhs(1);
            e->spread(1, dX.get(), s);
            SDT_HIP_CHECK(hipMemcpyAsync(mesh, e->d_real.get(), sizeof(double) * 3 * K3,
                                         hipMemcpyDeviceToHost, s));
            SDT_HIP_CHECK(hipStreamSynchronize(s));
        }
    });
}

int sdt_debug_spme_forms(sdt_mob_t mob, int interp_form, int z_form, int spread_form,
                         int fft_form) {
    SDT_CHECK_HANDLE(mob);
    return guard([&] {
        SDT_REQUIRE(fft_form == 0 || fft_form == 1, SDT_EINVAL, "fft form must be 0 or 1");
        SDT_REQUIRE(spread_form >= 0 && spread_form <= 15, SDT_EINVAL,
                    "spread form: 0 .. 3 (built-in, LDS, scalar, matrix cores), plus 4 for "
                    "row-tiled mesh stores, plus 8 for column blocks in turn (with 3)");
        SDT_REQUIRE(interp_form == 0 || interp_form == 1 || interp_form == 2 || interp_form == 4,
                    SDT_EINVAL, "interp meshes per pass must be 0 (built-in), 1, 2 or 4");
        SDT_REQUIRE(z_form >= 0 && z_form <= 5, SDT_EINVAL, "z transform form must be in [0, 5]");
        MobSpmeGpu* m = spme_of(mob);
        if (m->spme32()) {
            m->spme32()->interp_form = interp_form;
            m->spme32()->z_form = z_form;
            m->spme32()->spread_form = spread_form;
            m->spme32()->fft_form = fft_form;
        }
        if (m->spme64()) {
            m->spme64()->interp_form = interp_form;
            m->spme64()->z_form = z_form;
            m->spme64()->spread_form = spread_form;
            m->spme64()->fft_form = fft_form;
        }
    });
}

int sdt_debug_spme_tuning(sdt_mob_t mob, int fft_order, int interp_sxt, int interp_ch,
                          int spread_stage_kb) {
    SDT_CHECK_HANDLE(mob);
    return guard([&] {
        SpmeTuning t;
        t.fft_order = fft_order;
        t.interp_sxt = interp_sxt;
        t.interp_ch = interp_ch;
        t.spread_stage_kb = spread_stage_kb;
        // single-column and Lanczos-batch launches alike
        spme_of(mob)->set_tuning(t, 1);
        spme_of(mob)->set_tuning(t, 100);
    });
}

int sdt_debug_mob_launch_width(sdt_mob_t mob, int width) {
    SDT_CHECK_HANDLE(mob);
    return guard([&] {
        SDT_REQUIRE(width >= 1 && width <= 1024, SDT_EINVAL, "launch width must be in [1, 1024]");
        spme_of(mob)->set_launch_width(width);
    });
}

int sdt_debug_spme_interp(sdt_mob_t mob, const double* mesh, double alpha, double* v) {
    SDT_CHECK_HANDLE(mob);
    return guard([&] {
        MobSpmeGpu* m = spme_of(mob);
        SDT_REQUIRE(m->updated(), SDT_ESTATE, "mobility not updated");
        const int n3 = m->dim();
        const size_t K3 = (size_t)m->K() * m->K() * m->K();
        DevBuf<double> dY, dv;
        dY.reserve(n3);
        dv.reserve(n3);
        hipStream_t s = m->stream();
        // interp applies alpha / L^3 (the mesh normalisation of the matvec)
        const double L = m->box();
        const double a = alpha * L * L * L;
        if (m->precision() == kFP32) {
            SpmeGpu<float>* e = m->spme32();
            SDT_REQUIRE(!e->poly, SDT_EINVAL, "interp probe: monodisperse only");
            e->reserve_rhs(1);
            std::vector<float> tmp(3 * K3);
            for (size_t i = 0; i < 3 * K3; i++) tmp[i] = (float)mesh[i];
            SDT_HIP_CHECK(hipMemcpyAsync(e->d_real.get(), tmp.data(), sizeof(float) * 3 * K3,
                                         hipMemcpyHostToDevice, s));
            e->interp(1, a, 0.0, dY.get(), s);
            SDT_HIP_CHECK(hipStreamSynchronize(s));
        } else {
            SpmeGpu<double>* e = m->spme64();
            SDT_REQUIRE(!e->poly, SDT_EINVAL, "interp probe: monodisperse only");
            e->reserve_rhs(1);
            SDT_HIP_CHECK(hipMemcpyAsync(e->d_real.get(), mesh, sizeof(double) * 3 * K3,
                                         hipMemcpyHostToDevice, s));
            e->interp(1, a, 0.0, dY.get(), s);
        }
        unpack(m->npos(), 1, m->perm(), n3, dY.get(), dv.get(), s);
        SDT_HIP_CHECK(hipMemcpyAsync(v, dv.get(), sizeof(double) * n3, hipMemcpyDeviceToHost, s));
        SDT_HIP_CHECK(hipStreamSynchronize(s));
    });
}

}  // extern "C"

template <typename T>
static void stage_times(MobSpmeGpu* m, SpmeGpu<T>* e, int b, int reps, const double* X, double* Y,
                        double* t) {
    hipStream_t s = m->stream();
    e->reserve_rhs(b);
    // the spread stage as the full matvec runs it: with the column-blocked
    // input copy (timed with the spread) when the real space's blocks match
    const int rc = real_block_cols<T>(b);
    const bool share = rc > 0 && e->blocked_cols(b) == rc;
    DevBuf<double> xbuf;
    if (share) xbuf.reserve(ceil_div(real_scratch_bytes<T>(m->npos(), b), sizeof(double)));
    hipEvent_t ev[8];
    for (int i = 0; i < 8; i++) SDT_HIP_CHECK(hipEventCreate(&ev[i]));
    for (int i = 0; i < 7; i++) t[i] = 0.0;
    for (int r = 0; r < reps + 1; r++) {
        SDT_HIP_CHECK(hipEventRecord(ev[0], s));
        if (share) {
            real_block_copy<T>(m->npos(), b, X, xbuf.get(), s);
            e->spread(b, X, s, reinterpret_cast<const T*>(xbuf.get()));
        } else {
            e->spread(b, X, s);
        }
        SDT_HIP_CHECK(hipEventRecord(ev[1], s));
        e->fft_forward(b, s);
        SDT_HIP_CHECK(hipEventRecord(ev[2], s));
        e->influence(b, s);
        SDT_HIP_CHECK(hipEventRecord(ev[3], s));
        e->fft_backward(b, s);
        SDT_HIP_CHECK(hipEventRecord(ev[4], s));
        e->interp(b, 1.0, 0.0, Y, s);
        SDT_HIP_CHECK(hipEventRecord(ev[5], s));
        m->apply_packed(1, b, 1.0, X, 1.0, Y);
        SDT_HIP_CHECK(hipEventRecord(ev[6], s));
        m->apply_packed(0, b, 1.0, X, 0.0, Y);
        SDT_HIP_CHECK(hipEventRecord(ev[7], s));
        SDT_HIP_CHECK(hipEventSynchronize(ev[7]));
        if (r == 0) continue;   // warm-up
        for (int i = 0; i < 7; i++) {
            float ms = 0.f;
            SDT_HIP_CHECK(hipEventElapsedTime(&ms, ev[i], ev[i + 1]));
            t[i] += ms / reps;
        }
    }
    for (int i = 0; i < 8; i++) (void)hipEventDestroy(ev[i]);
}

extern "C" {

int sdt_bench_stage_times(sdt_mob_t mob, int b, int reps, double* t) {
    SDT_CHECK_HANDLE(mob);
    return guard([&] {
        MobSpmeGpu* m = spme_of(mob);
        SDT_REQUIRE(m->updated(), SDT_ESTATE, "mobility not updated");
        SDT_REQUIRE(b >= 1 && b <= 256 && reps >= 1 && t, SDT_EINVAL, "bad arguments");
        const size_t n = (size_t)m->dim() * b;
        DevBuf<double> dX, dY;
        dX.reserve(n);
        dY.reserve(n);
        gaussian_fill_device(99, 7, (long long)n, dX.get(), m->stream());
        fill((long long)n, 0.0, dY.get(), m->stream());
        if (m->precision() == kFP32) stage_times(m, m->spme32(), b, reps, dX.get(), dY.get(), t);
        else stage_times(m, m->spme64(), b, reps, dX.get(), dY.get(), t);
    });
}

}  // extern "C"

// Plain streaming copy of doubles: the reference kernel for scaling HBM
// byte counters (it reads and writes exactly 8 n bytes each).
__global__ void sdt_copy_kernel(long long n, const double* __restrict__ src,
                                double* __restrict__ dst) {
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        dst[i] = src[i];
}

extern "C" {

int sdt_bench_copy(int64_t n, int reps, double* ms_avg) {
    return guard([&] {
        SDT_REQUIRE(n > 0 && reps >= 1, SDT_EINVAL, "bad arguments");
        DevBuf<double> a, b;
        a.reserve((size_t)n);
        b.reserve((size_t)n);
        fill(n, 1.0, a.get(), 0);
        hipEvent_t e0, e1;
        SDT_HIP_CHECK(hipEventCreate(&e0));
        SDT_HIP_CHECK(hipEventCreate(&e1));
        SDT_HIP_CHECK(hipEventRecord(e0, 0));
        for (int r = 0; r < reps; r++)
            hipLaunchKernelGGL(sdt_copy_kernel, dim3(4096), dim3(256), 0, 0, (long long)n, a.get(),
                               b.get());
        SDT_HIP_CHECK(hipGetLastError());
        SDT_HIP_CHECK(hipEventRecord(e1, 0));
        SDT_HIP_CHECK(hipEventSynchronize(e1));
        float ms = 0.f;
        SDT_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
        if (ms_avg) *ms_avg = ms / reps;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    });
}

int sdt_bench_bd_steps(sdt_bd_t bd, int nsteps, double* ms_total) {
    SDT_CHECK_HANDLE(bd);
    return guard([&] {
        hipStream_t s = bd->bd->stream();
        hipEvent_t a, b;
        SDT_HIP_CHECK(hipEventCreate(&a));
        SDT_HIP_CHECK(hipEventCreate(&b));
        SDT_HIP_CHECK(hipEventRecord(a, s));
        bd->bd->advance(nsteps);
        SDT_HIP_CHECK(hipEventRecord(b, s));
        SDT_HIP_CHECK(hipEventSynchronize(b));
        float ms = 0.f;
        SDT_HIP_CHECK(hipEventElapsedTime(&ms, a, b));
        if (ms_total) *ms_total = ms;
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
    });
}

int sdt_bd_set_profiling(sdt_bd_t bd, int on) {
    SDT_CHECK_HANDLE(bd);
    StageTimer& t = bd->bd->timer();
    t.on = on != 0;
    if (t.on) t.reset();
    return SDT_OK;
}

int sdt_bd_profile(sdt_bd_t bd, double* sum_ms, int64_t* counts, int nstages) {
    SDT_CHECK_HANDLE(bd);
    StageTimer& t = bd->bd->timer();
    for (int i = 0; i < nstages && i < StageTimer::kStages; i++) {
        if (sum_ms) sum_ms[i] = t.sum_ms[i];
        if (counts) counts[i] = t.count[i];
    }
    return SDT_OK;
}

int sdt_bd_profile_columns(sdt_bd_t bd, int64_t* columns, int nstages) {
    SDT_CHECK_HANDLE(bd);
    StageTimer& t = bd->bd->timer();
    for (int i = 0; i < nstages && i < StageTimer::kStages; i++)
        if (columns) columns[i] = t.cols[i];
    return SDT_OK;
}

int sdt_util_rsa_positions(int n, double box, double a, uint64_t seed, int64_t max_attempts,
                           double* pos, int64_t* placed) {
    if (n <= 0 || box <= 0.0 || a <= 0.0 || !pos) { set_error("invalid arguments"); return SDT_EINVAL; }
    long long k = rsa_positions(n, box, a, seed, max_attempts, pos);
    if (placed) *placed = k;
    if (k < n) { set_error("random sequential addition ran out of attempts"); return SDT_ECAP; }
    return SDT_OK;
}

// ------------------------------------------------------ plain FP64 path --
int sdt_plain_mob_create(int npos, const double* rdi, double box_size, double xi, double rmax,
                         int dim, int porder, sdt_plain_mob_t* out) {
    if (!out) { set_error("null output handle"); return SDT_EINVAL; }
    *out = nullptr;
    return guard([&] {
        PlainMobGpu* m = new PlainMobGpu(npos, rdi, box_size, xi, rmax, dim, porder);
        try {
            m->init();
        } catch (...) {
            delete m;
            throw;
        }
        *out = new sdt_plain_mob_s{m, {}, {}, {}};
    });
}
int sdt_plain_mob_update(sdt_plain_mob_t mob, const double* pos) {
    SDT_CHECK_HANDLE(mob);
    return guard([&] {
        SDT_REQUIRE(pos != nullptr, SDT_EINVAL, "pos is null");
        const size_t n = (size_t)mob->m->dim();
        mob->d_pos.reserve(n);
        SDT_HIP_CHECK(sync_memcpy(mob->d_pos.get(), pos, sizeof(double) * n, hipMemcpyHostToDevice));
        mob->m->update(mob->d_pos.get());
    });
}
int sdt_plain_mob_mulvector(sdt_plain_mob_t mob, int mode, int num_rhs, double alpha, int ldf,
                            const double* f, double beta, int ldv, double* v) {
    SDT_CHECK_HANDLE(mob);
    return guard([&] {
        SDT_REQUIRE(f && v && num_rhs >= 1, SDT_EINVAL, "null vector or no right-hand side");
        const int n = mob->m->dim();
        SDT_REQUIRE(ldf >= n && ldv >= n, SDT_EINVAL, "leading dimensions must be at least 3*npos");
        mob->d_f.reserve((size_t)n * num_rhs);
        mob->d_v.reserve((size_t)n * num_rhs);
        SDT_HIP_CHECK(hipMemcpy2D(mob->d_f.get(), sizeof(double) * n, f, sizeof(double) * ldf,
                                  sizeof(double) * n, num_rhs, hipMemcpyHostToDevice));
        SDT_HIP_CHECK(hipMemcpy2D(mob->d_v.get(), sizeof(double) * n, v, sizeof(double) * ldv,
                                  sizeof(double) * n, num_rhs, hipMemcpyHostToDevice));
        mob->m->mul_vector(mode, num_rhs, alpha, n, mob->d_f.get(), beta, n, mob->d_v.get());
        SDT_HIP_CHECK(hipStreamSynchronize(mob->m->stream()));
        SDT_HIP_CHECK(hipMemcpy2D(v, sizeof(double) * ldv, mob->d_v.get(), sizeof(double) * n,
                                  sizeof(double) * n, num_rhs, hipMemcpyDeviceToHost));
    });
}
int sdt_plain_lanczos(sdt_plain_mob_t mob, int max_iters, int max_nrhs, double tol, int num_rhs,
                      int ldz, const double* z, int ldy, double* y, int* iters, int cap,
                      int* ngroups) {
    SDT_CHECK_HANDLE(mob);
    return guard([&] {
        SDT_REQUIRE(z && y, SDT_EINVAL, "null vector");
        const int n = mob->m->dim();
        SDT_REQUIRE(num_rhs >= 1 && ldz >= n && ldy >= n, SDT_EINVAL, "bad shapes");
        PlainLanczosGpu lz(n, max_iters, max_nrhs, tol);
        lz.init();
        DevBuf<double> dz, dy;
        dz.reserve((size_t)n * num_rhs);
        dy.reserve((size_t)n * num_rhs);
        SDT_HIP_CHECK(hipMemcpy2D(dz.get(), sizeof(double) * n, z, sizeof(double) * ldz,
                                  sizeof(double) * n, num_rhs, hipMemcpyHostToDevice));
        lz.compute(mob->m, num_rhs, n, dz.get(), n, dy.get());
        SDT_HIP_CHECK(hipMemcpy2D(y, sizeof(double) * ldy, dy.get(), sizeof(double) * n,
                                  sizeof(double) * n, num_rhs, hipMemcpyDeviceToHost));
        const std::vector<int>& it = lz.iters();
        if (ngroups) *ngroups = (int)it.size();
        for (int i = 0; i < (int)it.size() && i < cap && iters; i++) iters[i] = it[i];
    });
}
int sdt_plain_mob_destroy(sdt_plain_mob_t mob) {
    if (!mob) return SDT_OK;
    return guard([&] {
        delete mob->m;
        delete mob;
    });
}

int sdt_plain_bd_create(int npos, const double* pos, const double* rdi, double box_size, double xi,
                        double rmax, int dim, int porder, int max_iters, int max_nrhs, double tol,
                        double steric_r0, double steric_k0, double dt, int mob_interval,
                        uint64_t seed, sdt_plain_bd_t* out) {
    if (!out) { set_error("null output handle"); return SDT_EINVAL; }
    *out = nullptr;
    return guard([&] {
        PlainBDGpu* bd = new PlainBDGpu(npos, pos, rdi, box_size, xi, rmax, dim, porder, max_iters,
                                        max_nrhs, tol, steric_r0, steric_k0, dt, mob_interval, seed);
        *out = new sdt_plain_bd_s{bd};
    });
}
int sdt_plain_bd_init(sdt_plain_bd_t bd) {
    SDT_CHECK_HANDLE(bd);
    return guard([&] { bd->bd->init(); });
}
int sdt_plain_bd_advance(sdt_plain_bd_t bd, int nsteps) {
    SDT_CHECK_HANDLE(bd);
    return guard([&] { bd->bd->advance(nsteps); });
}
int sdt_plain_bd_get_positions(sdt_plain_bd_t bd, double* pos) {
    SDT_CHECK_HANDLE(bd);
    return guard([&] { bd->bd->get_positions(pos); });
}
int sdt_plain_bd_last_iters(sdt_plain_bd_t bd, int* iters, int cap, int* ngroups) {
    SDT_CHECK_HANDLE(bd);
    const std::vector<int>& it = bd->bd->last_iters();
    if (ngroups) *ngroups = (int)it.size();
    for (int i = 0; i < (int)it.size() && i < cap && iters; i++) iters[i] = it[i];
    return SDT_OK;
}
int sdt_plain_bd_set_profiling(sdt_plain_bd_t bd, int on) {
    SDT_CHECK_HANDLE(bd);
    StageTimer& t = bd->bd->timer();
    t.on = on != 0;
    if (t.on) t.reset();
    return SDT_OK;
}
int sdt_plain_bd_profile(sdt_plain_bd_t bd, double* sum_ms, int64_t* counts, int64_t* columns,
                         int nstages) {
    SDT_CHECK_HANDLE(bd);
    StageTimer& t = bd->bd->timer();
    for (int i = 0; i < nstages && i < StageTimer::kStages; i++) {
        if (sum_ms) sum_ms[i] = t.sum_ms[i];
        if (counts) counts[i] = t.count[i];
        if (columns) columns[i] = t.cols[i];
    }
    return SDT_OK;
}
int sdt_plain_bd_destroy(sdt_plain_bd_t bd) {
    if (!bd) return SDT_OK;
    return guard([&] {
        delete bd->bd;
        delete bd;
    });
}

// ------------------------------------------------------------- Ensemble ---
int sdt_ensemble_create(int nsys, const int* npos, const double* pos, const double* rdi,
                        const double* box_size, const double* xi, const double* rmax,
                        const int* dim, int porder, int precision, int max_iters, int max_nrhs,
                        double tol, double steric_r0, double steric_k0, double dt,
                        int mob_interval, uint64_t seed, sdt_ensemble_t* out) {
    if (!out) { set_error("null output handle"); return SDT_EINVAL; }
    *out = nullptr;
    if (nsys <= 0 || !npos || !pos || !rdi || !box_size || !xi || !rmax || !dim) {
        set_error("invalid array arguments");
        return SDT_EINVAL;
    }
    return guard([&] {
        std::vector<EnsembleSystem> sys(nsys);
        for (int s = 0; s < nsys; s++) {
            sys[s].npos = npos[s];
            sys[s].box = box_size[s];
            sys[s].xi = xi[s];
            sys[s].rmax = rmax[s];
            sys[s].dim = dim[s];
        }
        EnsembleGpu* e = new EnsembleGpu(sys, pos, rdi, porder, precision, max_iters, max_nrhs,
                                         tol, steric_r0, steric_k0, dt, mob_interval, seed);
        *out = new sdt_ensemble_s{e};
    });
}
int sdt_ensemble_init(sdt_ensemble_t e, int autotune, const char* tune_cache) {
    SDT_CHECK_HANDLE(e);
    return guard([&] { e->e->init(autotune != 0, tune_cache ? tune_cache : ""); });
}
int sdt_ensemble_set_graphs(sdt_ensemble_t e, int on) {
    SDT_CHECK_HANDLE(e);
    e->e->set_graphs(on != 0);
    return SDT_OK;
}
int sdt_ensemble_advance(sdt_ensemble_t e, int nsteps, double* step_ms) {
    SDT_CHECK_HANDLE(e);
    return guard([&] { e->e->advance(nsteps, step_ms); });
}
int sdt_ensemble_get_positions(sdt_ensemble_t e, int sys, double* pos) {
    SDT_CHECK_HANDLE(e);
    return guard([&] { e->e->get_positions(sys, pos); });
}
int sdt_ensemble_tuning(sdt_ensemble_t e, int sys, int* choices, double* ms) {
    SDT_CHECK_HANDLE(e);
    return guard([&] {
        SDT_REQUIRE(sys >= 0 && sys < e->e->size(), SDT_EINVAL, "system index out of range");
        const std::vector<TuneResult>& t = e->e->tuning();
        for (int w = 0; w < 2; w++) {
            const TuneResult& r = t[2 * sys + w];
            if (choices) {
                choices[4 * w] = r.tuning.fft_order;
                choices[4 * w + 1] = r.tuning.interp_sxt;
                choices[4 * w + 2] = r.tuning.interp_ch;
                choices[4 * w + 3] = r.tuning.spread_stage_kb;
            }
            if (ms) {
                ms[2 * w] = r.ms_default;
                ms[2 * w + 1] = r.ms_best;
            }
        }
    });
}
int sdt_ensemble_last_iters(sdt_ensemble_t e, int sys, int* iters, int cap, int* ngroups) {
    SDT_CHECK_HANDLE(e);
    return guard([&] {
        SDT_REQUIRE(sys >= 0 && sys < e->e->size(), SDT_EINVAL, "system index out of range");
        const std::vector<int>& it = e->e->system(sys)->last_iters();
        if (ngroups) *ngroups = (int)it.size();
        for (int i = 0; i < (int)it.size() && i < cap && iters; i++) iters[i] = it[i];
    });
}
int sdt_ensemble_destroy(sdt_ensemble_t e) {
    if (!e) return SDT_OK;
    return guard([&] {
        delete e->e;
        delete e;
    });
}
int sdt_bd_autotune(sdt_bd_t bd, const char* tune_cache, int* choices, double* ms) {
    SDT_CHECK_HANDLE(bd);
    return guard([&] {
        BDGpu* b = bd->bd;
        MobSpmeGpu* m = b->mob();
        SDT_REQUIRE(m != nullptr, SDT_EINVAL, "autotune needs the SPME mobility");
        std::vector<double> pos(3 * (size_t)b->npos());
        b->get_positions(pos.data());
        m->update(pos.data(), nullptr, false);
        const int wide = std::min(b->mob_interval(), BrwnLanczosGpu::kMaxBatchCols);
        const std::string cache = tune_cache ? tune_cache : "";
        const TuneResult r[2] = {autotune_spme(m, 1, cache), autotune_spme(m, wide, cache)};
        for (int w = 0; w < 2; w++) {
            if (choices) {
                choices[4 * w] = r[w].tuning.fft_order;
                choices[4 * w + 1] = r[w].tuning.interp_sxt;
                choices[4 * w + 2] = r[w].tuning.interp_ch;
                choices[4 * w + 3] = r[w].tuning.spread_stage_kb;
            }
            if (ms) {
                ms[2 * w] = r[w].ms_default;
                ms[2 * w + 1] = r[w].ms_best;
            }
        }
    });
}
int sdt_bd_set_graphs(sdt_bd_t bd, int on) {
    SDT_CHECK_HANDLE(bd);
    bd->bd->set_graphs(on != 0);
    return SDT_OK;
}

// ---------------------------------------------------------------- Serving --
int sdt_serve_create(int kind, double capacity_bytes, int max_active, int porder, int max_iters,
                     int max_nrhs, double tol, double steric_r0, double steric_k0, double dt,
                     int mob_interval, sdt_serve_t* out) {
    if (!out) { set_error("null output handle"); return SDT_EINVAL; }
    *out = nullptr;
    return guard([&] {
        ServeParams p;
        p.kind = kind;
        p.capacity = capacity_bytes;
        p.max_active = max_active;
        p.porder = porder;
        p.max_iters = max_iters;
        p.max_nrhs = max_nrhs;
        p.tol = tol;
        p.r0 = steric_r0;
        p.k0 = steric_k0;
        p.dt = dt;
        p.mob_interval = mob_interval;
        *out = new sdt_serve_s{new ServeGpu(p), p};
    });
}
int sdt_serve_footprint(sdt_serve_t sv, int npos, int dim, double box_size, double rmax,
                        double* bytes) {
    SDT_CHECK_HANDLE(sv);
    if (!bytes || npos <= 0 || dim <= 0 || box_size <= 0.0) {
        set_error("invalid arguments");
        return SDT_EINVAL;
    }
    *bytes = (double)serve_footprint(sv->p, npos, dim, box_size, rmax);
    return SDT_OK;
}
int sdt_serve_submit(sdt_serve_t sv, int npos, const double* pos, const double* rdi,
                     double box_size, double xi, double rmax, int dim, int nsteps, uint64_t seed,
                     double arrival_s, int* job) {
    SDT_CHECK_HANDLE(sv);
    return guard([&] {
        const int id = sv->s->submit(npos, pos, rdi, box_size, xi, rmax, dim, nsteps, seed,
                                     arrival_s);
        if (job) *job = id;
    });
}
int sdt_serve_run(sdt_serve_t sv, double deadline_s) {
    SDT_CHECK_HANDLE(sv);
    return guard([&] { sv->s->run(deadline_s); });
}
int sdt_serve_job(sdt_serve_t sv, int job, int* state, double* times, int* steps_done,
                  double* footprint_bytes) {
    SDT_CHECK_HANDLE(sv);
    if (job < 0 || job >= sv->s->njobs()) { set_error("job index out of range"); return SDT_EINVAL; }
    const ServeJob& j = sv->s->job(job);
    if (state) *state = j.state;
    if (times) {
        times[0] = j.arrival;
        times[1] = j.admit;
        times[2] = j.finish;
    }
    if (steps_done) *steps_done = j.steps_done;
    if (footprint_bytes) *footprint_bytes = (double)j.bytes;
    if (j.state == ServeJob::kFailed) set_error(j.error);
    return SDT_OK;
}
int sdt_serve_job_positions(sdt_serve_t sv, int job, double* pos) {
    SDT_CHECK_HANDLE(sv);
    if (job < 0 || job >= sv->s->njobs() || !pos) { set_error("invalid arguments"); return SDT_EINVAL; }
    const ServeJob& j = sv->s->job(job);
    std::memcpy(pos, j.pos.data(), sizeof(double) * j.pos.size());
    return SDT_OK;
}
int sdt_serve_job_step_ms(sdt_serve_t sv, int job, double* ms, int cap, int* n) {
    SDT_CHECK_HANDLE(sv);
    if (job < 0 || job >= sv->s->njobs()) { set_error("job index out of range"); return SDT_EINVAL; }
    const ServeJob& j = sv->s->job(job);
    if (n) *n = (int)j.step_ms.size();
    for (int i = 0; i < (int)j.step_ms.size() && i < cap && ms; i++) ms[i] = j.step_ms[i];
    return SDT_OK;
}
int sdt_serve_stats(sdt_serve_t sv, double* stats) {
    SDT_CHECK_HANDLE(sv);
    if (!stats) { set_error("null output"); return SDT_EINVAL; }
    stats[0] = sv->s->wall();
    stats[1] = (double)sv->s->peak_bytes();
    stats[2] = sv->s->peak_active();
    return SDT_OK;
}
int sdt_device_memory(double* free_bytes, double* total_bytes) {
    return guard([&] {
        size_t f = 0, t = 0;
        SDT_HIP_CHECK(hipMemGetInfo(&f, &t));
        if (free_bytes) *free_bytes = (double)f;
        if (total_bytes) *total_bytes = (double)t;
    });
}
int sdt_serve_destroy(sdt_serve_t sv) {
    if (!sv) return SDT_OK;
    return guard([&] {
        delete sv->s;
        delete sv;
    });
}

}  // extern "C"
