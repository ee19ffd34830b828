// Oracle C ABI for the Python test harness (test infrastructure only; see
// orc_common.hpp). Also holds the steric force and the BD step driver.
#include <cmath>
#include <cstring>
#include <vector>
#include <omp.h>
#include "orc_common.hpp"
#include "orc_lanczos.hpp"
#include "orc_pairlist.hpp"
#include "orc_spme.hpp"

namespace orc {
void ewald_matrix(int npos, const double* pos, const double* rdi, double box,
                  double xi, double tol, int ldm, double* mat, double* mat_real,
                  double* mat_recip, int* nr_out, int* nk_out);
double ewald_auto_xi(double box);

// StericForce::ForceKernel (src/force/force_steric.cc:14-53) over the S
// pair list with cutoff r0.
static void steric_accumulate(int npos, const double* pos, const double* rdi,
                              double box, double r0, double k0,
                              const int* rowptr, const int* colidx, double* f) {
    #pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < npos; i++) {
        double x1 = pos[3 * i], y1 = pos[3 * i + 1], z1 = pos[3 * i + 2];
        double aa = rdi[i];
        for (int k = rowptr[i]; k < rowptr[i + 1]; k++) {
            int j = colidx[k];
            if (j == i) continue;
            double rx = remainder(x1 - pos[3 * j], box);
            double ry = remainder(y1 - pos[3 * j + 1], box);
            double rz = remainder(z1 - pos[3 * j + 2], box);
            double r = sqrt(rx * rx + ry * ry + rz * rz);
            double ex = rx / r, ey = ry / r, ez = rz / r;
            double force = k0 * ((aa + rdi[j]) * r0 / 2.0 - r);
            f[3 * i] += force * ex;
            f[3 * i + 1] += force * ey;
            f[3 * i + 2] += force * ez;
        }
    }
}

// BDSimBox::Advance (src/simbox/simbox_bd.cc:296-366) with SPME mobility,
// block Lanczos displacements and steric forces; kT = 1. Gaussian block
// `blk` (steps blk*M_I .. ) is Philox stream `blk`, column-major M_I x 3N.
struct BDSim {
    int npos;
    double box, dt;
    int mob_interval;
    uint64_t seed;
    std::vector<double> pos, rdi;
    MobSpme mob;
    BrwnLanczos brwn;
    PairList steric_pairs;
    double r0, k0;
    int ldm;
    std::vector<double> Z, D, f;
    std::vector<int> rowptr, colidx;
    long istep = 0;
    BDSim(int n, const double* p, const double* r, double L, double xi, double rmax,
          int K, int porder, int max_iters, int max_nrhs, double tol, double r0_,
          double k0_, double dt_, int mi, uint64_t seed_)
        : npos(n), box(L), dt(dt_), mob_interval(mi), seed(seed_), pos(p, p + 3 * n),
          rdi(r, r + n), mob(n, r, L, xi, rmax, K, porder),
          brwn(3 * n, max_iters, max_nrhs, tol), steric_pairs(n, r, L, r0_, 'S'),
          r0(r0_), k0(k0_) {
        ldm = (int)pad_len(3 * n, sizeof(double));
    }
    bool init() {
        if (!mob.init()) return false;
        if (!brwn.init()) return false;
        if (steric_pairs.init() <= 0) return false;
        // M_I columns (the reference allocates mob_interval * dim * ldm,
        // src/simbox/simbox_bd.cc:171-175, 3N times what it uses)
        Z.assign((size_t)mob_interval * ldm, 0.0);
        D.assign((size_t)mob_interval * ldm, 0.0);
        f.assign(ldm, 0.0);
        rowptr.assign(npos + 1, 0);
        return true;
    }
    void advance(int nsteps) {
        const int n3 = 3 * npos;
        for (int s = 0; s < nsteps; s++) {
            int irhs = (int)(istep % mob_interval);
            if (irhs == 0) {
                long blk = istep / mob_interval;
                std::vector<double> g((size_t)mob_interval * n3);
                gaussian_fill(seed, (uint32_t)blk, (int64_t)g.size(), g.data());
                for (int c = 0; c < mob_interval; c++)
                    memcpy(&Z[(size_t)c * ldm], &g[(size_t)c * n3], sizeof(double) * n3);
                mob.update(pos.data(), rdi.data());
                brwn.compute(&mob, mob_interval, ldm, Z.data(), ldm, D.data());
            }
            double* disp = &D[(size_t)irhs * ldm];
            std::fill(f.begin(), f.end(), 0.0);
            int nnz = steric_pairs.build(pos.data());
            colidx.assign(nnz, 0);
            steric_pairs.get_pairs(rowptr.data(), colidx.data());
            steric_accumulate(npos, pos.data(), rdi.data(), box, r0, k0, rowptr.data(),
                              colidx.data(), f.data());
            mob.mul_vector(1, dt, ldm, f.data(), sqrt(2.0 * dt), ldm, disp);
            for (int i = 0; i < n3; i++) pos[i] += disp[i];
            istep++;
        }
    }
};

}  // namespace orc

using namespace orc;

extern "C" {

double orc_bspline(int p, double x) { return bspline(p, x); }

void orc_philox(const uint32_t* ctr, const uint32_t* key, uint32_t* out) {
    philox4x32_10(ctr, key, out);
}

void orc_gaussian(uint64_t seed, uint32_t stream, int64_t n, double* out) {
    gaussian_fill(seed, stream, n, out);
}

int orc_num_threads(void) { return omp_get_max_threads(); }
void orc_set_num_threads(int n) { omp_set_num_threads(n); }

// in-place R2C / C2R on a [z][y][2(K/2+1)] mesh
void orc_fft3d_forward(int K, double* mesh) {
    FFT3D f; f.init(K); f.forward(mesh);
}
void orc_fft3d_backward(int K, double* mesh) {
    FFT3D f; f.init(K); f.backward(mesh);
}
void orc_fft1d(int n, double* xy, int sign) {
    FFTPlan1D p; p.init(n);
    std::vector<cplx> w(n);
    p.exec(reinterpret_cast<cplx*>(xy), w.data(), sign);
}

void orc_sym_eig(int n, const double* A, double* w, double* V) {
    sym_eig(n, A, n, w, V, n);
}
void orc_tridiag_eig(int n, const double* d, const double* e, double* w, double* V) {
    tridiag_eig(n, d, e, w, V, n);
}
void orc_householder_qr(int m, int n, double* A, double* R) {
    householder_qr(m, n, A, m, R, n);
}
int orc_cholesky(int n, double* A) { return cholesky_lower(n, A, n); }

// ---- SPME engine pieces (for unit tests) ----
void* orc_spme_create(int npos, const double* rdi, double box, double xi, int dim,
                      int porder) {
    SpmeEngine* e = new SpmeEngine();
    if (!e->create(npos, rdi, box, xi, dim, porder)) { delete e; return nullptr; }
    return e;
}
void orc_spme_destroy(void* h) { delete static_cast<SpmeEngine*>(h); }
void orc_spme_update(void* h, const double* pos) { static_cast<SpmeEngine*>(h)->update(pos); }
int orc_spme_ld3(void* h) { return static_cast<SpmeEngine*>(h)->ld3; }
// spread f (3N) into the three meshes and copy them out (3*ld3 doubles)
void orc_spme_spread(void* h, const double* f, double* mesh_out) {
    SpmeEngine* e = static_cast<SpmeEngine*>(h);
    e->spread(nullptr, f);
    memcpy(mesh_out, e->grid.data(), sizeof(double) * 3 * (size_t)e->ld3);
}
// interpolate from given meshes
void orc_spme_interp(void* h, const double* mesh_in, double alpha, double* v) {
    SpmeEngine* e = static_cast<SpmeEngine*>(h);
    memcpy(e->grid.data(), mesh_in, sizeof(double) * 3 * (size_t)e->ld3);
    e->interpolate(nullptr, alpha, 0.0, v);
}
void orc_spme_lm2(void* h, double* out) {
    SpmeEngine* e = static_cast<SpmeEngine*>(h);
    memcpy(out, e->lm2.data(), sizeof(double) * e->lm2.size());
}
void orc_spme_recip(void* h, int nrhs, double alpha, int ldin, const double* in,
                    double beta, int ldout, double* out) {
    static_cast<SpmeEngine*>(h)->recip(nrhs, alpha, ldin, in, beta, ldout, out);
}

// ---- MobSpme ----
void* orc_mob_spme_create(int npos, const double* rdi, double box, double xi,
                          double rmax, int dim, int porder) {
    return new MobSpme(npos, rdi, box, xi, rmax, dim, porder);
}
int orc_mob_spme_init(void* h) { return static_cast<MobSpme*>(h)->init() ? 1 : 0; }
void orc_mob_spme_update(void* h, const double* pos, const double* rdi) {
    static_cast<MobSpme*>(h)->update(pos, rdi);
}
void orc_mob_spme_mulvector(void* h, int mode, int nrhs, double alpha, int ldf,
                            const double* f, double beta, int ldv, double* v) {
    MobSpme* m = static_cast<MobSpme*>(h);
    if (mode == 0) m->mul_vector(nrhs, alpha, ldf, f, beta, ldv, v);
    else if (mode == 1) m->real_mul_vector(nrhs, alpha, ldf, f, beta, ldv, v);
    else m->recip_mul_vector(nrhs, alpha, ldf, f, beta, ldv, v);
}
int orc_mob_spme_real_nnz(void* h) {
    return (int)static_cast<MobSpme*>(h)->real_matrix().colidx.size();
}
void orc_mob_spme_destroy(void* h) { delete static_cast<MobSpme*>(h); }

// ---- pair lists ----
void* orc_pairlist_create(int npos, const double* rdi, double box, double cutoff, char mode) {
    return new PairList(npos, rdi, box, cutoff, mode);
}
int orc_pairlist_init(void* h) { return static_cast<PairList*>(h)->init(); }
int orc_pairlist_build(void* h, const double* pos) { return static_cast<PairList*>(h)->build(pos); }
void orc_pairlist_get(void* h, int* rowptr, int* colidx) {
    static_cast<PairList*>(h)->get_pairs(rowptr, colidx);
}
int orc_pairlist_nc1(void* h) { return static_cast<PairList*>(h)->nc1(); }
void orc_pairlist_destroy(void* h) { delete static_cast<PairList*>(h); }

// ---- steric ----
// ForceBase::Compute (src/force/force_base.cc:24-28): zero then accumulate
int orc_steric_compute(int npos, const double* pos, const double* rdi, double box,
                       double r0, double k0, double* f) {
    PairList pl(npos, rdi, box, r0, 'S');
    if (pl.init() <= 0) return -1;
    int nnz = pl.build(pos);
    std::vector<int> rowptr(npos + 1), colidx(nnz);
    pl.get_pairs(rowptr.data(), colidx.data());
    memset(f, 0, sizeof(double) * 3 * npos);
    steric_accumulate(npos, pos, rdi, box, r0, k0, rowptr.data(), colidx.data(), f);
    return nnz;
}

// ---- dense Ewald ----
double orc_ewald_auto_xi(double box) { return ewald_auto_xi(box); }
void orc_ewald_matrix(int npos, const double* pos, const double* rdi, double box,
                      double xi, double tol, double* mat, double* mat_real,
                      double* mat_recip, int* nr, int* nk) {
    ewald_matrix(npos, pos, rdi, box, xi, tol, 3 * npos, mat, mat_real, mat_recip, nr, nk);
}

// ---- Lanczos / Cholesky ----
// mob_kind 0: dense row-major matrix `op`; 1: MobSpme handle `op`
int orc_lanczos(int mob_kind, void* op, int dim, int max_iters, int max_nrhs,
                double tol, int nrhs, int ldz, const double* z, int ldy, double* y,
                int* iters_out) {
    DenseMob* dm = nullptr;
    Mob* mob;
    if (mob_kind == 0) { dm = new DenseMob(dim, static_cast<const double*>(op), dim); mob = dm; }
    else mob = static_cast<MobSpme*>(op);
    BrwnLanczos bl(dim, max_iters, max_nrhs, tol);
    if (!bl.init()) { delete dm; return -1; }
    bl.compute(mob, nrhs, ldz, z, ldy, y);
    int ngroups = (int)bl.iters().size();
    for (int g = 0; g < ngroups && iters_out; g++) iters_out[g] = bl.iters()[g];
    delete dm;
    return ngroups;
}
int orc_chol_disp(int n, const double* mat, int nrhs, int ldz, const double* z, int ldy,
                  double* y) {
    return chol_displacements(n, mat, n, nrhs, ldz, z, ldy, y);
}

// ---- BD driver ----
void* orc_bd_create(int npos, const double* pos, const double* rdi, double box,
                    double xi, double rmax, int dim, int porder, int max_iters,
                    int max_nrhs, double tol, double r0, double k0, double dt,
                    int mob_interval, uint64_t seed) {
    BDSim* s = new BDSim(npos, pos, rdi, box, xi, rmax, dim, porder, max_iters,
                         max_nrhs, tol, r0, k0, dt, mob_interval, seed);
    if (!s->init()) { delete s; return nullptr; }
    return s;
}
void orc_bd_advance(void* h, int nsteps) { static_cast<BDSim*>(h)->advance(nsteps); }
void orc_bd_get_pos(void* h, double* pos) {
    BDSim* s = static_cast<BDSim*>(h);
    memcpy(pos, s->pos.data(), sizeof(double) * s->pos.size());
}
int orc_bd_last_iters(void* h, int* out, int cap) {
    BDSim* s = static_cast<BDSim*>(h);
    int n = (int)s->brwn.iters().size();
    for (int i = 0; i < n && i < cap; i++) out[i] = s->brwn.iters()[i];
    return n;
}
void orc_bd_destroy(void* h) { delete static_cast<BDSim*>(h); }

}  // extern "C"
