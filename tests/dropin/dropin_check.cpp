// Drives the C++ drop-in classes of include/stokesdt_amd.hpp through the
// reference's own base-class interfaces (MobBase, BrwnBase, ForceBase,
// PairListBase), as a reference caller would, and writes every input and
// output as text lines "name count v0 v1 ..." (%.17g) to the file given as
// argv[1]. Inputs are drawn from a fixed linear congruential sequence.
// Exit status: 0 on success, 2 when a class reports a failure (for instance
// when no GPU is present), with the message on stderr.
#include <cstdio>
#include <cstdlib>
#include <vector>
#include "stokesdt_amd.hpp"

namespace {

struct Lcg {
    unsigned long long s;
    explicit Lcg(unsigned long long seed) : s(seed) {}
    double next() {   // uniform in [0, 1)
        s = s * 6364136223846793005ULL + 1442695040888963407ULL;
        return (double)(s >> 11) / 9007199254740992.0;
    }
};

void put(FILE* out, const char* name, const double* v, size_t n) {
    std::fprintf(out, "%s %zu", name, n);
    for (size_t i = 0; i < n; i++) std::fprintf(out, " %.17g", v[i]);
    std::fprintf(out, "\n");
}

void puti(FILE* out, const char* name, const int* v, size_t n) {
    std::fprintf(out, "%s %zu", name, n);
    for (size_t i = 0; i < n; i++) std::fprintf(out, " %d", v[i]);
    std::fprintf(out, "\n");
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s <output file>\n", argv[0]);
        return 1;
    }
    const int n = 200, K = 32, p = 6, nrhs = 4;
    const double box = 16.0, xi = 0.9, rmax = 5.0;
    Lcg rng(12345);
    std::vector<double> pos(3 * n), rdi(n), f(3 * n * nrhs), v(3 * n * nrhs);
    for (int i = 0; i < 3 * n; i++) pos[i] = box * rng.next();
    for (int i = 0; i < n; i++) rdi[i] = 0.8 + 0.4 * rng.next();
    for (int i = 0; i < 3 * n * nrhs; i++) f[i] = rng.next() - 0.5;
    FILE* out = std::fopen(argv[1], "w");
    if (!out) return 1;
    put(out, "pos", &pos[0], pos.size());
    put(out, "rdi", &rdi[0], rdi.size());
    put(out, "f", &f[0], f.size());

    // MobSpme through MobBase
    stokesdt_amd::MobSpmeHip spme(n, &rdi[0], box, xi, rmax, K, p);
    stokesdt::MobBase* mob = &spme;
    if (!mob->Init()) return 2;
    mob->Update(&pos[0], &rdi[0]);
    for (size_t i = 0; i < v.size(); i++) v[i] = 1.0;
    mob->MulVector(nrhs, 2.0, 3 * n, &f[0], 0.5, 3 * n, &v[0]);
    if (spme.last_status() != SDT_OK) return 2;
    put(out, "mf", &v[0], v.size());
    std::vector<double> vr(3 * n, 0.0), vk(3 * n, 0.0);
    spme.RealMulVector(1, 1.0, 3 * n, &f[0], 0.0, 3 * n, &vr[0]);
    spme.RecipMulVector(1, 1.0, 3 * n, &f[0], 0.0, 3 * n, &vk[0]);
    put(out, "mf_real", &vr[0], vr.size());
    put(out, "mf_recip", &vk[0], vk.size());
    const int d = mob->dim();
    put(out, "dim", std::vector<double>(1, (double)d).data(), 1);

    // BrwnLanczos through BrwnBase
    stokesdt_amd::BrwnLanczosHip lz(3 * n, 100, 4, 1e-8);
    stokesdt::BrwnBase* brwn = &lz;
    if (!brwn->Init()) return 2;
    std::vector<double> y(3 * n * nrhs, 0.0);
    brwn->Compute(mob, nrhs, 3 * n, &f[0], 3 * n, &y[0]);
    if (lz.last_status() != SDT_OK) return 2;
    put(out, "lanczos", &y[0], y.size());
    std::vector<int> it = lz.iters();
    puti(out, "lanczos_iters", &it[0], it.size());

    // StericForce through ForceBase (Compute zeroes, then accumulates)
    stokesdt_amd::StericForceHip st(n, &rdi[0], box, 2.0, 125.0);
    stokesdt::ForceBase* force = &st;
    if (!force->Init()) return 2;
    std::vector<double> fs(3 * n, 7.0);
    force->Compute(&pos[0], &rdi[0], &fs[0]);
    if (st.last_status() != SDT_OK) return 2;
    put(out, "steric", &fs[0], fs.size());

    // PairListR / PairListS through PairListBase
    const char modes[2] = {'R', 'S'};
    const double cut[2] = {5.0, 2.2};
    for (int m = 0; m < 2; m++) {
        stokesdt_amd::PairListHip pl(n, &rdi[0], box, cut[m], modes[m]);
        stokesdt::PairListBase* base = &pl;
        if (base->Init() <= 0) return 2;
        const int nnz = base->Build(&pos[0]);
        if (pl.last_status() != SDT_OK || nnz <= 0) return 2;
        std::vector<int> rowptr(n + 1), colidx(nnz);
        base->GetPairs(&rowptr[0], &colidx[0]);
        puti(out, m == 0 ? "pairs_R_rowptr" : "pairs_S_rowptr", &rowptr[0], rowptr.size());
        puti(out, m == 0 ? "pairs_R_colidx" : "pairs_S_colidx", &colidx[0], colidx.size());
    }

    // MobEwald + BrwnChol
    stokesdt_amd::MobEwaldHip ew(n, &rdi[0], box, 1e-10);
    if (!ew.Init()) return 2;
    ew.Update(&pos[0], &rdi[0]);
    std::vector<double> ve(3 * n, 0.0);
    ew.MulVector(1, 1.0, 3 * n, &f[0], 0.0, 3 * n, &ve[0]);
    put(out, "ewald_mf", &ve[0], ve.size());
    stokesdt_amd::BrwnCholHip ch(3 * n);
    if (!ch.Init()) return 2;
    std::vector<double> yc(3 * n, 0.0);
    static_cast<stokesdt::BrwnBase*>(&ch)->Compute(&ew, 1, 3 * n, &f[0], 3 * n, &yc[0]);
    if (ch.last_status() != SDT_OK) return 2;
    put(out, "chol", &yc[0], yc.size());
    std::fclose(out);
    std::printf("drop-in classes: all calls succeeded\n");
    return 0;
}
