// Oracle SPME engine, real-space BSR and MobSpme (test infrastructure only;
// see orc_common.hpp).
#pragma once
#include <vector>
#include "orc_common.hpp"
#include "orc_pairlist.hpp"

namespace orc {

// B-spline M_p(x + 1) by the Cox-de Boor recursion
// (src/mob/spme.cc:38-68, `BSpline`).
double bspline(int porder, double x);

struct SpmeEngine {
    // parameters (src/mob/spme.h:24-81)
    double xi = 0, box = 0;
    int npos = 0, porder = 0, dim = 0;
    bool poly = false;                 // rdi2 != NULL in the reference
    bool wrap_fix = true;              // wrap positions into [0, L) first
    std::vector<double> rdi2;
    // mesh layout: in-place R2C, ld1 = 2(K/2+1), ld2 = K*ld1, ld3 = K*ld2
    int ld1 = 0, ld2 = 0, ld3 = 0;
    std::vector<double> grid;          // 3 * ld3
    std::vector<double> map;           // K wavenumbers
    std::vector<double> lm2;           // K*K*(K/2+1)
    // sparse interpolation matrix (src/mob/spme.cc:625-633)
    int ldP = 0;
    std::vector<double> P;
    std::vector<int> ind;
    // independent-set blocks (src/mob/spme.cc:635-677)
    int sizeb = 0, nb = 0;
    std::vector<int> block_start, block_members;   // CSR of particles/block
    FFT3D fft;

    bool create(int npos, const double* rdi, double box, double xi, int dim,
                int porder);
    void update(const double* pos);
    void spread(const double* rdi2_or_null, const double* f);
    void apply_influence(int flag);
    void interpolate(const double* rdi2_or_null, double alpha, double beta,
                     double* v) const;
    void recip(int nrhs, double alpha, int ldin, const double* in,
               double beta, int ldout, double* out);
};

// 3x3 block CSR (src/util/sparse.h:21-38) and SpMV3x3
// (src/util/sparse.cc:79-153).
struct BSR3 {
    int nrowb = 0;
    std::vector<int> rowptr, colidx;
    std::vector<double> val;     // 9 per block, row-major
    void spmv(int nrhs, double alpha, int ldx, const double* x, double beta,
              int ldy, double* y) const;
};

// MobSpme (src/mob/mob_spme.cc:19-262).
class MobSpme : public Mob {
  public:
    MobSpme(int npos, const double* rdi, double box, double xi, double rmax,
            int dim, int porder);
    bool init();
    void update(const double* pos, const double* rdi);
    int dim() const override { return 3 * npos_; }
    void mul_vector(int nrhs, double alpha, int ldf, const double* f,
                    double beta, int ldv, double* v) override;
    void real_mul_vector(int nrhs, double alpha, int ldf, const double* f,
                         double beta, int ldv, double* v);
    void recip_mul_vector(int nrhs, double alpha, int ldf, const double* f,
                          double beta, int ldv, double* v);
    const BSR3& real_matrix() const { return real_; }
    SpmeEngine& engine() { return spme_; }

  private:
    int npos_;
    double box_, xi_, rmax_;
    int dim_, porder_;
    std::vector<double> rdi_;
    PairList pairs_;
    SpmeEngine spme_;
    BSR3 real_;
};

// scalars_ewald_F (src/mob/ewald.h:55-85)
void scalars_ewald(double xi, double r, double aa, double ab, double* xa,
                   double* ya);

}  // namespace orc
