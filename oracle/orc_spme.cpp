// Oracle SPME (test infrastructure only; see orc_common.hpp).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <omp.h>
#include "orc_spme.hpp"

namespace orc {

// src/mob/spme.cc:38-68. spline[j] holds M_i(x - j + 1) while the order i
// is raised from 1 to porder; the result is M_porder(x + 1).
double bspline(int porder, double x) {
    std::vector<double> s(porder + 1, 0.0);
    int p = (int)floor(x) + 1;
    if (p >= 0 && p < porder) s[p] = 1.0;
    for (int i = 2; i <= porder; i++) {
        for (int j = 0; j < porder + 1 - i; j++) {
            s[j] = ((x - j + 1.0) * s[j] + (i - x + j - 1.0) * s[j + 1]) / (i - 1.0);
        }
    }
    return s[0];
}

// src/mob/spme.h:114-123
static inline double scalar_recip(double k, double xi) {
    double kx = k / xi;
    double kx2 = kx * kx;
    double k2 = k * k;
    return 6.0 * M_PI * exp(-kx2 / 4.0) / k2 * (1.0 + kx2 * (1.0 / 4.0 + kx2 / 8.0));
}

bool SpmeEngine::create(int npos_, const double* rdi, double box_, double xi_,
                        int dim_, int porder_) {
    // src/mob/spme.cc:512-690
    npos = npos_; box = box_; xi = xi_; dim = dim_; porder = porder_;
    poly = (rdi != nullptr);
    if (poly) {
        rdi2.resize(npos);
        for (int i = 0; i < npos; i++) rdi2[i] = rdi[i] * rdi[i];
    }
    const int K = dim;
    ld1 = 2 * (K / 2 + 1);
    ld2 = K * ld1;
    ld3 = K * ld2;
    grid.assign((size_t)3 * ld3, 0.0);
    fft.init(K);

    // ComputeSpline (src/mob/spme.cc:71-160): wavenumbers, B-spline moduli
    // and lm2 = ScalarRecip(|k|) * |b_x b_y b_z|^2
    map.resize(K);
    for (int i = 0; i < K; i++) {
        int m = (i <= K / 2) ? i : i - K;
        map[i] = m * 2.0 * M_PI / box;
    }
    std::vector<double> sv(porder - 1);
    for (int i = 0; i < porder - 1; i++) sv[i] = bspline(porder, i);
    std::vector<cplx> b(K);
    std::vector<double> den2(K);
    for (int x = 0; x < K; x++) {
        double kx = map[x];
        cplx den(0.0, 0.0);
        for (int i = 0; i < porder - 1; i++) {
            double phi = i * kx * box / K;
            den += sv[i] * cplx(cos(phi), sin(phi));
        }
        den2[x] = std::norm(den);
    }
    // Deviation from src/mob/spme.cc:110-128 (a defect there): for odd
    // porder the sum vanishes at the Nyquist index and rounding leaves
    // |den|^2 ~ 1e-33, amplifying that plane by ~1e33. |den|^2 is replaced
    // there by the mean of its neighbours (Essmann et al. 1995); even porder
    // is unchanged (|den|^2 > 1e-3 everywhere).
    for (int x = 0; x < K; x++)
        if (den2[x] < 1e-7) den2[x] = 0.5 * (den2[(x + K - 1) % K] + den2[(x + 1) % K]);
    for (int x = 0; x < K; x++) {
        // only |b|^2 enters lm2: b = e^{i phi0} / den with |den|^2 as above
        double phi0 = box * map[x] * (porder - 1) / K;
        b[x] = cplx(cos(phi0), sin(phi0)) / sqrt(den2[x]);
    }
    const int Kh = K / 2 + 1;
    lm2.assign((size_t)K * K * Kh, 0.0);
    for (int j = 0; j < K * K; j++) {
        int z = j / K, y = j % K;
        double kz = map[z], ky = map[y];
        for (int x = 0; x < Kh; x++) {
            double kx = map[x];
            double k = sqrt(kx * kx + ky * ky + kz * kz);
            double m2 = scalar_recip(k, xi);
            cplx bbb = b[x] * b[y] * b[z];
            lm2[(size_t)j * Kh + x] = m2 * std::norm(bbb);
        }
    }
    const int p3 = porder * porder * porder;
    ldP = (int)pad_len(p3, sizeof(double));
    P.assign((size_t)npos * ldP, 0.0);
    ind.assign((size_t)npos * ldP, 0);
    // independent sets: block side = smallest divisor of K that is >= p
    sizeb = porder;
    while (K % sizeb != 0) sizeb++;
    nb = K / sizeb;
    if (nb < 4) return false;   // src/mob/spme.cc:643-645
    return true;
}

// src/mob/spme.cc:366-509. With wrap_fix the position is first wrapped into
// [0, L) (the reference computes the wrap and then discards it,
// src/mob/spme.cc:404-412; for positions already in [0, L) the two agree).
void SpmeEngine::update(const double* pos) {
    const int K = dim, p = porder, p2 = p * p;
    std::vector<int> blk(npos);
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < npos; i++) {
        double c[3];
        for (int d = 0; d < 3; d++) {
            double x = pos[3 * i + d];
            if (wrap_fix) {
                x = fmod(x, box);
                x = (x >= 0.0 ? x : box + x);
            }
            c[d] = x / box * K;
        }
        int bidx3[3];
        int base[3];
        double q[3][16];
        for (int d = 0; d < 3; d++) {
            bidx3[d] = (int)(c[d] / sizeb);
            if (bidx3[d] >= nb) bidx3[d] = nb - 1;
            double g = floor(c[d]);
            int b0 = (p % 2 == 0) ? (int)(g - p / 2 + 1) : (int)(g - (p - 1) / 2);
            base[d] = ((b0 % K) + K) % K;
            for (int x = 0; x < p; x++) q[d][x] = bspline(p, g - c[d] + x);
        }
        blk[i] = (bidx3[2] * nb + bidx3[1]) * nb + bidx3[0];
        for (int z = 0; z < p; z++) {
            int zz = (z + base[2]) % K;
            for (int y = 0; y < p; y++) {
                int yy = (y + base[1]) % K;
                for (int x = 0; x < p; x++) {
                    int xx = (x + base[0]) % K;
                    P[(size_t)i * ldP + z * p2 + y * p + x] = q[0][x] * q[1][y] * q[2][z];
                    ind[(size_t)i * ldP + z * p2 + y * p + x] = zz * ld2 + yy * ld1 + xx;
                }
            }
        }
    }
    // per-block particle lists (counting sort by block)
    const int nb3 = nb * nb * nb;
    block_start.assign(nb3 + 1, 0);
    for (int i = 0; i < npos; i++) block_start[blk[i] + 1]++;
    for (int b = 0; b < nb3; b++) block_start[b + 1] += block_start[b];
    block_members.assign(npos, 0);
    std::vector<int> fill(block_start.begin(), block_start.end() - 1);
    for (int i = 0; i < npos; i++) block_members[fill[blk[i]]++] = i;
}

// src/mob/spme.cc:313-363: zero the mesh, then sweep the 8 parity classes of
// blocks; blocks of one class are >= 2 blocks apart, so their stencils never
// overlap and each class is spread in parallel without atomics. For odd nb
// the classes are not independent (a reference defect); the sweep is then
// done serially.
void SpmeEngine::spread(const double* scale2, const double* f) {
    const int p3 = porder * porder * porder;
    double* g0 = grid.data();
    #pragma omp parallel for schedule(static)
    for (size_t i = 0; i < (size_t)3 * ld3; i++) g0[i] = 0.0;
    auto spread_one = [&](int i) {
        double f0 = f[3 * i], f1 = f[3 * i + 1], f2 = f[3 * i + 2];
        if (scale2) { f0 *= scale2[i]; f1 *= scale2[i]; f2 *= scale2[i]; }
        const double* Pi = &P[(size_t)i * ldP];
        const int* Ii = &ind[(size_t)i * ldP];
        for (int j = 0; j < p3; j++) {
            int idx = Ii[j];
            double pv = Pi[j];
            g0[idx] += pv * f0;
            g0[ld3 + idx] += pv * f1;
            g0[2 * (size_t)ld3 + idx] += pv * f2;
        }
    };
    if (nb % 2 == 0) {
        for (int set = 0; set < 8; set++) {
            int sx = set & 1, sy = (set >> 1) & 1, sz = (set >> 2) & 1;
            #pragma omp parallel for collapse(3) schedule(dynamic)
            for (int bz = sz; bz < nb; bz += 2)
                for (int by = sy; by < nb; by += 2)
                    for (int bx = sx; bx < nb; bx += 2) {
                        int b = (bz * nb + by) * nb + bx;
                        for (int k = block_start[b]; k < block_start[b + 1]; k++)
                            spread_one(block_members[k]);
                    }
        }
    } else {
        for (int i = 0; i < npos; i++) spread_one(i);
    }
}

// src/mob/spme.cc:201-310 (flag 0: mono with s = 1 - k^2/3; flag 1: s = 1;
// flag 2: s = -k^2/6), then k = 0 is zeroed.
void SpmeEngine::apply_influence(int flag) {
    const int K = dim, Kh = K / 2 + 1;
    double* g = grid.data();
    #pragma omp parallel for schedule(static)
    for (int j = 0; j < K * K; j++) {
        int z = j / K, y = j % K;
        double kz = map[z], ky = map[y];
        double kyz = ky * ky + kz * kz;
        const double* mm = &lm2[(size_t)j * Kh];
        for (int x = 0; x < Kh; x++) {
            double kx = map[x];
            double m2 = mm[x];
            double kk = kx * kx + kyz;
            double kk_1 = 1.0 / kk;
            if (flag == 0) m2 = m2 * (1.0 - kk / 3.0);
            else if (flag == 2) m2 = m2 * (-kk / 6.0);
            double exx = kx * kx * kk_1, eyy = ky * ky * kk_1, ezz = kz * kz * kk_1;
            double exy = kx * ky * kk_1, eyz = ky * kz * kk_1, exz = kx * kz * kk_1;
            double B0 = m2 - m2 * exx, B1 = -m2 * exy, B2 = -m2 * exz;
            double B3 = m2 - m2 * eyy, B4 = -m2 * eyz, B5 = m2 - m2 * ezz;
            size_t i = (size_t)z * ld2 + (size_t)y * ld1 + 2 * x;
            for (int part = 0; part < 2; part++) {
                double v0 = g[i + part], v1 = g[ld3 + i + part], v2 = g[2 * (size_t)ld3 + i + part];
                g[i + part] = B0 * v0 + B1 * v1 + B2 * v2;
                g[ld3 + i + part] = B1 * v0 + B3 * v1 + B4 * v2;
                g[2 * (size_t)ld3 + i + part] = B2 * v0 + B4 * v1 + B5 * v2;
            }
        }
    }
    for (int c = 0; c < 3; c++) {
        g[(size_t)c * ld3] = 0.0;
        g[(size_t)c * ld3 + 1] = 0.0;
    }
}

// src/mob/spme.cc:163-198 and spme.h:219-265
void SpmeEngine::interpolate(const double* scale2, double alpha, double beta,
                             double* v) const {
    const int p3 = porder * porder * porder;
    const double* g = grid.data();
    #pragma omp parallel for schedule(static)
    for (int i = 0; i < npos; i++) {
        double a0 = scale2 ? scale2[i] * alpha : alpha;
        const double* Pi = &P[(size_t)i * ldP];
        const int* Ii = &ind[(size_t)i * ldP];
        double t0 = 0.0, t1 = 0.0, t2 = 0.0;
        for (int j = 0; j < p3; j++) {
            int idx = Ii[j];
            double pv = Pi[j];
            t0 += g[idx] * pv;
            t1 += g[ld3 + idx] * pv;
            t2 += g[2 * (size_t)ld3 + idx] * pv;
        }
        v[3 * i + 0] = beta * v[3 * i + 0] + a0 * t0;
        v[3 * i + 1] = beta * v[3 * i + 1] + a0 * t1;
        v[3 * i + 2] = beta * v[3 * i + 2] + a0 * t2;
    }
}

// src/mob/spme.cc:714-839: a per-RHS loop; mono = one mesh round trip,
// polydisperse = three.
void SpmeEngine::recip(int nrhs, double alpha, int ldin, const double* in,
                       double beta, int ldout, double* out) {
    const double K3 = (double)dim * dim * dim;
    const double alpha0 = alpha * K3 / (box * box * box);
    auto fwd = [&]() { for (int c = 0; c < 3; c++) fft.forward(&grid[(size_t)c * ld3]); };
    auto bwd = [&]() { for (int c = 0; c < 3; c++) fft.backward(&grid[(size_t)c * ld3]); };
    for (int r = 0; r < nrhs; r++) {
        const double* vin = in + (size_t)ldin * r;
        double* vout = out + (size_t)ldout * r;
        if (!poly) {
            spread(nullptr, vin); fwd(); apply_influence(0); bwd();
            interpolate(nullptr, alpha0, beta, vout);
        } else {
            spread(nullptr, vin); fwd(); apply_influence(1); bwd();
            interpolate(nullptr, alpha0, beta, vout);
            spread(rdi2.data(), vin); fwd(); apply_influence(2); bwd();
            interpolate(nullptr, alpha0, 1.0, vout);
            spread(nullptr, vin); fwd(); apply_influence(2); bwd();
            interpolate(rdi2.data(), alpha0, 1.0, vout);
        }
    }
}

// src/util/sparse.cc:79-153 (beta applies at the first block of each row;
// the self block guarantees every row has one)
void BSR3::spmv(int nrhs, double alpha, int ldx, const double* x, double beta,
                int ldy, double* y) const {
    #pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < nrowb; i++) {
        int row = 3 * i;
        double b = beta;
        for (int j = rowptr[i]; j < rowptr[i + 1]; j++) {
            int col = 3 * colidx[j];
            const double* a = &val[(size_t)j * 9];
            for (int k = 0; k < nrhs; k++) {
                const double* xk = x + (size_t)k * ldx + col;
                double* yk = y + (size_t)k * ldy + row;
                double t0 = a[0] * xk[0] + a[3] * xk[1] + a[6] * xk[2];
                double t1 = a[1] * xk[0] + a[4] * xk[1] + a[7] * xk[2];
                double t2 = a[2] * xk[0] + a[5] * xk[1] + a[8] * xk[2];
                yk[0] = b * yk[0] + alpha * t0;
                yk[1] = b * yk[1] + alpha * t1;
                yk[2] = b * yk[2] + alpha * t2;
            }
            b = 1.0;
        }
    }
}

void scalars_ewald(double xi, double r, double aa, double ab, double* xa,
                   double* ya) {
    // src/mob/ewald.h:55-85
    double xir = xi * r;
    double xir2 = xir * xir;
    double r2 = r * r;
    double r3 = r2 * r;
    double aa2 = aa * aa, ab2 = ab * ab;
    double erfcxir = erfc(xir);
    double expxir2 = xi / sqrt(M_PI) * exp(-xir2);
    double a0 = 2.0 * (-3.0 + xir2 * 2.0) * expxir2 + erfcxir / r;
    double b0 = 2.0 * (1.0 + xir2 * -2.0) * expxir2 + erfcxir / r;
    double a2 = 4.0 / r2 * (1.0 + xir2 * (14.0 + xir2 * (-20.0 + xir2 * 4.0))) * expxir2 +
                2.0 / r3 * erfcxir;
    double b2 = -4.0 / r2 * (3.0 + xir2 * (2.0 + xir2 * (-16.0 + xir2 * 4.0))) * expxir2 -
                6.0 / r3 * erfcxir;
    double a12 = a2 * (aa2 + ab2) / 6.0;
    double a22 = b2 * (aa2 + ab2) / 6.0;
    *ya = 0.75 * (a0 + a12);
    *xa = 0.75 * (b0 + a22) + *ya;
}

MobSpme::MobSpme(int npos, const double* rdi, double box, double xi, double rmax,
                 int dim, int porder)
    : npos_(npos), box_(box), xi_(xi), rmax_(rmax), dim_(dim), porder_(porder),
      rdi_(rdi, rdi + npos), pairs_(npos, rdi, box, rmax, 'R') {}

// src/mob/mob_spme.cc:53-132
bool MobSpme::init() {
    if (npos_ <= 0 || box_ <= 0.0 || rmax_ <= 0.0 || xi_ <= 0.0 || dim_ <= 0) return false;
    if (porder_ <= 1) return false;
    if (2.0 * rmax_ >= box_) return false;
    if (4 * porder_ > dim_) return false;
    double mx = *std::max_element(rdi_.begin(), rdi_.end());
    double mn = *std::min_element(rdi_.begin(), rdi_.end());
    bool mono = (mx == 1.0 && mn == 1.0);
    if (!spme_.create(npos_, mono ? nullptr : rdi_.data(), box_, xi_, dim_, porder_))
        return false;
    if (pairs_.init() <= 0) return false;
    real_.nrowb = npos_;
    return true;
}

// src/mob/mob_spme.cc:135-216
void MobSpme::update(const double* pos, const double* rdi) {
    spme_.update(pos);
    int nnz = pairs_.build(pos);
    real_.rowptr.assign(npos_ + 1, 0);
    real_.colidx.assign(nnz, 0);
    real_.val.assign((size_t)nnz * 9, 0.0);
    pairs_.get_pairs(real_.rowptr.data(), real_.colidx.data());
    const double xi2 = xi_ * xi_;
    const double xiaspi = xi_ / sqrt(M_PI);
    #pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < npos_; i++) {
        double x1 = pos[3 * i], y1 = pos[3 * i + 1], z1 = pos[3 * i + 2];
        double aa = rdi[i], aa2 = aa * aa;
        double self_a = 1.0 / aa - xiaspi * (6.0 - 40.0 / 3.0 * xi2 * aa2);
        for (int k = real_.rowptr[i]; k < real_.rowptr[i + 1]; k++) {
            int j = real_.colidx[k];
            double* v = &real_.val[(size_t)k * 9];
            if (j != i) {
                double ab = rdi[j], ab2 = ab * ab;
                double rx = remainder(pos[3 * j] - x1, box_);
                double ry = remainder(pos[3 * j + 1] - y1, box_);
                double rz = remainder(pos[3 * j + 2] - z1, box_);
                double rr = rx * rx + ry * ry + rz * rz;
                double r = sqrt(rr);
                double xa, ya;
                scalars_ewald(xi_, r, aa, ab, &xa, &ya);
                if (r < aa + ab) {
                    xa += -(1.5 - (aa2 + ab2) * 0.5 / rr) / r + 2.0 / (aa + ab) -
                          3.0 * r / 8.0 / (aa2 + ab2);
                    ya += -(0.75 + (aa2 + ab2) * 0.25 / rr) / r + 2.0 / (aa + ab) -
                          9.0 * r / 16.0 / (aa2 + ab2);
                }
                double ex = rx / r, ey = ry / r, ez = rz / r;
                double d = xa - ya;
                v[0] = d * ex * ex + ya; v[1] = d * ex * ey; v[2] = d * ex * ez;
                v[3] = d * ex * ey; v[4] = d * ey * ey + ya; v[5] = d * ey * ez;
                v[6] = d * ex * ez; v[7] = d * ey * ez; v[8] = d * ez * ez + ya;
            } else {
                for (int t = 0; t < 9; t++) v[t] = 0.0;
                v[0] = v[4] = v[8] = self_a;
            }
        }
    }
}

// src/mob/mob_spme.cc:219-236
void MobSpme::mul_vector(int nrhs, double alpha, int ldf, const double* f,
                         double beta, int ldv, double* v) {
    spme_.recip(nrhs, alpha, ldf, f, beta, ldv, v);
    real_.spmv(nrhs, alpha, ldf, f, 1.0, ldv, v);
}

void MobSpme::real_mul_vector(int nrhs, double alpha, int ldf, const double* f,
                              double beta, int ldv, double* v) {
    real_.spmv(nrhs, alpha, ldf, f, beta, ldv, v);
}

void MobSpme::recip_mul_vector(int nrhs, double alpha, int ldf, const double* f,
                               double beta, int ldv, double* v) {
    spme_.recip(nrhs, alpha, ldf, f, beta, ldv, v);
}

}  // namespace orc
