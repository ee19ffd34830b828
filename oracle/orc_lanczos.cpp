// Oracle Krylov Brownian displacements (test infrastructure only; see
// orc_common.hpp). Restates BrwnLanczos (src/brwn/brwn_lanczos.cc:16-304)
// and BrwnChol (src/brwn/brwn_chol.cc:52-83).
#include <cmath>
#include <cstring>
#include <vector>
#include "orc_common.hpp"
#include "orc_lanczos.hpp"

namespace orc {

static double nrm2(int n, const double* x) {
    double s = 0.0;
    for (int i = 0; i < n; i++) s += x[i] * x[i];
    return sqrt(s);
}

BrwnLanczos::BrwnLanczos(int dim, int max_iters, int max_nrhs, double tol)
    : dim_(dim), max_iters_(max_iters), max_nrhs_(max_nrhs), tol_(tol) {}

// src/brwn/brwn_lanczos.cc:224-275
bool BrwnLanczos::init() {
    if (dim_ <= 0 || max_iters_ <= 0 || max_nrhs_ <= 0 || tol_ <= 0.0) return false;
    if (max_nrhs_ > dim_ / 3) max_nrhs_ = dim_ / 3;
    ldv_ = (int)pad_len(dim_, sizeof(double));
    v_.assign((size_t)max_nrhs_ * (max_iters_ + 1) * ldv_, 0.0);
    y_old_.assign((size_t)max_nrhs_ * ldv_, 0.0);
    return true;
}

// single-vector Lanczos, src/brwn/brwn_lanczos.cc:16-82
int BrwnLanczos::lanczos(Mob* mob, const double* z, double* y) {
    const int n = dim_;
    std::vector<double> d0(max_iters_ + 1), e0(max_iters_ + 1);
    double normz = nrm2(n, z);
    for (int i = 0; i < n; i++) v_[i] = z[i] / normz;
    int iter;
    for (iter = 0; iter < max_iters_; iter++) {
        double* w = &v_[(size_t)(iter + 1) * ldv_];
        const double* vi = &v_[(size_t)iter * ldv_];
        mob->mul_vector(1, 1.0, ldv_, vi, 0.0, ldv_, w);
        if (iter > 0) {
            const double* vp = &v_[(size_t)(iter - 1) * ldv_];
            for (int i = 0; i < n; i++) w[i] -= e0[iter - 1] * vp[i];
        }
        double h = 0.0;
        for (int i = 0; i < n; i++) h += w[i] * vi[i];
        d0[iter] = h;
        for (int i = 0; i < n; i++) w[i] -= h * vi[i];
        double normw = nrm2(n, w);
        e0[iter] = normw;
        for (int i = 0; i < n; i++) w[i] /= normw;
        // sqrt of the (iter+1)-tridiagonal, first column
        int m = iter + 1;
        std::vector<double> w_eig(m), Q((size_t)m * m);
        tridiag_eig(m, d0.data(), e0.data(), w_eig.data(), Q.data(), m);
        std::vector<double> s(m, 0.0);
        for (int a = 0; a < m; a++) {
            double acc = 0.0;
            for (int k = 0; k < m; k++) acc += Q[(size_t)a * m + k] * sqrt(w_eig[k]) * Q[k];
            s[a] = acc;   // (Q Lambda^1/2 Q^T)[a][0]
        }
        for (int i = 0; i < n; i++) {
            double acc = 0.0;
            for (int a = 0; a < m; a++) acc += v_[(size_t)a * ldv_ + i] * s[a];
            y[i] = normz * acc;
        }
        if (iter > 0) {
            double dn = 0.0;
            for (int i = 0; i < n; i++) {
                double d = y_old_[i] - y[i];
                dn += d * d;
            }
            if (sqrt(dn) / normz < tol_) break;
        }
        for (int i = 0; i < n; i++) y_old_[i] = y[i];
    }
    last_iters_ = iter;
    return iter == max_iters_ ? -1 : iter;
}

// block Lanczos, src/brwn/brwn_lanczos.cc:85-201. T is the block
// tridiagonal with diagonal blocks H_i = V_i^T (M V_i - V_{i-1} B_{i-1}^T)
// and subdiagonal blocks B_i from QR(W) = V_{i+1} B_i; after each iteration
// Y = V_{0..i} T^{1/2} E_1 R with Z = V_0 R. The reference writes Y with its
// internal leading dimension regardless of ldy (a defect); here ldy is
// honoured, which is identical whenever ldy equals that leading dimension.
int BrwnLanczos::block_lanczos(Mob* mob, int b, int ldz, const double* z,
                               int ldy, double* y) {
    const int n = dim_;
    const int ldv = ldv_;
    std::vector<double> R((size_t)b * b), H((size_t)b * b), Bsub((size_t)b * b);
    // Z Frobenius norm and V_0 R = QR(Z)
    double normz = 0.0;
    for (int j = 0; j < b; j++)
        for (int i = 0; i < n; i++) normz += z[(size_t)j * ldz + i] * z[(size_t)j * ldz + i];
    normz = sqrt(normz);
    for (int j = 0; j < b; j++)
        memcpy(&v_[(size_t)j * ldv], &z[(size_t)j * ldz], sizeof(double) * n);
    householder_qr(n, b, &v_[0], ldv, R.data(), b);
    const int nmax = (max_iters_ + 1) * b;
    std::vector<double> T((size_t)nmax * nmax, 0.0);   // row-major, full
    std::vector<double> Yold((size_t)b * n, 0.0);
    int iter;
    for (iter = 0; iter < max_iters_; iter++) {
        double* W = &v_[(size_t)(iter + 1) * b * ldv];
        const double* Vi = &v_[(size_t)iter * b * ldv];
        double beta;
        if (iter > 0) {
            // W = V_{i-1} B_{i-1}^T (Bsub holds B_{i-1}, upper triangular)
            const double* Vp = &v_[(size_t)(iter - 1) * b * ldv];
            #pragma omp parallel for schedule(static)
            for (int i = 0; i < n; i++) {
                for (int c = 0; c < b; c++) {
                    double acc = 0.0;
                    for (int k = 0; k < b; k++) acc += Vp[(size_t)k * ldv + i] * Bsub[(size_t)k * b + c];
                    W[(size_t)c * ldv + i] = acc;
                }
            }
            beta = -1.0;
        } else {
            beta = 0.0;
        }
        // W = M V_i + beta W
        mob->mul_vector(b, 1.0, ldv, Vi, beta, ldv, W);
        // H = V_i^T W
        for (int a = 0; a < b; a++) {
            for (int c = 0; c < b; c++) {
                double acc = 0.0;
                #pragma omp parallel for reduction(+ : acc) schedule(static)
                for (int i = 0; i < n; i++) acc += Vi[(size_t)a * ldv + i] * W[(size_t)c * ldv + i];
                H[(size_t)c * b + a] = acc;   // column-major H(a, c)
            }
        }
        // W -= V_i H
        #pragma omp parallel for schedule(static)
        for (int i = 0; i < n; i++) {
            double vi[64];
            for (int a = 0; a < b; a++) vi[a] = Vi[(size_t)a * ldv + i];
            for (int c = 0; c < b; c++) {
                double acc = 0.0;
                for (int a = 0; a < b; a++) acc += vi[a] * H[(size_t)c * b + a];
                W[(size_t)c * ldv + i] -= acc;
            }
        }
        // [V_{i+1}, B_i] = QR(W)
        householder_qr(n, b, W, ldv, Bsub.data(), b);
        // assemble T: diagonal block H_i (lower part used, as dsbev 'L')
        const int o = iter * b;
        for (int c = 0; c < b; c++)
            for (int a = c; a < b; a++) {
                double hv = H[(size_t)c * b + a];   // H(a, c), a >= c
                T[(size_t)(o + a) * nmax + o + c] = hv;
                T[(size_t)(o + c) * nmax + o + a] = hv;
            }
        // subdiagonal block B_i at rows (i+1)b.., cols ib..
        if (iter + 1 < max_iters_ + 1) {
            for (int c = 0; c < b; c++)
                for (int a = 0; a <= c; a++) {
                    double bv = Bsub[(size_t)c * b + a];   // B(a, c), upper
                    T[(size_t)(o + b + a) * nmax + o + c] = bv;
                    T[(size_t)(o + c) * nmax + o + b + a] = bv;
                }
        }
        // S = T_m^{1/2} E_1 with T_m the leading m = (iter+1)b block
        const int m = (iter + 1) * b;
        std::vector<double> w_eig(m), Q((size_t)m * m);
        sym_eig(m, T.data(), nmax, w_eig.data(), Q.data(), m);
        std::vector<double> S((size_t)m * b), D((size_t)m * b);
        for (int a = 0; a < m; a++)
            for (int c = 0; c < b; c++) {
                double acc = 0.0;
                for (int k = 0; k < m; k++)
                    acc += Q[(size_t)a * m + k] * sqrt(w_eig[k]) * Q[(size_t)c * m + k];
                S[(size_t)a * b + c] = acc;
            }
        // D = S R
        for (int a = 0; a < m; a++)
            for (int c = 0; c < b; c++) {
                double acc = 0.0;
                for (int k = 0; k <= c; k++) acc += S[(size_t)a * b + k] * R[(size_t)c * b + k];
                D[(size_t)a * b + c] = acc;
            }
        // Y = V_{0..iter} D
        #pragma omp parallel for schedule(static)
        for (int i = 0; i < n; i++) {
            for (int c = 0; c < b; c++) {
                double acc = 0.0;
                for (int a = 0; a < m; a++) acc += v_[(size_t)a * ldv + i] * D[(size_t)a * b + c];
                y[(size_t)c * ldy + i] = acc;
            }
        }
        if (iter > 0) {
            double dn = 0.0;
            for (int c = 0; c < b; c++)
                for (int i = 0; i < n; i++) {
                    double d = Yold[(size_t)c * n + i] - y[(size_t)c * ldy + i];
                    dn += d * d;
                }
            if (sqrt(dn) / normz < tol_) break;
        }
        for (int c = 0; c < b; c++)
            for (int i = 0; i < n; i++) Yold[(size_t)c * n + i] = y[(size_t)c * ldy + i];
    }
    last_iters_ = iter;
    return iter == max_iters_ ? -1 : iter;
}

// src/brwn/brwn_lanczos.cc:278-304
void BrwnLanczos::compute(Mob* mob, int num_rhs, int ldz, const double* z,
                          int ldy, double* y) {
    iters_.clear();
    for (int irhs = 0; irhs < num_rhs; irhs += max_nrhs_) {
        int nrhs = (irhs + max_nrhs_ > num_rhs) ? num_rhs - irhs : max_nrhs_;
        int it;
        if (nrhs == 1) it = lanczos(mob, &z[(size_t)ldz * irhs], &y[(size_t)ldy * irhs]);
        else it = block_lanczos(mob, nrhs, ldz, &z[(size_t)ldz * irhs], ldy, &y[(size_t)ldy * irhs]);
        iters_.push_back(it);
    }
}

// Dense Cholesky displacements y = L z with M = L L^T
// (src/brwn/brwn_chol.cc:52-83). Returns the dpotrf-style info code.
int chol_displacements(int n, const double* mat, int ldm, int nrhs, int ldz,
                       const double* z, int ldy, double* y) {
    std::vector<double> L((size_t)n * n);
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) L[(size_t)i * n + j] = mat[(size_t)i * ldm + j];
    int info = cholesky_lower(n, L.data(), n);
    if (info != 0) return info;
    #pragma omp parallel for schedule(dynamic, 16)
    for (int i = 0; i < n; i++) {
        for (int r = 0; r < nrhs; r++) {
            double acc = 0.0;
            for (int k = 0; k <= i; k++) acc += L[(size_t)i * n + k] * z[(size_t)r * ldz + k];
            y[(size_t)r * ldy + i] = acc;
        }
    }
    return 0;
}

}  // namespace orc
