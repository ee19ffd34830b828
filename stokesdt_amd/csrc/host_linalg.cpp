// Small dense FP64 host linear algebra for the Krylov solver.
//
// The reference obtains T^{1/2} E_1 from LAPACKE_dsbev with a full
// eigenvector matrix every iteration (src/brwn/brwn_lanczos.cc:163-177),
// O(m^3). Only m x b entries of the result are needed, so this form keeps
// the rotations of the band reduction and of the QL iteration and applies
// them to m x b matrices.
#include "host_linalg.hpp"
#include <algorithm>
#include <cmath>
#include <cstring>

namespace sdt {

bool chol_upper(int b, const double* G, double* R) {
    for (int i = 0; i < b * b; i++) R[i] = 0.0;
    for (int j = 0; j < b; j++) {
        double s = G[j * b + j];
        for (int k = 0; k < j; k++) s -= R[k * b + j] * R[k * b + j];
        if (!(s > 0.0)) return false;
        const double rjj = sqrt(s);
        R[j * b + j] = rjj;
        for (int c = j + 1; c < b; c++) {
            double t = G[j * b + c];
            for (int k = 0; k < j; k++) t -= R[k * b + j] * R[k * b + c];
            R[j * b + c] = t / rjj;
        }
    }
    return true;
}

void inv_upper(int b, const double* R, double* Ri) {
    for (int i = 0; i < b * b; i++) Ri[i] = 0.0;
    for (int c = 0; c < b; c++) {
        // solve R x = e_c (back substitution), column c of R^{-1}
        for (int r = c; r >= 0; r--) {
            double s = (r == c) ? 1.0 : 0.0;
            for (int k = r + 1; k <= c; k++) s -= R[r * b + k] * Ri[k * b + c];
            Ri[r * b + c] = s / R[r * b + r];
        }
    }
}

void mul_upper(int b, const double* A, const double* B, double* C) {
    for (int i = 0; i < b; i++)
        for (int j = 0; j < b; j++) {
            double s = 0.0;
            for (int k = i; k <= j; k++) s += A[i * b + k] * B[k * b + j];
            C[i * b + j] = (j >= i) ? s : 0.0;
        }
}

void BandSqrt::compute(int m, int kd, const double* T, int kb, double* S) {
    const int ldt = 2 * kd + 1;
    std::vector<double> Tb((size_t)m * ldt, 0.0);
    for (int i = 0; i < m; i++)
        for (int j = std::max(0, i - kd); j <= std::min(m - 1, i + kd); j++)
            Tb[(size_t)i * ldt + (j - i + kd)] = T[(size_t)i * m + j];
    compute_band(m, kd, Tb.data(), ldt, kb, S);
}

void BandSqrt::compute_band(int m, int kd, const double* Tb, int ldt, int kb, double* S) {
    // band storage of half-width w (the bandwidth kd plus room for the
    // bulge chased during the reduction): element (i, j), |i - j| <= w, at
    // A[i * ld + j - i + w]
    const int w = kd + 2;
    const int ld = 2 * w + 1;
    A_.assign((size_t)m * ld, 0.0);
    double* A = A_.data();
    auto at = [&](int i, int j) -> double& { return A[(size_t)i * ld + (j - i + w)]; };
    for (int i = 0; i < m; i++)
        for (int j = std::max(0, i - kd); j <= std::min(m - 1, i + kd); j++)
            at(i, j) = Tb[(size_t)i * ldt + (j - i + kd)];
    band_rot_.clear();
    ql_rot_.clear();
    auto rotate = [&](int p, double c, double s) {
        // A <- G A G^T with G = [[c, s], [-s, c]] on rows/cols (p, p+1)
        const int q = p + 1;
        const int lo = std::max(0, q - w), hi = std::min(m - 1, p + w);
        // rows p and q (columns lo..hi lie within both bands)
        double* rp = &at(p, lo);
        double* rq = &at(q, lo);
        for (int l = 0; l <= hi - lo; l++) {
            double ap = rp[l], aq = rq[l];
            rp[l] = c * ap + s * aq;
            rq[l] = -s * ap + c * aq;
        }
        // columns p and q, by symmetry, outside the 2x2 block
        for (int l = lo; l <= hi; l++) {
            if (l == p || l == q) continue;
            at(l, p) = at(p, l);
            at(l, q) = at(q, l);
        }
        // the 2x2 block: rows were rotated, now the columns
        double app = at(p, p), apq = at(p, q), aqp = at(q, p), aqq = at(q, q);
        at(p, p) = c * app + s * apq;
        at(p, q) = -s * app + c * apq;
        at(q, p) = c * aqp + s * aqq;
        at(q, q) = -s * aqp + c * aqq;
        band_rot_.push_back(Rot{p, c, s});
    };
    auto zero_with = [&](int p, int col) {
        // zero A[p+1][col] against A[p][col]
        const double a = at(p, col), b = at(p + 1, col);
        if (b == 0.0) return;
        const double r = sqrt(a * a + b * b);
        rotate(p, a / r, b / r);
    };
    // band -> tridiagonal (bulge chasing)
    if (kd > 1) {
        for (int j = 0; j + 2 < m; j++) {
            for (int k = std::min(kd, m - 1 - j); k >= 2; k--) {
                int p = j + k - 1;
                zero_with(p, j);
                int col = p, i = p + 1 + kd;
                while (i < m) {
                    zero_with(i - 1, col);
                    col = i - 1;
                    i += kd;
                }
            }
        }
    }
    d_.resize(m);
    e_.assign(m, 0.0);
    for (int i = 0; i < m; i++) d_[i] = at(i, i);
    for (int i = 0; i + 1 < m; i++) e_[i] = at(i + 1, i);
    // implicit QL with Wilkinson-type shifts, rotations recorded
    {
        double* d = d_.data();
        double* e = e_.data();
        const double eps = 2.220446049250313e-16;
        for (int l = 0; l < m; l++) {
            int iter = 0, mm;
            do {
                for (mm = l; mm < m - 1; mm++) {
                    double dd = fabs(d[mm]) + fabs(d[mm + 1]);
                    if (fabs(e[mm]) <= eps * dd) break;
                }
                if (mm != l) {
                    if (++iter > 60) break;
                    double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
                    double r = sqrt(g * g + 1.0);
                    g = d[mm] - d[l] + e[l] / (g + (g >= 0.0 ? r : -r));
                    double s = 1.0, c = 1.0, p = 0.0;
                    bool early = false;
                    for (int i = mm - 1; i >= l; i--) {
                        double f = s * e[i], bb = c * e[i];
                        r = sqrt(f * f + g * g);
                        e[i + 1] = r;
                        if (r == 0.0) {
                            d[i + 1] -= p;
                            e[mm] = 0.0;
                            early = true;
                            break;
                        }
                        s = f / r;
                        c = g / r;
                        g = d[i + 1] - p;
                        r = (d[i] - g) * s + 2.0 * c * bb;
                        p = s * r;
                        d[i + 1] = g + p;
                        g = c * r - bb;
                        ql_rot_.push_back(Rot{i, c, s});
                    }
                    if (early) continue;
                    d[l] -= p;
                    e[l] = g;
                    e[mm] = 0.0;
                }
            } while (mm != l);
        }
    }
    // X = E_1 (m x kb, row-major)
    X_.assign((size_t)m * kb, 0.0);
    double* X = X_.data();
    for (int c = 0; c < kb && c < m; c++) X[(size_t)c * kb + c] = 1.0;
    // X <- Q_band^T X : apply each band rotation G (rows p, p+1) in order
    for (const Rot& r : band_rot_) {
        double* xp = X + (size_t)r.p * kb;
        double* xq = xp + kb;
        for (int c = 0; c < kb; c++) {
            double a = xp[c], b = xq[c];
            xp[c] = r.c * a + r.s * b;
            xq[c] = -r.s * a + r.c * b;
        }
    }
    // X <- Z^T X : QL rotation R^T = [[c, -s], [s, c]] on rows (i, i+1)
    for (const Rot& r : ql_rot_) {
        double* xi = X + (size_t)r.p * kb;
        double* xj = xi + kb;
        for (int c = 0; c < kb; c++) {
            double a = xi[c], b = xj[c];
            xi[c] = r.c * a - r.s * b;
            xj[c] = r.s * a + r.c * b;
        }
    }
    // g(Lambda) = sqrt
    for (int i = 0; i < m; i++) {
        double g = sqrt(std::max(d_[i], 0.0));
        for (int c = 0; c < kb; c++) X[(size_t)i * kb + c] *= g;
    }
    // X <- Z X : replay QL rotations in reverse applying R
    for (size_t t = ql_rot_.size(); t-- > 0;) {
        const Rot& r = ql_rot_[t];
        double* xi = X + (size_t)r.p * kb;
        double* xj = xi + kb;
        for (int c = 0; c < kb; c++) {
            double a = xi[c], b = xj[c];
            xi[c] = r.c * a + r.s * b;
            xj[c] = -r.s * a + r.c * b;
        }
    }
    // X <- Q_band X : replay band rotations in reverse applying G^T
    for (size_t t = band_rot_.size(); t-- > 0;) {
        const Rot& r = band_rot_[t];
        double* xp = X + (size_t)r.p * kb;
        double* xq = xp + kb;
        for (int c = 0; c < kb; c++) {
            double a = xp[c], b = xq[c];
            xp[c] = r.c * a - r.s * b;
            xq[c] = r.s * a + r.c * b;
        }
    }
    memcpy(S, X, sizeof(double) * (size_t)m * kb);
}

}  // namespace sdt
