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

namespace {

typedef BandSqrt::Rot Rot;

// rows (p, p+1) of the m x KB row-major X by [[c, sg s], [-sg s, c]] for each
// recorded rotation, in order or in reverse (sg = +-1: the products by -s
// equal those of the written-out forms, sign changes being exact)
template <int KB>
void replay_fixed(const Rot* r, size_t n, bool reverse, double sg, double* X, int kb) {
    const int nk = KB ? KB : kb;   // KB = 0: the width at run time
    for (size_t t = 0; t < n; t++) {
        const Rot& g = r[reverse ? n - 1 - t : t];
        const double c = g.c, s = sg * g.s;
        double* xp = X + (size_t)g.p * nk;
        double* xq = xp + nk;
        for (int k = 0; k < nk; k++) {
            const double a = xp[k], b = xq[k];
            xp[k] = c * a + s * b;
            xq[k] = -s * a + c * b;
        }
    }
}

void replay(const std::vector<Rot>& r, bool reverse, double sg, double* X, int kb) {
    switch (kb) {
        case 4: replay_fixed<4>(r.data(), r.size(), reverse, sg, X, kb); break;
        case 8: replay_fixed<8>(r.data(), r.size(), reverse, sg, X, kb); break;
        case 16: replay_fixed<16>(r.data(), r.size(), reverse, sg, X, kb); break;
        default: replay_fixed<0>(r.data(), r.size(), reverse, sg, X, kb); break;
    }
}

// implicit QL with Wilkinson-type shifts on the tridiagonal (d, e), the
// rotations recorded. Each rotation R^T = [[c, -s], [s, c]] is also applied
// to rows (i, i+1) of X as it is made (X <- Z^T X): the row updates do not
// feed the sweep's chain of square roots and divisions and run beside it.
template <int KB>
void ql_sweeps(int m, double* d, double* e, std::vector<Rot>& rot, double* X, int kb) {
    const int nk = KB ? KB : kb;
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
                    rot.push_back(Rot{i, c, s});
                    double* xi = X + (size_t)i * nk;
                    double* xj = xi + nk;
                    for (int k = 0; k < nk; k++) {
                        const double xa = xi[k], xb = xj[k];
                        xi[k] = c * xa - s * xb;
                        xj[k] = s * xa + c * xb;
                    }
                }
                if (early) continue;
                d[l] -= p;
                e[l] = g;
                e[mm] = 0.0;
            }
        } while (mm != l);
    }
}

}  // namespace

void BandSqrt::compute_band(int m, int kd, const double* Tb, int ldt, int kb, double* S) {
    // lower band storage of half-width w (the bandwidth kd plus room for the
    // bulge chased during the reduction): A(i, j), 0 <= i - j <= w, at
    // A[i * ld + i - j]; the rotations keep A symmetric, so the upper half is
    // never formed
    const int w = kd + 2;
    const int ld = w + 1;
    A_.assign((size_t)m * ld, 0.0);
    double* A = A_.data();
    auto at = [&](int i, int j) -> double& { return A[(size_t)i * ld + (i - j)]; };
    for (int i = 0; i < m; i++)
        for (int j = std::max(0, i - kd); j <= i; j++) at(i, j) = Tb[(size_t)i * ldt + (j - i + kd)];
    band_rot_.clear();
    ql_rot_.clear();
    auto rotate = [&](int p, double c, double s) {
        // A <- G A G^T with G = [[c, s], [-s, c]] on rows/cols (p, p+1)
        const int q = p + 1;
        double* rp = A + (size_t)p * ld;   // rp[p - l] = A(p, l)
        double* rq = A + (size_t)q * ld;   // rq[q - l] = A(q, l)
        // rows p and q left of the 2x2 block
        for (int l = std::max(0, q - w); l < p; l++) {
            const double ap = rp[p - l], aq = rq[q - l];
            rp[p - l] = c * ap + s * aq;
            rq[q - l] = -s * ap + c * aq;
        }
        // columns p and q below it (rows l, entries A(l, p) and A(l, q))
        const int l1 = std::min(m - 1, p + w);
        for (int l = q + 1; l <= l1; l++) {
            double* rl = A + (size_t)l * ld;
            const double ap = rl[l - p], aq = rl[l - q];
            rl[l - p] = c * ap + s * aq;
            rl[l - q] = -s * ap + c * aq;
        }
        // the 2x2 block: its rows, then its columns
        const double app = rp[0], apq = rq[1], aqq = rq[0];
        const double x1 = c * app + s * apq, x2 = c * apq + s * aqq;
        const double x3 = -s * app + c * apq, x4 = -s * apq + c * aqq;
        rp[0] = c * x1 + s * x2;
        rq[1] = c * x3 + s * x4;
        rq[0] = -s * x3 + c * x4;
        band_rot_.push_back(Rot{p, c, s});
    };
    auto zero_with = [&](int p, int col) {
        // zero A(p+1, col) against A(p, col), col < p
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
    // X = E_1 (m x kb, row-major)
    X_.assign((size_t)m * kb, 0.0);
    double* X = X_.data();
    for (int c = 0; c < kb && c < m; c++) X[(size_t)c * kb + c] = 1.0;
    // X <- Q_band^T X : each band rotation G (rows p, p+1) in order
    replay(band_rot_, false, 1.0, X, kb);
    // X <- Z^T X inside the QL sweeps
    switch (kb) {
        case 4: ql_sweeps<4>(m, d_.data(), e_.data(), ql_rot_, X, kb); break;
        case 8: ql_sweeps<8>(m, d_.data(), e_.data(), ql_rot_, X, kb); break;
        case 16: ql_sweeps<16>(m, d_.data(), e_.data(), ql_rot_, X, kb); break;
        default: ql_sweeps<0>(m, d_.data(), e_.data(), ql_rot_, X, kb); break;
    }
    // g(Lambda) = sqrt
    for (int i = 0; i < m; i++) {
        double g = sqrt(std::max(d_[i], 0.0));
        for (int c = 0; c < kb; c++) X[(size_t)i * kb + c] *= g;
    }
    // X <- Z X : the QL rotations in reverse, applying R
    replay(ql_rot_, true, 1.0, X, kb);
    // X <- Q_band X : the band rotations in reverse, applying G^T
    replay(band_rot_, true, -1.0, X, kb);
    memcpy(S, X, sizeof(double) * (size_t)m * kb);
}

}  // namespace sdt
