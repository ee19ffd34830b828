// Oracle dense FP64 linear algebra (test infrastructure only; see
// orc_common.hpp).
//
// The reference calls MKL LAPACKE for these (dgeqrf/dorgqr at
// src/brwn/brwn_lanczos.cc:107,111,146,161; dsbev :167; dstev :56; dpotrf
// src/brwn/brwn_chol.cc:61; dsyev tests/lanczos/testlanczos.cc:115). MKL is
// not available; these are textbook algorithms (Golub & Van Loan, "Matrix
// Computations", 4th ed.: Householder tridiagonalisation 8.3.1, implicit
// symmetric QL/QR 8.3.3, Householder QR 5.2.1, Cholesky 4.2.1), pinned in
// tests against numpy/scipy.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>
#include "orc_common.hpp"

namespace orc {

// Implicit QL on a symmetric tridiagonal matrix (d diagonal, e[i] couples
// i and i+1, e[n-1] unused) accumulating plane rotations into the columns
// of Z (row-major n x n, ldz).
static void ql_implicit(int n, double* d, double* e, double* Z, int ldz) {
    if (n <= 1) return;
    e[n - 1] = 0.0;
    const double eps = 2.220446049250313e-16;
    for (int l = 0; l < n; l++) {
        int iter = 0;
        int m;
        do {
            for (m = l; m < n - 1; m++) {
                double dd = fabs(d[m]) + fabs(d[m + 1]);
                if (fabs(e[m]) <= eps * dd) break;
            }
            if (m != l) {
                if (++iter > 60) break;
                double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
                double r = hypot(g, 1.0);
                g = d[m] - d[l] + e[l] / (g + (g >= 0.0 ? fabs(r) : -fabs(r)));
                double s = 1.0, c = 1.0, p = 0.0;
                int i;
                bool deflated = false;
                for (i = m - 1; i >= l; i--) {
                    double f = s * e[i];
                    double b = c * e[i];
                    r = hypot(f, g);
                    e[i + 1] = r;
                    if (r == 0.0) {
                        d[i + 1] -= p;
                        e[m] = 0.0;
                        deflated = true;
                        break;
                    }
                    s = f / r;
                    c = g / r;
                    g = d[i + 1] - p;
                    r = (d[i] - g) * s + 2.0 * c * b;
                    p = s * r;
                    d[i + 1] = g + p;
                    g = c * r - b;
                    for (int k = 0; k < n; k++) {
                        double zf = Z[k * ldz + i + 1];
                        Z[k * ldz + i + 1] = s * Z[k * ldz + i] + c * zf;
                        Z[k * ldz + i] = c * Z[k * ldz + i] - s * zf;
                    }
                }
                if (deflated) continue;
                d[l] -= p;
                e[l] = g;
                e[m] = 0.0;
            }
        } while (m != l);
    }
}

static void sort_eig(int n, double* w, double* V, int ldv) {
    std::vector<int> idx(n);
    for (int i = 0; i < n; i++) idx[i] = i;
    std::sort(idx.begin(), idx.end(), [&](int a, int b) { return w[a] < w[b]; });
    std::vector<double> w2(n), V2((size_t)n * n);
    for (int j = 0; j < n; j++) {
        w2[j] = w[idx[j]];
        for (int i = 0; i < n; i++) V2[(size_t)i * n + j] = V[(size_t)i * ldv + idx[j]];
    }
    for (int j = 0; j < n; j++) w[j] = w2[j];
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) V[(size_t)i * ldv + j] = V2[(size_t)i * n + j];
}

void tridiag_eig(int n, const double* d, const double* e, double* w,
                 double* V, int ldv) {
    std::vector<double> dd(d, d + n), ee(n, 0.0);
    for (int i = 0; i + 1 < n; i++) ee[i] = e[i];
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) V[(size_t)i * ldv + j] = (i == j) ? 1.0 : 0.0;
    ql_implicit(n, dd.data(), ee.data(), V, ldv);
    for (int i = 0; i < n; i++) w[i] = dd[i];
    sort_eig(n, w, V, ldv);
}

void sym_eig(int n, const double* A_in, int lda, double* w, double* V, int ldv) {
    // working copy, row-major n x n
    std::vector<double> A((size_t)n * n);
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) A[(size_t)i * n + j] = A_in[(size_t)i * lda + j];
    // Q accumulates the Householder reflectors: A_orig = Q T Q^T
    std::vector<double> Q((size_t)n * n, 0.0);
    for (int i = 0; i < n; i++) Q[(size_t)i * n + i] = 1.0;
    std::vector<double> v(n), p(n), wv(n);
    for (int k = 0; k + 2 < n; k++) {
        // x = A[k+1:n, k]
        double alpha = 0.0;
        for (int i = k + 1; i < n; i++) alpha += A[(size_t)i * n + k] * A[(size_t)i * n + k];
        alpha = sqrt(alpha);
        if (alpha == 0.0) continue;
        double x0 = A[(size_t)(k + 1) * n + k];
        double sgn = (x0 >= 0.0) ? 1.0 : -1.0;
        // v = x + sgn*||x|| e1, beta = 2 / v^T v
        for (int i = 0; i < n; i++) v[i] = 0.0;
        for (int i = k + 1; i < n; i++) v[i] = A[(size_t)i * n + k];
        v[k + 1] += sgn * alpha;
        double vtv = 0.0;
        for (int i = k + 1; i < n; i++) vtv += v[i] * v[i];
        double beta = 2.0 / vtv;
        // p = beta * A v (only trailing block matters, but the full row k
        // also changes: apply to rows/cols k..n-1)
        for (int i = k; i < n; i++) {
            double s = 0.0;
            for (int j = k + 1; j < n; j++) s += A[(size_t)i * n + j] * v[j];
            p[i] = beta * s;
        }
        double ptv = 0.0;
        for (int i = k + 1; i < n; i++) ptv += p[i] * v[i];
        for (int i = k; i < n; i++) wv[i] = p[i] - 0.5 * beta * ptv * v[i];
        wv[k] = p[k];  // v[k] == 0
        // A := A - v w^T - w v^T on rows/cols k..n-1
        for (int i = k; i < n; i++)
            for (int j = k; j < n; j++)
                A[(size_t)i * n + j] -= v[i] * wv[j] + wv[i] * v[j];
        // Q := Q P with P = I - beta v v^T
        for (int i = 0; i < n; i++) {
            double s = 0.0;
            for (int j = k + 1; j < n; j++) s += Q[(size_t)i * n + j] * v[j];
            s *= beta;
            for (int j = k + 1; j < n; j++) Q[(size_t)i * n + j] -= s * v[j];
        }
    }
    std::vector<double> d(n), e(n, 0.0);
    for (int i = 0; i < n; i++) d[i] = A[(size_t)i * n + i];
    for (int i = 0; i + 1 < n; i++) e[i] = A[(size_t)(i + 1) * n + i];
    ql_implicit(n, d.data(), e.data(), Q.data(), n);
    for (int i = 0; i < n; i++) w[i] = d[i];
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) V[(size_t)i * ldv + j] = Q[(size_t)i * n + j];
    sort_eig(n, w, V, ldv);
}

void householder_qr(int m, int n, double* A, int lda, double* R, int ldr) {
    std::vector<double> vs((size_t)m * n, 0.0), betas(n, 0.0);
    for (int j = 0; j < n; j++) {
        double alpha = 0.0;
        for (int i = j; i < m; i++) alpha += A[(size_t)j * lda + i] * A[(size_t)j * lda + i];
        alpha = sqrt(alpha);
        double* v = &vs[(size_t)j * m];
        if (alpha == 0.0) { betas[j] = 0.0; continue; }
        double x0 = A[(size_t)j * lda + j];
        double sgn = (x0 >= 0.0) ? 1.0 : -1.0;
        for (int i = j; i < m; i++) v[i] = A[(size_t)j * lda + i];
        v[j] += sgn * alpha;
        double vtv = 0.0;
        for (int i = j; i < m; i++) vtv += v[i] * v[i];
        double beta = 2.0 / vtv;
        betas[j] = beta;
        for (int c = j; c < n; c++) {
            double s = 0.0;
            for (int i = j; i < m; i++) s += v[i] * A[(size_t)c * lda + i];
            s *= beta;
            for (int i = j; i < m; i++) A[(size_t)c * lda + i] -= s * v[i];
        }
    }
    for (int c = 0; c < n; c++)
        for (int r = 0; r < n; r++)
            R[(size_t)c * ldr + r] = (r <= c) ? A[(size_t)c * lda + r] : 0.0;
    // explicit Q = H_0 H_1 ... H_{n-1} [I_n; 0]
    for (int c = 0; c < n; c++)
        for (int i = 0; i < m; i++) A[(size_t)c * lda + i] = (i == c) ? 1.0 : 0.0;
    for (int j = n - 1; j >= 0; j--) {
        double beta = betas[j];
        if (beta == 0.0) continue;
        const double* v = &vs[(size_t)j * m];
        for (int c = 0; c < n; c++) {
            double s = 0.0;
            for (int i = j; i < m; i++) s += v[i] * A[(size_t)c * lda + i];
            s *= beta;
            for (int i = j; i < m; i++) A[(size_t)c * lda + i] -= s * v[i];
        }
    }
}

int cholesky_lower(int n, double* A, int lda) {
    for (int j = 0; j < n; j++) {
        double s = A[(size_t)j * lda + j];
        for (int k = 0; k < j; k++) s -= A[(size_t)j * lda + k] * A[(size_t)j * lda + k];
        if (!(s > 0.0)) return j + 1;
        double ljj = sqrt(s);
        A[(size_t)j * lda + j] = ljj;
        #pragma omp parallel for schedule(static) if (n - j > 256)
        for (int i = j + 1; i < n; i++) {
            double t = A[(size_t)i * lda + j];
            for (int k = 0; k < j; k++) t -= A[(size_t)i * lda + k] * A[(size_t)j * lda + k];
            A[(size_t)i * lda + j] = t / ljj;
        }
        for (int k = j + 1; k < n; k++) A[(size_t)j * lda + k] = 0.0;
    }
    return 0;
}

double frob_rel_err(int nrows, int ncols, int ld1, const double* a,
                    int ld2, const double* b) {
    // relative Frobenius error ||a - b||_F / ||a||_F over nrows vectors of
    // length ncols (the reference metric, src/util/matrix_io.cc:91-117)
    double num = 0.0, den = 0.0;
    for (int r = 0; r < nrows; r++) {
        for (int c = 0; c < ncols; c++) {
            double x = a[(size_t)r * ld1 + c];
            double d = x - b[(size_t)r * ld2 + c];
            num += d * d;
            den += x * x;
        }
    }
    return sqrt(num) / sqrt(den);
}

}  // namespace orc
