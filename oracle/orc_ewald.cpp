// Oracle dense Ewald-summed RPY matrix (test infrastructure only; see
// orc_common.hpp). Restates src/mob/ewald.cc:19-83 (EwaldReal/EwaldRecip,
// XitoRmax/XitoKmax), :86-330 (EwaldKernel) and :643-797 (tables), and the
// automatic splitting of src/mob/mob_ewald.cc:39-53.
#include <cmath>
#include <vector>
#include <omp.h>
#include "orc_common.hpp"
#include "orc_spme.hpp"

namespace orc {

static double ewald_real_bound(double r, double xi) {
    double xir = xi * r, xir2 = xir * xir, r2 = r * r;
    return erfc(xir) * (0.75 + 1.5 / r2) / r +
           exp(-xir2) * xi / sqrt(M_PI) *
               (4.5 + 3.0 * xir2 + (3.0 + xir2 * (14.0 + xir2 * (4.0 + xir2))) / r2);
}

static double ewald_recip_bound(double k, double xi) {
    double kx = k / xi, kx2 = kx * kx, k2 = k * k;
    return 6.0 * M_PI * exp(-kx2 / 4.0) * (1.0 + k2 / 3.0) / k2 *
           (1.0 + kx2 * (1.0 / 4.0 + kx2 / 8.0));
}

static double xi_to_rmax(double xi, double tol, double box, double safe) {
    const double red = 0.99;
    double xi0 = xi / safe;
    double rmax = 10.0 * sqrt(-log(tol)) / xi;
    while (true) {
        double e0 = ewald_real_bound(rmax * red, xi0);
        double e1 = 6.0 * ewald_real_bound(rmax * red + box, xi0);
        if (e0 + e1 >= tol) break;
        rmax *= red;
    }
    return rmax;
}

static double xi_to_kmax(double xi, double tol, double box, double safe) {
    const double red = 0.99;
    double dk = 2.0 * M_PI / box;
    double xi0 = xi * safe;
    double kmax = 10.0 * sqrt(-log(tol)) * 2.0 * xi;
    while (true) {
        double e0 = ewald_recip_bound(kmax * red, xi0);
        double e1 = 6.0 * ewald_recip_bound(kmax * red + dk, xi0);
        if (e0 + e1 >= tol) break;
        kmax *= red;
    }
    return kmax;
}

double ewald_auto_xi(double box) {
    return pow(10.0, 1.0 / 6.0) * sqrt(M_PI) / box;
}

// Dense 3N x 3N row-major matrices (any of mat/mat_real/mat_recip may be
// null); ldm is the row stride.
void ewald_matrix(int npos, const double* pos, const double* rdi, double box,
                  double xi, double tol, int ldm, double* mat, double* mat_real,
                  double* mat_recip, int* nr_out, int* nk_out) {
    // real-space image table
    double rmax = xi_to_rmax(xi, tol, box, 1.15);
    std::vector<double> rlx, rly, rlz;
    int rmaxi = (int)(rmax / box) + 1;
    double rl = sqrt(3 * box * box);
    for (int m1 = -rmaxi; m1 <= rmaxi; m1++) {
        double xx = box * m1 + (m1 > 0 ? -0.5 * box : (m1 < 0 ? 0.5 * box : 0.0));
        for (int m2 = -rmaxi; m2 <= rmaxi; m2++) {
            double yy = box * m2 + (m2 > 0 ? -0.5 * box : (m2 < 0 ? 0.5 * box : 0.0));
            for (int m3 = -rmaxi; m3 <= rmaxi; m3++) {
                double zz = box * m3 + (m3 > 0 ? -0.5 * box : (m3 < 0 ? 0.5 * box : 0.0));
                if (sqrt(xx * xx + yy * yy + zz * zz) <= rmax + rl) {
                    rlx.push_back(box * m1);
                    rly.push_back(box * m2);
                    rlz.push_back(box * m3);
                }
            }
        }
    }
    const int nr = (int)rlx.size();
    // reciprocal-space table
    double kmax = xi_to_kmax(xi, tol, box, 1.15);
    int kmaxi = (int)(kmax * box / (2.0 * M_PI));
    double pivol = M_PI / box / box / box;
    double xi2 = xi * xi;
    std::vector<double> kv1, kv2, kv3, kabs, kexp;
    for (int m1 = -kmaxi; m1 <= kmaxi; m1++) {
        double k1 = 2.0 * M_PI * m1 / box;
        for (int m2 = -kmaxi; m2 <= kmaxi; m2++) {
            double k2 = 2.0 * M_PI * m2 / box;
            for (int m3 = -kmaxi; m3 <= kmaxi; m3++) {
                double k3 = 2.0 * M_PI * m3 / box;
                if (m1 == 0 && m2 == 0 && m3 == 0) continue;
                double kk = k1 * k1 + k2 * k2 + k3 * k3;
                double k = sqrt(kk);
                if (k > kmax) continue;
                double k4z = kk / (4.0 * xi2);
                double k4zp = 1.0 + k4z * (1.0 + 2.0 * k4z);
                kv1.push_back(k1); kv2.push_back(k2); kv3.push_back(k3);
                kabs.push_back(k);
                kexp.push_back(pivol * k4zp / kk * exp(-k4z));
            }
        }
    }
    const int nk = (int)kabs.size();
    if (nr_out) *nr_out = nr;
    if (nk_out) *nk_out = nk;
    const double rmax2 = rmax * rmax;
    const double xiaspi = xi / sqrt(M_PI);

    #pragma omp parallel for schedule(dynamic, 4)
    for (int i = 0; i < npos; i++) {
        const double* p1 = &pos[3 * i];
        double aa = rdi[i], aa2 = aa * aa;
        for (int j = i; j < npos; j++) {
            const double* p2 = &pos[3 * j];
            double ab = rdi[j], ab2 = ab * ab;
            double x0 = remainder(p2[0] - p1[0], box);
            double y0 = remainder(p2[1] - p1[1], box);
            double z0 = remainder(p2[2] - p1[2], box);
            double re[6] = {0, 0, 0, 0, 0, 0};
            if (mat || mat_real) {
                if (i == j) {
                    double self_a = 1.0 / aa - xiaspi * (6.0 - 40.0 / 3.0 * xi2 * aa2);
                    re[0] = re[3] = re[5] = self_a;
                }
                for (int m = 0; m < nr; m++) {
                    double xx = x0 + rlx[m], yy = y0 + rly[m], zz = z0 + rlz[m];
                    double rr = xx * xx + yy * yy + zz * zz;
                    if (rr == 0.0 || rr > rmax2) continue;
                    double r = sqrt(rr);
                    double ex = xx / r, ey = yy / r, ez = zz / r;
                    double xa, ya;
                    if (r < aa + ab) {
                        // overlap correction (src/mob/ewald.cc:177-195)
                        double cxa = -(1.5 - (aa2 + ab2) * 0.5 / rr) / r + 2.0 / (aa + ab) -
                                     3.0 * r / 8.0 / (aa2 + ab2);
                        double cya = -(0.75 + (aa2 + ab2) * 0.25 / rr) / r + 2.0 / (aa + ab) -
                                     9.0 * r / 16.0 / (aa2 + ab2);
                        re[0] += (cxa - cya) * ex * ex + cya;
                        re[1] += (cxa - cya) * ex * ey;
                        re[2] += (cxa - cya) * ex * ez;
                        re[3] += (cxa - cya) * ey * ey + cya;
                        re[4] += (cxa - cya) * ey * ez;
                        re[5] += (cxa - cya) * ez * ez + cya;
                    }
                    scalars_ewald(xi, r, aa, ab, &xa, &ya);
                    re[0] += (xa - ya) * ex * ex + ya;
                    re[1] += (xa - ya) * ex * ey;
                    re[2] += (xa - ya) * ex * ez;
                    re[3] += (xa - ya) * ey * ey + ya;
                    re[4] += (xa - ya) * ey * ez;
                    re[5] += (xa - ya) * ez * ez + ya;
                }
            }
            double rc[6] = {0, 0, 0, 0, 0, 0};
            if (mat || mat_recip) {
                for (int m = 0; m < nk; m++) {
                    double k = kabs[m];
                    double ex = kv1[m] / k, ey = kv2[m] / k, ez = kv3[m] / k;
                    double kk = k * k;
                    double ya = 6.0 * (1.0 - kk * (aa2 + ab2) / 6.0) * kexp[m];
                    double cf = cos(kv1[m] * x0 + kv2[m] * y0 + kv3[m] * z0);
                    rc[0] += cf * ya * (1.0 - ex * ex);
                    rc[1] += -cf * ya * ex * ey;
                    rc[2] += -cf * ya * ex * ez;
                    rc[3] += cf * ya * (1.0 - ey * ey);
                    rc[4] += -cf * ya * ey * ez;
                    rc[5] += cf * ya * (1.0 - ez * ez);
                }
            }
            auto store = [&](double* M, const double* s) {
                if (!M) return;
                const int i0 = 3 * i, j0 = 3 * j;
                double b[3][3] = {{s[0], s[1], s[2]}, {s[1], s[3], s[4]}, {s[2], s[4], s[5]}};
                for (int a = 0; a < 3; a++)
                    for (int c = 0; c < 3; c++) {
                        M[(size_t)(i0 + a) * ldm + j0 + c] = b[a][c];
                        M[(size_t)(j0 + a) * ldm + i0 + c] = b[a][c];
                    }
            };
            store(mat_real, re);
            store(mat_recip, rc);
            if (mat) {
                double s[6];
                for (int t = 0; t < 6; t++) s[t] = re[t] + rc[t];
                store(mat, s);
            }
        }
    }
}

}  // namespace orc
