// Oracle FFT (test infrastructure only; see orc_common.hpp).
//
// The reference calls Intel MKL DFTI for its 3D R2C/C2R transforms
// (src/mob/spme.cc:549-618 set-up, :754-765 use). MKL is not available, so
// this is a from-the-mathematics mixed-radix Stockham (self-sorting,
// decimation in frequency) FFT. It is pinned in tests against numpy.fft and
// against a brute-force DFT.
#include <cmath>
#include <cstring>
#include <omp.h>
#include "orc_common.hpp"

namespace orc {

void FFTPlan1D::init(int n_) {
    n = n_;
    radix.clear();
    int m = n;
    // prefer radix 4, then 2, 3, 5, then any remaining prime factor
    while (m % 4 == 0) { radix.push_back(4); m /= 4; }
    while (m % 2 == 0) { radix.push_back(2); m /= 2; }
    while (m % 3 == 0) { radix.push_back(3); m /= 3; }
    while (m % 5 == 0) { radix.push_back(5); m /= 5; }
    for (int f = 7; m > 1; f += 2) {
        while (m % f == 0) { radix.push_back(f); m /= f; }
    }
    tw.clear();
    tw_off.clear();
    int len = n;
    for (size_t s = 0; s < radix.size(); s++) {
        int r = radix[s];
        int mm = len / r;
        tw_off.push_back((int)tw.size());
        for (int p = 0; p < mm; p++) {
            for (int t = 0; t < r; t++) {
                double ang = -2.0 * M_PI * (double)((long long)p * t % len) / len;
                tw.push_back(cplx(cos(ang), sin(ang)));
            }
        }
        len = mm;
    }
}

// one DFT of radix r on a[0..r-1] (in place), sign -1 forward
static inline void small_dft(int r, cplx* a, int sign) {
    if (r == 2) {
        cplx u = a[0], v = a[1];
        a[0] = u + v;
        a[1] = u - v;
    } else if (r == 4) {
        cplx a0 = a[0], a1 = a[1], a2 = a[2], a3 = a[3];
        cplx s02 = a0 + a2, d02 = a0 - a2;
        cplx s13 = a1 + a3, d13 = a1 - a3;
        // multiply d13 by -i (forward) or +i (backward)
        cplx jd13 = (sign < 0) ? cplx(d13.imag(), -d13.real())
                               : cplx(-d13.imag(), d13.real());
        a[0] = s02 + s13;
        a[1] = d02 + jd13;
        a[2] = s02 - s13;
        a[3] = d02 - jd13;
    } else {
        cplx out[64];
        cplx* o = out;
        std::vector<cplx> big;
        if (r > 64) { big.resize(r); o = big.data(); }
        for (int t = 0; t < r; t++) {
            cplx acc(0.0, 0.0);
            for (int j = 0; j < r; j++) {
                double ang = sign * 2.0 * M_PI * (double)((j * t) % r) / r;
                acc += a[j] * cplx(cos(ang), sin(ang));
            }
            o[t] = acc;
        }
        for (int t = 0; t < r; t++) a[t] = o[t];
    }
}

void FFTPlan1D::exec(cplx* x, cplx* work, int sign) const {
    cplx* src = x;
    cplx* dst = work;
    int len = n;
    int s = 1;
    cplx a[64];
    for (size_t st = 0; st < radix.size(); st++) {
        int r = radix[st];
        int m = len / r;
        const cplx* w = &tw[tw_off[st]];
        for (int p = 0; p < m; p++) {
            for (int q = 0; q < s; q++) {
                for (int j = 0; j < r; j++) a[j] = src[q + s * (p + j * m)];
                small_dft(r, a, sign);
                for (int t = 0; t < r; t++) {
                    cplx wt = w[p * r + t];
                    if (sign > 0) wt = std::conj(wt);
                    dst[q + s * (r * p + t)] = a[t] * wt;
                }
            }
        }
        len = m;
        s *= r;
        cplx* tmp = src; src = dst; dst = tmp;
    }
    if (src != x) memcpy(x, src, sizeof(cplx) * n);
}

void FFT3D::init(int K_) {
    K = K_;
    ldx = 2 * (K / 2 + 1);
    plan.init(K);
}

// R2C along x is computed as a full complex transform of the real row
// (imaginary part zero) keeping bins 0..K/2; y and z are complex transforms
// over the half spectrum.
void FFT3D::forward(double* mesh) const {
    const int Kh = K / 2 + 1;
    const size_t ldy = (size_t)ldx;          // doubles per row
    const size_t ldz = (size_t)K * ldx;      // doubles per plane
    #pragma omp parallel
    {
        std::vector<cplx> buf(K), work(K);
        #pragma omp for collapse(2) schedule(static)
        for (int z = 0; z < K; z++) {
            for (int y = 0; y < K; y++) {
                double* row = mesh + z * ldz + y * ldy;
                for (int x = 0; x < K; x++) buf[x] = cplx(row[x], 0.0);
                plan.exec(buf.data(), work.data(), -1);
                for (int x = 0; x < Kh; x++) {
                    row[2 * x] = buf[x].real();
                    row[2 * x + 1] = buf[x].imag();
                }
            }
        }
        #pragma omp for collapse(2) schedule(static)
        for (int z = 0; z < K; z++) {
            for (int x = 0; x < Kh; x++) {
                double* base = mesh + z * ldz + 2 * x;
                for (int y = 0; y < K; y++)
                    buf[y] = cplx(base[y * ldy], base[y * ldy + 1]);
                plan.exec(buf.data(), work.data(), -1);
                for (int y = 0; y < K; y++) {
                    base[y * ldy] = buf[y].real();
                    base[y * ldy + 1] = buf[y].imag();
                }
            }
        }
        #pragma omp for collapse(2) schedule(static)
        for (int y = 0; y < K; y++) {
            for (int x = 0; x < Kh; x++) {
                double* base = mesh + y * ldy + 2 * x;
                for (int z = 0; z < K; z++)
                    buf[z] = cplx(base[z * ldz], base[z * ldz + 1]);
                plan.exec(buf.data(), work.data(), -1);
                for (int z = 0; z < K; z++) {
                    base[z * ldz] = buf[z].real();
                    base[z * ldz + 1] = buf[z].imag();
                }
            }
        }
    }
}

void FFT3D::backward(double* mesh) const {
    const int Kh = K / 2 + 1;
    const size_t ldy = (size_t)ldx;
    const size_t ldz = (size_t)K * ldx;
    const double scale = 1.0 / ((double)K * K * K);
    #pragma omp parallel
    {
        std::vector<cplx> buf(K), work(K);
        #pragma omp for collapse(2) schedule(static)
        for (int y = 0; y < K; y++) {
            for (int x = 0; x < Kh; x++) {
                double* base = mesh + y * ldy + 2 * x;
                for (int z = 0; z < K; z++)
                    buf[z] = cplx(base[z * ldz], base[z * ldz + 1]);
                plan.exec(buf.data(), work.data(), +1);
                for (int z = 0; z < K; z++) {
                    base[z * ldz] = buf[z].real();
                    base[z * ldz + 1] = buf[z].imag();
                }
            }
        }
        #pragma omp for collapse(2) schedule(static)
        for (int z = 0; z < K; z++) {
            for (int x = 0; x < Kh; x++) {
                double* base = mesh + z * ldz + 2 * x;
                for (int y = 0; y < K; y++)
                    buf[y] = cplx(base[y * ldy], base[y * ldy + 1]);
                plan.exec(buf.data(), work.data(), +1);
                for (int y = 0; y < K; y++) {
                    base[y * ldy] = buf[y].real();
                    base[y * ldy + 1] = buf[y].imag();
                }
            }
        }
        // C2R along x: rebuild the Hermitian row, inverse, keep real part.
        // The imaginary parts of bins 0 and K/2 (K even) are ignored, as a
        // C2R transform does.
        #pragma omp for collapse(2) schedule(static)
        for (int z = 0; z < K; z++) {
            for (int y = 0; y < K; y++) {
                double* row = mesh + z * ldz + y * ldy;
                for (int x = 0; x < Kh; x++) buf[x] = cplx(row[2 * x], row[2 * x + 1]);
                buf[0] = cplx(buf[0].real(), 0.0);
                if (K % 2 == 0) buf[K / 2] = cplx(buf[K / 2].real(), 0.0);
                for (int x = Kh; x < K; x++) buf[x] = std::conj(buf[K - x]);
                plan.exec(buf.data(), work.data(), +1);
                for (int x = 0; x < K; x++) row[x] = buf[x].real() * scale;
                for (int x = K; x < ldx; x++) row[x] = 0.0;
            }
        }
    }
}

}  // namespace orc
