// Oracle random numbers (test infrastructure only; see orc_common.hpp).
//
// The reference draws its Gaussian right-hand sides with MKL VSL
// (SFMT19937 + Box-Muller, src/util/rnd_stream.cc:20-65). That stream cannot
// be reproduced without MKL, so both this oracle and the product use
// Philox-4x32-10 (Salmon et al., "Parallel random numbers: as easy as 1, 2,
// 3", SC'11) with the Box-Muller transform in FP64, and the same counters.
// Pinned against the Random123 known-answer vectors in tests.
#include <cmath>
#include "orc_common.hpp"

namespace orc {

static inline void mulhilo32(uint32_t a, uint32_t b, uint32_t* hi, uint32_t* lo) {
    uint64_t p = (uint64_t)a * (uint64_t)b;
    *hi = (uint32_t)(p >> 32);
    *lo = (uint32_t)p;
}

void philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
    const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; r++) {
        uint32_t hi0, lo0, hi1, lo1;
        mulhilo32(M0, c0, &hi0, &lo0);
        mulhilo32(M1, c2, &hi1, &lo1);
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n1 = lo1;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        uint32_t n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += W0;
        k1 += W1;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

// Element pair q of stream `stream` uses counter {q_lo, q_hi, stream, 0}
// and key {seed_lo, seed_hi}. The four output words give two 53-bit
// uniforms in (0,1); Box-Muller turns them into elements 2q and 2q+1.
void gaussian_fill(uint64_t seed, uint32_t stream, int64_t n, double* out) {
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    const int64_t npairs = (n + 1) / 2;
    #pragma omp parallel for schedule(static)
    for (int64_t q = 0; q < npairs; q++) {
        uint32_t ctr[4] = {(uint32_t)q, (uint32_t)((uint64_t)q >> 32), stream, 0u};
        uint32_t w[4];
        philox4x32_10(ctr, key, w);
        uint64_t a = ((uint64_t)w[0] << 32 | w[1]) >> 11;
        uint64_t b = ((uint64_t)w[2] << 32 | w[3]) >> 11;
        const double inv53 = 1.0 / 9007199254740992.0;  // 2^-53
        double u1 = ((double)a + 0.5) * inv53;
        double u2 = ((double)b + 0.5) * inv53;
        double r = sqrt(-2.0 * log(u1));
        double t = 2.0 * M_PI * u2;
        out[2 * q] = r * cos(t);
        if (2 * q + 1 < n) out[2 * q + 1] = r * sin(t);
    }
}

}  // namespace orc
