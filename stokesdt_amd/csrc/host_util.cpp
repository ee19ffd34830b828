// Host utilities for synthetic workloads: random sequential addition of
// non-overlapping equal spheres in a periodic cubic box.
#include <cmath>
#include <cstdint>
#include <vector>

namespace sdt {

// splitmix64 (Steele, Lea, Flood 2014) as the proposal generator
static inline uint64_t splitmix64(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Returns the number of particles placed (n on success). Centres are at
// least 2a apart (minimum image); proposals are uniform in [0, L)^3.
long long rsa_positions(int n, double box, double a, uint64_t seed, long long max_attempts,
                        double* pos) {
    const double dmin = 2.0 * a, dmin2 = dmin * dmin;
    int nc = (int)(box / dmin);
    if (nc < 3) nc = 3;
    const double cs = box / nc;
    std::vector<std::vector<int>> cells((size_t)nc * nc * nc);
    uint64_t st = seed * 0x2545F4914F6CDD1Dull + 1;
    int placed = 0;
    long long attempts = 0;
    auto cell_of = [&](double x) {
        int c = (int)(x / cs);
        return c >= nc ? nc - 1 : c;
    };
    while (placed < n && attempts < max_attempts) {
        attempts++;
        double c[3];
        for (int d = 0; d < 3; d++) c[d] = (double)(splitmix64(st) >> 11) * (1.0 / 9007199254740992.0) * box;
        int ic[3] = {cell_of(c[0]), cell_of(c[1]), cell_of(c[2])};
        bool ok = true;
        for (int dx = -1; dx <= 1 && ok; dx++)
            for (int dy = -1; dy <= 1 && ok; dy++)
                for (int dz = -1; dz <= 1 && ok; dz++) {
                    int jx = (ic[0] + dx + nc) % nc, jy = (ic[1] + dy + nc) % nc,
                        jz = (ic[2] + dz + nc) % nc;
                    for (int j : cells[((size_t)jx * nc + jy) * nc + jz]) {
                        double r2 = 0.0;
                        for (int d = 0; d < 3; d++) {
                            double v = c[d] - pos[3 * j + d];
                            v -= box * std::nearbyint(v / box);
                            r2 += v * v;
                        }
                        if (r2 < dmin2) { ok = false; break; }
                    }
                }
        if (!ok) continue;
        for (int d = 0; d < 3; d++) pos[3 * placed + d] = c[d];
        cells[((size_t)ic[0] * nc + ic[1]) * nc + ic[2]].push_back(placed);
        placed++;
    }
    return placed;
}

}  // namespace sdt
