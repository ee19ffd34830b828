// Autotuner of the SPME launch choices: every candidate computes the same
// operator (the FFT stage order changes only its rounding), so the choice is
// made by timing alone, on the caller's own system.
#include <cmath>
#include <cstdio>
#include <fstream>
#include <sstream>
#include "block_gpu.hpp"
#include "tune.hpp"

namespace sdt {

static std::string tune_key(MobSpmeGpu* m, int b) {
    char buf[160];
    const int oct = (int)std::lround(std::log2((double)m->npos()));
    std::snprintf(buf, sizeof(buf), "spme-v1 prec=%d K=%d p=%d poly=%d noct=%d b=%d",
                  m->precision(), m->K(), m->porder(), m->mono() ? 0 : 1, oct, b);
    return buf;
}

static bool cache_lookup(const std::string& path, const std::string& key, SpmeTuning* t) {
    if (path.empty()) return false;
    std::ifstream in(path);
    std::string line;
    bool found = false;
    while (std::getline(in, line)) {
        if (line.compare(0, key.size(), key) != 0 || line.size() <= key.size() ||
            line[key.size()] != ' ')
            continue;
        std::istringstream ss(line.substr(key.size() + 1));
        SpmeTuning c;
        if (ss >> c.fft_order >> c.interp_sxt >> c.interp_ch >> c.spread_stage_kb) {
            *t = c;
            found = true;   // the last entry of a key wins
        }
    }
    return found;
}

static void cache_append(const std::string& path, const std::string& key, const SpmeTuning& t,
                         double ms0, double ms1) {
    if (path.empty()) return;
    std::ofstream out(path, std::ios::app);
    if (!out) return;   // an unwritable cache only costs a re-tune
    out << key << ' ' << t.fft_order << ' ' << t.interp_sxt << ' ' << t.interp_ch << ' '
        << t.spread_stage_kb << "  # ms " << ms0 << " -> " << ms1 << '\n';
}

TuneResult autotune_spme(MobSpmeGpu* mob, int b, const std::string& cache_path) {
    SDT_REQUIRE(mob && mob->updated(), SDT_ESTATE, "autotune: the mobility must be updated first");
    SDT_REQUIRE(b >= 1, SDT_EINVAL, "autotune: at least one right-hand side");
    TuneResult r;
    const std::string key = tune_key(mob, b);
    if (cache_lookup(cache_path, key, &r.tuning)) {
        mob->set_tuning(r.tuning, b);
        r.cached = true;
        return r;
    }
    hipStream_t s = mob->stream();
    const size_t n = (size_t)mob->dim() * b;
    DevBuf<double> X, Y;
    X.reserve(n);
    Y.reserve(n);
    gaussian_fill_device(0x7e57ull, 0, (long long)n, X.get(), s);
    hipEvent_t e0, e1;
    SDT_HIP_CHECK(hipEventCreate(&e0));
    SDT_HIP_CHECK(hipEventCreate(&e1));
    auto time_of = [&](const SpmeTuning& t) {
        mob->set_tuning(t, b);
        mob->apply_packed(0, b, 1.0, X.get(), 0.0, Y.get());   // first use of the launch
        const int reps = 3;
        SDT_HIP_CHECK(hipEventRecord(e0, s));
        for (int k = 0; k < reps; k++) mob->apply_packed(0, b, 1.0, X.get(), 0.0, Y.get());
        SDT_HIP_CHECK(hipEventRecord(e1, s));
        SDT_HIP_CHECK(hipEventSynchronize(e1));
        float ms = 0.f;
        SDT_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
        r.tried++;
        return (double)ms / reps;
    };
    try {
        SpmeTuning best;
        double tbest = time_of(best);
        r.ms_default = tbest;
        auto consider = [&](const SpmeTuning& c) {
            if (c == best) return;
            const double t = time_of(c);
            if (t < 0.98 * tbest) {
                tbest = t;
                best = c;
            }
        };
        FFTPlan p0;
        make_fft_plan(mob->K(), &p0, 0);
        for (int order = 1; order <= 2; order++) {
            FFTPlan p;
            if (!make_fft_plan(mob->K(), &p, order)) continue;
            bool same = p.nst == p0.nst;
            for (int i = 0; same && i < p.nst; i++) same = p.radix[i] == p0.radix[i];
            if (same) continue;
            SpmeTuning c = best;
            c.fft_order = order;
            consider(c);
        }
        const int strips[][2] = {{1, 64}, {2, 64}, {2, 128}, {4, 64}, {4, 128}, {1, 128}};
        for (const auto& sc : strips) {
            SpmeTuning c = best;
            c.interp_sxt = sc[0];
            c.interp_ch = sc[1];
            consider(c);
        }
        for (int kb : {32, 48, 64}) {
            SpmeTuning c = best;
            c.spread_stage_kb = kb;
            consider(c);
        }
        mob->set_tuning(best, b);
        r.tuning = best;
        r.ms_best = tbest;
    } catch (...) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    cache_append(cache_path, key, r.tuning, r.ms_default, r.ms_best);
    return r;
}

}  // namespace sdt
