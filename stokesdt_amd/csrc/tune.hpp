// Project autotuner of the SPME launch choices (FFT stage order, interp
// strip and particle chunk, spread stage size) for one mobility engine.
#pragma once
#include <string>
#include "mob_gpu.hpp"

namespace sdt {

struct TuneResult {
    SpmeTuning tuning;
    double ms_default = 0.0;   // packed matvec of b columns with the heuristics
    double ms_best = 0.0;      // with the chosen tuning
    int tried = 0;             // candidates timed (0 when read from the cache)
    bool cached = false;
};

// Times the packed b-column matvec of an updated mobility for candidate
// choices (coordinate descent: FFT order, then interp, then spread; a
// candidate replaces the current one when it is more than 2 % faster),
// applies the best to the width class of b (narrow b < 8, wide b >= 8) and
// returns it. cache_path (may be empty) names a text
// file of earlier results, one "key tuning" line each, keyed by precision,
// mesh, order, polydispersity, particle-count octave and b; a hit is applied
// without timing, a new result is appended.
TuneResult autotune_spme(MobSpmeGpu* mob, int b, const std::string& cache_path);

}  // namespace sdt
