// tune.h — the autotuner's table, consulted by the dispatchers.
#pragma once

#include "internal.h"

namespace blasdsp {

struct TuneChoice {
    int algo;     // BLASDSP_ALGO_TILE or BLASDSP_ALGO_SKINNY
    int tile;     // tile edge for the tile kernel (64 or 128)
    int splitk;   // K slices for the skinny kernel
    double us;    // measured time of one call, microseconds
};

// True and *out filled when the shape's class has been tuned (or loaded).
bool tune_lookup(bool f64, bool ta, bool tb, int64_t M, int64_t N, int64_t K, TuneChoice *out);

}  // namespace blasdsp
