// cm_tuned.hpp -- launch-shape defaults of a new context on gfx950, written
// by cumetis_amd/tune.py from its measurements (see that script for the
// workloads and candidates; the record is cumetis_amd/tuned_gfx950.json).
// Every value changes launch shapes only, never a result.
#pragma once
#include <stdint.h>

namespace cm {
constexpr int32_t TUNED_FIRST_ROUND_BATCH = 9;
constexpr int32_t TUNED_LEVEL_BATCH_MARGIN = 2;
constexpr int32_t TUNED_LARGE_LEVEL_BATCH = 13;
constexpr int32_t TUNED_COMPACT_ROUND = 5;
constexpr int64_t TUNED_COMPACT_MIN_N = int64_t(1048576);
constexpr int64_t TUNED_SPECULATE_EDGES = int64_t(16777216);
constexpr int32_t TUNED_NUMBER_ONE_TILE_MAX = 4096;
constexpr int32_t TUNED_SEG_WAVE_BINS = 6;
constexpr int32_t TUNED_GRAPH_ROUNDS = 0;
}  // namespace cm
