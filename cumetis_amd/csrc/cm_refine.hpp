// cm_refine.hpp -- one deterministic k-way refinement pass (cm_refine.hip).
#pragma once
#include <stdint.h>

#include <vector>

#include "cm_internal.hpp"

namespace cm {
// Refines d_where of level L in place; pw (part weights) is updated; returns
// the number of accepted moves.
int refine_level(cm_ctx *ctx, const Level &L, int nparts, int32_t *d_where, std::vector<int64_t> &pw,
                 int64_t tvwgt);
// Moves vertices out of parts above int(w/k * 1.03) (the refinement's upper
// bound), highest connectivity gain first, into parts with room; no
// counterpart in the reference, which does not enforce the bound.
int balance_level(cm_ctx *ctx, const Level &L, int nparts, int32_t *d_where, std::vector<int64_t> &pw,
                  int64_t tvwgt);
}  // namespace cm
