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
// Enqueues the edge cut of d_where on level L (cu:356-366 counts it on the
// host): *d_sum receives the weight of the directed entries whose ends lie in
// different parts, each undirected edge twice; the caller halves it.
void level_cut(cm_ctx *ctx, const Level &L, const int32_t *d_where, unsigned long long *d_sum);
}  // namespace cm
