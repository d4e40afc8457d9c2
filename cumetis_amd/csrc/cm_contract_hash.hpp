// cm_contract_hash.hpp -- hash-based first-seen contraction of the coarse
// segments above REG_SEG entries (see cm_contract_hash.hip).
#pragma once
#include <stdint.h>

#include <vector>

struct cm_ctx;

namespace cm {

struct Level;

constexpr int REG_SEG = 16;  // coarse segments up to this merge in registers
// Hashed segments up to this length have their staged rows copied to the
// final place by k_contract_small (with the short rows of its tile); only
// longer ones need the emit pass.
constexpr int TILE_COPY_SEG = 64;

// Device state between the two halves: everything the emit step reads.
struct HashContract {
  std::vector<void *> blocks;  // pool blocks to release
  void *chunks = nullptr;
  unsigned *ctr = nullptr;     // item counters and the hub-table overflow flag
  int32_t *tc = nullptr, *tw = nullptr;  // compacted rows, at the segment offsets
  unsigned gchunk = 1;
  int64_t seg_max = -1;  // bound on the segment length (-1 unknown)
};

// Rows of every segment above REG_SEG into the staging area, their lengths
// into cdeg[c].  Queues work only; no host synchronisation.  seg_max bounds
// the segment length (-1 unknown); passes for longer classes are skipped.
template <typename OffT>
void hash_prepare_impl(cm_ctx *ctx, const Level &L, const int32_t *match, const int32_t *cmap, int64_t nc_bound,
                       const int64_t *d_nc, const int32_t *rep, const OffT *segoff, int32_t *cdeg,
                       int64_t seg_max, HashContract &hc);
// Staging area -> final rows once cxadj is written; queues the overflow flag
// into h_small[6] for the caller's next synchronisation.
template <typename OffT>
void hash_emit_impl(cm_ctx *ctx, const OffT *segoff, const int32_t *cdeg, const OffT *cxadj, int32_t *cadjncy,
                    int32_t *cadjwgt, const HashContract &hc);
void hash_release(cm_ctx *ctx, HashContract &hc);

}  // namespace cm
