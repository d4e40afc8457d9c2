// cm_contract_hash.hpp -- hash-based first-seen contraction of the coarse
// segments above REG_SEG entries (see cm_contract_hash.hip): size classes,
// item lists, and the select step shared by its own kernel and the
// numbering kernel.
#pragma once
#include <stdint.h>

#include <vector>

#include "cm_common.hpp"

struct cm_ctx;

namespace cm {

struct Level;

constexpr int REG_SEG = 16;  // coarse segments up to this merge in registers
// When no segment can be longer than this, k_contract_small copies the
// staged hashed rows to their final place with the short rows of its tile
// and the emit pass is skipped (longer segments make the emit pass, which
// copies cooperatively, the cheaper one for all rows).
constexpr int TILE_COPY_SEG = 64;

// Size classes of the segments above REG_SEG: up to HC_WAVE_MAX entries one
// wave (a half-size table up to HC_WAVE_SPLIT), up to HC_BLOCK_MAX one block
// (a half-size table up to HC_BLOCK_SPLIT), longer ones ("hubs") one block
// per key partition of about HC_PART entries.  Every segment is also cut
// into chunks of HC_CHUNK entries for the gather, count and copy passes.
constexpr int HC_CHUNK = 2048;
constexpr int HC_WAVE_SPLIT = 64;
constexpr int HC_WAVE_MAX = 256;
constexpr int HC_BLOCK_SPLIT = 1024;
constexpr int HC_BLOCK_MAX = 2048;
constexpr int HC_PART = 4096;
// Hubs of at most HC_PMAX key partitions have their entries bucketed by
// partition inside each chunk, so a partition block reads only its own
// slices; longer ones are read whole by every partition block.
constexpr int HC_PMAX = 256;

struct SegChunk {
  int32_t c, first, nch, k;  // chunk k of the nch chunks of segment c, from chunk `first`
};
struct HubPart {
  int32_t c, p, P, hfirst;  // key partition p of P of hub segment c, whose chunks start at hfirst
};

// item counters (HC_N_HCHUNK: chunks of hub segments, listed again densely)
enum { HC_N_CHUNK = 0, HC_N_WAVE, HC_N_BLOCK, HC_N_PART, HC_N_HUB, HC_N_BLOCKL, HC_N_WAVEL, HC_N_HCHUNK, HC_N_COUNTERS };
constexpr int HC_OVERFLOW = 8;  // ctr[HC_OVERFLOW]: a hub table filled
static_assert(HC_N_COUNTERS <= HC_OVERFLOW, "counters below the overflow word");
constexpr int HC_CTR_WORDS = 16;

// The item lists (device pointers) and their counters.
struct HcLists {
  SegChunk *chunks = nullptr;
  SegChunk *hchunks = nullptr;  // the hub chunks again, one dense run per hub
  int32_t *wave = nullptr, *wavel = nullptr, *block = nullptr, *blockl = nullptr, *hubs = nullptr;
  HubPart *parts = nullptr;
  unsigned *ctr = nullptr;
};

// Items a segment of len entries adds to each list.  Every counter is
// updated by its own predicated add (a class chosen by branches would merge
// into one dynamically indexed update, which puts the array in scratch).
__device__ __forceinline__ void hc_count(int64_t len, unsigned (&cnt)[HC_N_COUNTERS])
{
  if (len <= REG_SEG) return;
  const bool hub = len > HC_BLOCK_MAX;
  cnt[HC_N_CHUNK] += (unsigned)((len + HC_CHUNK - 1) / HC_CHUNK);
  cnt[HC_N_WAVE] += len <= HC_WAVE_SPLIT ? 1u : 0u;
  cnt[HC_N_WAVEL] += len > HC_WAVE_SPLIT && len <= HC_WAVE_MAX ? 1u : 0u;
  cnt[HC_N_BLOCK] += len > HC_WAVE_MAX && len <= HC_BLOCK_SPLIT ? 1u : 0u;
  cnt[HC_N_BLOCKL] += len > HC_BLOCK_SPLIT && !hub ? 1u : 0u;
  cnt[HC_N_HUB] += hub ? 1u : 0u;
  cnt[HC_N_PART] += hub ? (unsigned)((len + HC_PART - 1) / HC_PART) : 0u;
  cnt[HC_N_HCHUNK] += hub ? (unsigned)((len + HC_CHUNK - 1) / HC_CHUNK) : 0u;
}

// The items of segment c at the positions at[] (advanced past them).
__device__ __forceinline__ void hc_write(int32_t c, int64_t len, unsigned (&at)[HC_N_COUNTERS], const HcLists &hl)
{
  if (len <= REG_SEG) return;
  const int nch = (int)((len + HC_CHUNK - 1) / HC_CHUNK);
  const unsigned first = at[HC_N_CHUNK];
  for (int k = 0; k < nch; ++k) hl.chunks[first + k] = SegChunk{c, (int32_t)first, nch, k};
  at[HC_N_CHUNK] += nch;
  const bool hub = len > HC_BLOCK_MAX;
  const bool cw = len <= HC_WAVE_SPLIT, cwl = !cw && len <= HC_WAVE_MAX;
  const bool cb = len > HC_WAVE_MAX && len <= HC_BLOCK_SPLIT, cbl = len > HC_BLOCK_SPLIT && !hub;
  if (cw) hl.wave[at[HC_N_WAVE]] = c;
  if (cwl) hl.wavel[at[HC_N_WAVEL]] = c;
  if (cb) hl.block[at[HC_N_BLOCK]] = c;
  if (cbl) hl.blockl[at[HC_N_BLOCKL]] = c;
  at[HC_N_WAVE] += cw ? 1u : 0u;
  at[HC_N_WAVEL] += cwl ? 1u : 0u;
  at[HC_N_BLOCK] += cb ? 1u : 0u;
  at[HC_N_BLOCKL] += cbl ? 1u : 0u;
  if (hub) {
    const unsigned hf = at[HC_N_HCHUNK];
    hl.hubs[at[HC_N_HUB]] = (int32_t)hf;  // a hub is named by its first chunk in hchunks
    for (int k = 0; k < nch; ++k) hl.hchunks[hf + k] = SegChunk{c, (int32_t)hf, nch, k};
    const int P = (int)((len + HC_PART - 1) / HC_PART);
    for (int p = 0; p < P; ++p) hl.parts[at[HC_N_PART] + p] = HubPart{c, p, P, (int32_t)hf};
  }
  at[HC_N_HUB] += hub ? 1u : 0u;
  at[HC_N_PART] += hub ? (unsigned)((len + HC_PART - 1) / HC_PART) : 0u;
  at[HC_N_HCHUNK] += hub ? (unsigned)nch : 0u;
}

// Block of NW waves, called by all its threads: reserves every thread's
// counts in the lists with one counter update per block and list; at[]
// receives the thread's first positions.
template <int NW>
__device__ __forceinline__ void hc_reserve(const unsigned (&cnt)[HC_N_COUNTERS], unsigned (&at)[HC_N_COUNTERS],
                                           unsigned *ctr)
{
  __shared__ unsigned s_w[HC_N_COUNTERS][NW], s_base[HC_N_COUNTERS];
  const int w = threadIdx.x >> 6;
  unsigned inc[HC_N_COUNTERS];
#pragma unroll
  for (int i = 0; i < HC_N_COUNTERS; ++i) {
    inc[i] = wave_incl_scan<unsigned>(cnt[i]);
    if (lane() == CM_WAVE - 1) s_w[i][w] = inc[i];
  }
  __syncthreads();
  if (threadIdx.x < HC_N_COUNTERS) {
    const int i = threadIdx.x;
    unsigned tot = 0;
#pragma unroll
    for (int j = 0; j < NW; ++j) tot += s_w[i][j];
    s_base[i] = tot ? atomicAdd(&ctr[i], tot) : 0u;
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < HC_N_COUNTERS; ++i) {
    at[i] = s_base[i] + inc[i] - cnt[i];
    for (int j = 0; j < w; ++j) at[i] += s_w[i][j];
  }
}

// Packed staging.  On a level with unit edge weights a coarse neighbour's
// weight counts the fine edges between the two pairs: at most 4 in a graph
// without repeated edges.  Such a level stages a hashed row as one word per
// entry, (weight - 1) in the top two bits above the coarse id (the weight
// array is then not used); a larger weight (repeated edges in the input)
// sets the overflow word and the level is redone on the sort-based path.
constexpr int HC_PACK_SHIFT = 30;
constexpr int64_t HC_PACK_MAX_ID = (int64_t)1 << HC_PACK_SHIFT;  // coarse ids below this
__device__ __forceinline__ int32_t hc_pack(int32_t id, int32_t w)
{
  return (int32_t)(((uint32_t)(w - 1) << HC_PACK_SHIFT) | (uint32_t)id);
}
__device__ __forceinline__ int32_t hc_unpack_id(int32_t x) { return x & (int32_t)(HC_PACK_MAX_ID - 1); }
__device__ __forceinline__ int32_t hc_unpack_w(int32_t x) { return (int32_t)((uint32_t)x >> HC_PACK_SHIFT) + 1; }
// One staged entry: both words, or the packed one (tw == nullptr; a weight
// above 4 sets *ovf).
__device__ __forceinline__ void hc_stage(int32_t *tc, int32_t *tw, int64_t at, int32_t id, int32_t w, unsigned *ovf)
{
  if (tw) {
    tc[at] = id;
    tw[at] = w;
  } else {
    if (w > 4) *ovf = 1u;
    tc[at] = hc_pack(id, w);
  }
}

// Device state between the two halves: everything the emit step reads.
struct HashContract {
  std::vector<void *> blocks;  // pool blocks to release
  HcLists lists;
  int32_t *tc = nullptr, *tw = nullptr;  // compacted rows, at the segment offsets (tw null: packed)
  unsigned gchunk = 1;
  int64_t seg_max = -1;   // bound on the segment length (-1 unknown)
  int64_t nc_bound = 0;   // the lists hold this many coarse vertices at most
  bool selected = false;  // the lists were filled by the numbering kernel
};

// Whether level L can have segments above REG_SEG (each joins two rows).
bool hash_needed(const Level &L);
// The item lists and staging rows for nc_bound coarse vertices of L (no
// launches); number() can then fill the lists.
void hash_alloc(cm_ctx *ctx, const Level &L, int64_t nc_bound, HashContract &hc);
// Rows of every segment above REG_SEG into the staging area, their lengths
// into cdeg[c].  Queues work only; no host synchronisation.  hc may come
// from hash_alloc (its lists filled by number() or not); otherwise it is set
// up here.  seg_max bounds the segment length (-1 unknown); passes for
// longer classes are skipped.
template <typename OffT>
void hash_prepare_impl(cm_ctx *ctx, const Level &L, const int32_t *match, const int32_t *cmap, int64_t nc_bound,
                       const int64_t *d_nc, const int32_t *rep, const OffT *segoff, int32_t *cdeg,
                       int64_t seg_max, HashContract &hc);
// Staging area -> final rows once cxadj is written.
template <typename OffT>
void hash_emit_impl(cm_ctx *ctx, const OffT *segoff, const int32_t *cdeg, const OffT *cxadj, int32_t *cadjncy,
                    int32_t *cadjwgt, const HashContract &hc);
// The device word that is nonzero if a hub table filled (only hubs can
// overflow), or nullptr when no hub pass ran.
const unsigned *hash_overflow_flag(const HashContract &hc);
void hash_release(cm_ctx *ctx, HashContract &hc);

}  // namespace cm
