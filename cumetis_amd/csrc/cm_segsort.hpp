// cm_segsort.hpp -- binned segmented sort (see cm_segsort.hip).
#pragma once
#include <stdint.h>

struct cm_ctx;

namespace cm {

constexpr int SEG_NBINS = 13;   // 0..5 wave bins, 6..11 LDS bins, 12 hub rows
constexpr int HUB_TILE = 4096;  // LDS tile of the hub path
constexpr int SEG_LARGE_BIN = 6;  // first bin with more than 64 keys

// Work unit of the hub path: merge runs A=[a,a+alen) and B=[a+alen,a+alen+blen)
// and write outputs [d0, d0+HUB_TILE) of the merged range (blen 0 = copy);
// for tile sorts, [a+d0, a+d0+alen) is one tile.
struct MergeDesc {
  int64_t a;
  int32_t alen, blen, d0, pad;
};

// Segments grouped by size class: list[off[c] .. off[c+1]) are the segment
// ids of class c (segments of length <= 1 are not listed).
struct SegBins {
  int32_t *list = nullptr;
  int64_t off[SEG_NBINS + 1] = {0};
};

// Where the sort takes its scratch: the context's pool (arena == nullptr)
// or a bump arena over device memory the caller provides, which is never
// freed piecewise (the whole arena is dropped after the call).
struct SegArena {
  char *base = nullptr;
  size_t cap = 0, used = 0;
};
void *seg_alloc(cm_ctx *ctx, SegArena *ar, size_t bytes);
void seg_release(cm_ctx *ctx, SegArena *ar, void *p);
// Arena bytes segsort_i32 may take for n keys in `length` segments.
size_t segsort_i32_arena_bytes(int64_t n, int64_t length);

void bin_segments(cm_ctx *ctx, int64_t nseg, const void *segoff, bool off64, SegBins &b,
                  SegArena *ar = nullptr);
void release_bins(cm_ctx *ctx, SegBins &b, SegArena *ar = nullptr);
void sort_binned(cm_ctx *ctx, const SegBins &b, const void *segoff, bool off64, int64_t total,
                 uint64_t *keys, uint32_t *vals, int first_bin = 0, SegArena *ar = nullptr);
void segsort_i32(cm_ctx *ctx, int32_t *keys, int32_t *vals, int32_t n, const int32_t *segs,
                 int32_t length, SegArena *ar = nullptr);

}  // namespace cm
