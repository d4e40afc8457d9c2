// cm_segsort.hip -- binned segmented sort of 64-bit keys carrying 32-bit
// values.  Plays the role of the reference's bb_segsort
// (include/bb_segsort.h:35-192, bb_bin.h, bb_comput_s.h, bb_comput_l.h),
// redesigned for wave64:
//   bin 0..5   (2..64 keys)   one sub-wave group of W = 2..64 lanes per
//                             segment, one key per lane, bitonic network over
//                             cross-lane shuffles;
//   bin 6..11  (65..4096)     one 256-thread workgroup per segment, bitonic
//                             network in LDS (padded to 128..4096);
//   bin 12     (> 4096)       hub rows: 4096-key LDS tile sorts, then
//                             merge-path passes that double the run width.
// Keys are unique within a segment wherever this engine uses the sort (they
// carry the original position in their low bits), so the result is fully
// determined.
#include "cm_common.hpp"
#include "cm_internal.hpp"
#include "cm_segsort.hpp"

#include <algorithm>

namespace cm {

__device__ __forceinline__ int seg_class(int64_t len)
{
  if (len <= 1) return -1;
  if (len > 4096) return 12;
  // ceil(log2(len)) - 1 for len in [2, 4096]
  const int lg = 64 - __clzll((unsigned long long)(len - 1));
  return lg - 1;
}

// Binning without contended atomics: every workgroup of 256 segments writes
// its per-class counts to hist[c * nblocks + block]; one exclusive scan over
// that class-major array gives each (class, block) its first slot, so the
// lists are grouped by class with ascending segment ids inside a class.
template <typename OffT>
__global__ __launch_bounds__(256) void k_bin_hist(int64_t nseg, const OffT *segoff, unsigned nb,
                                                  int32_t *hist)
{
  __shared__ unsigned sh[SEG_NBINS];
  if (threadIdx.x < SEG_NBINS) sh[threadIdx.x] = 0;
  __syncthreads();
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int cls = s < nseg ? seg_class((int64_t)(segoff[s + 1] - segoff[s])) : -1;
  for (int c = 0; c < SEG_NBINS; ++c) {
    const uint64_t mk = __ballot(cls == c);
    if (mk && lane() == 0) atomicAdd(&sh[c], (unsigned)__popcll(mk));
  }
  __syncthreads();
  if (threadIdx.x < SEG_NBINS) hist[(int64_t)threadIdx.x * nb + blockIdx.x] = (int32_t)sh[threadIdx.x];
}

template <typename OffT>
__global__ __launch_bounds__(256) void k_bin_scatter(int64_t nseg, const OffT *segoff, unsigned nb,
                                                     const int32_t *base, int32_t *list,
                                                     int64_t *class_off)
{
  __shared__ unsigned wcnt[SEG_NBINS][4];
  const int w = threadIdx.x >> 6;
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int cls = s < nseg ? seg_class((int64_t)(segoff[s + 1] - segoff[s])) : -1;
  uint64_t mine = 0;
  for (int c = 0; c < SEG_NBINS; ++c) {
    const uint64_t mk = __ballot(cls == c);
    if (lane() == 0) wcnt[c][w] = (unsigned)__popcll(mk);
    if (cls == c) mine = mk;
  }
  __syncthreads();
  if (cls >= 0) {
    unsigned before = 0;
    for (int i = 0; i < w; ++i) before += wcnt[cls][i];
    list[base[(int64_t)cls * nb + blockIdx.x] + before + mask_rank(mine)] = (int32_t)s;
  }
  if (blockIdx.x == 0 && threadIdx.x <= SEG_NBINS)
    class_off[threadIdx.x] = base[(int64_t)threadIdx.x * nb];
}

// ------------------------------------------------ bins 0..5: in registers
template <int W, typename OffT>
__global__ __launch_bounds__(256) void k_sort_wave(const int32_t *list, int64_t count,
                                                   const OffT *segoff, uint64_t *keys,
                                                   uint32_t *vals)
{
  const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t g = gid / W;
  const int t = (int)(gid % W);
  uint64_t key = ~0ull;
  uint32_t val = 0;
  int64_t s0 = 0;
  int len = 0;
  if (g < count) {
    const int32_t s = list[g];
    s0 = (int64_t)segoff[s];
    len = (int)((int64_t)segoff[s + 1] - s0);
    if (t < len) {
      key = keys[s0 + t];
      val = vals[s0 + t];
    }
  }
#pragma unroll
  for (int k = 2; k <= W; k <<= 1) {
#pragma unroll
    for (int j = k >> 1; j > 0; j >>= 1) {
      const uint64_t ok = __shfl_xor(key, j, CM_WAVE);
      const uint32_t ov = __shfl_xor(val, j, CM_WAVE);
      const bool up = (t & k) == 0;
      const bool lower = (t & j) == 0;
      const bool take = (lower == up) ? (ok < key) : (ok > key);
      if (take) {
        key = ok;
        val = ov;
      }
    }
  }
  if (g < count && t < len) {
    keys[s0 + t] = key;
    vals[s0 + t] = val;
  }
}

// ------------------------------------------------ bins 6..11: LDS bitonic
template <int S>
__device__ __forceinline__ void lds_sort(uint64_t *keys, uint32_t *vals, int64_t s0, int len,
                                         uint64_t *sk, uint32_t *sv)
{
  for (int i = threadIdx.x; i < S; i += blockDim.x) {
    if (i < len) {
      sk[i] = keys[s0 + i];
      sv[i] = vals[s0 + i];
    } else {
      sk[i] = ~0ull;
      sv[i] = 0;
    }
  }
  __syncthreads();
  for (int k = 2; k <= S; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int t = threadIdx.x; t < S / 2; t += blockDim.x) {
        const int i = ((t & ~(j - 1)) << 1) | (t & (j - 1));  // 2j * (t / j) + t % j, j a power of two
        const int l = i + j;
        const bool up = (i & k) == 0;
        const uint64_t a = sk[i], b = sk[l];
        if (up ? (a > b) : (a < b)) {
          sk[i] = b;
          sk[l] = a;
          const uint32_t x = sv[i];
          sv[i] = sv[l];
          sv[l] = x;
        }
      }
      __syncthreads();
    }
  }
  for (int i = threadIdx.x; i < len; i += blockDim.x) {
    keys[s0 + i] = sk[i];
    vals[s0 + i] = sv[i];
  }
}

template <int S, typename OffT>
__global__ __launch_bounds__(256) void k_sort_lds(const int32_t *list, int64_t count,
                                                  const OffT *segoff, uint64_t *keys,
                                                  uint32_t *vals)
{
  __shared__ uint64_t sk[S];
  __shared__ uint32_t sv[S];
  for (int64_t g = blockIdx.x; g < count; g += gridDim.x) {
    const int32_t s = list[g];
    const int64_t s0 = (int64_t)segoff[s];
    const int len = (int)((int64_t)segoff[s + 1] - s0);
    lds_sort<S>(keys, vals, s0, len, sk, sv);
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) void k_sort_tiles(const MergeDesc *tiles, int64_t count,
                                                    uint64_t *keys, uint32_t *vals)
{
  __shared__ uint64_t sk[HUB_TILE];
  __shared__ uint32_t sv[HUB_TILE];
  for (int64_t g = blockIdx.x; g < count; g += gridDim.x) {
    const MergeDesc d = tiles[g];
    lds_sort<HUB_TILE>(keys, vals, d.a + d.d0, d.alen, sk, sv);
    __syncthreads();
  }
}

// ------------------------------------------------ bin 12: merge-path passes
// One workgroup writes outputs [d0, d1) (at most HUB_TILE) of the merge of
// runs A=[a, a+alen) and B=[a+alen, a+alen+blen), A first on equal keys.
// Two threads find the block's co-ranks in global memory; the block then
// stages the A and B keys those outputs come from (at most HUB_TILE of them
// together) in LDS, and each thread finds its own co-rank and merges
// MERGE_ITEMS outputs there.  (Searching and merging straight from global
// memory made every step a dependent load.)
constexpr int MERGE_ITEMS = HUB_TILE / 256;

__device__ __forceinline__ int merge_corank(const uint64_t *A, int alen, const uint64_t *B, int blen, int d)
{
  int lo = max(0, d - blen), hi = min(d, alen);
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (A[mid] <= B[d - mid - 1]) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

__global__ __launch_bounds__(256) void k_merge(const MergeDesc *descs, int64_t count,
                                               const uint64_t *sk, const uint32_t *sv,
                                               uint64_t *dk, uint32_t *dv)
{
  __shared__ uint64_t s_k[HUB_TILE];  // the block's A keys, then its B keys
  __shared__ int s_c[2];
  for (int64_t g = blockIdx.x; g < count; g += gridDim.x) {
    const MergeDesc q = descs[g];
    const int alen = q.alen, blen = q.blen, total = alen + blen;
    const uint64_t *A = sk + q.a;
    const uint64_t *B = A + alen;
    const int d0 = q.d0, d1 = min(d0 + HUB_TILE, total);
    if (threadIdx.x < 2) s_c[threadIdx.x] = merge_corank(A, alen, B, blen, threadIdx.x ? d1 : d0);
    __syncthreads();
    const int i0 = s_c[0], j0 = d0 - i0;
    const int na = s_c[1] - i0, nb = (d1 - s_c[1]) - j0;
    for (int x = threadIdx.x; x < na + nb; x += blockDim.x) s_k[x] = x < na ? A[i0 + x] : B[j0 + x - na];
    __syncthreads();
    const uint64_t *LA = s_k, *LB = s_k + na;
    const int d = threadIdx.x * MERGE_ITEMS;
    if (d < na + nb) {
      const int dend = min(d + MERGE_ITEMS, na + nb);
      int i = merge_corank(LA, na, LB, nb, d), j = d - i;
      for (int o = d; o < dend; ++o) {
        const bool ta = j >= nb || (i < na && LA[i] <= LB[j]);
        const int64_t out = q.a + d0 + o;
        if (ta) {
          dk[out] = LA[i];
          dv[out] = sv[q.a + i0 + i];
          ++i;
        } else {
          dk[out] = LB[j];
          dv[out] = sv[q.a + alen + j0 + j];
          ++j;
        }
      }
    }
    __syncthreads();  // s_k and s_c are reused
  }
}

template <typename OffT>
__global__ __launch_bounds__(256) void k_hub_info(const int32_t *list, int64_t count,
                                                  const OffT *segoff, int64_t *info)
{
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= count) return;
  const int32_t s = list[g];
  info[2 * g] = (int64_t)segoff[s];
  info[2 * g + 1] = (int64_t)segoff[s + 1] - (int64_t)segoff[s];
}

// ---------------------------------------------------------------- host side
constexpr size_t ARENA_ALIGN = 256;

static size_t arena_round(size_t bytes) { return (bytes + ARENA_ALIGN - 1) / ARENA_ALIGN * ARENA_ALIGN; }

void *seg_alloc(cm_ctx *ctx, SegArena *ar, size_t bytes)
{
  if (!ar) return ctx->pool.alloc(bytes);
  const size_t b = arena_round(bytes ? bytes : 1);
  CM_REQUIRE(ar->used + b <= ar->cap, CM_ERR_UNSUPPORTED, "segmented sort scratch exhausted");
  void *p = ar->base + ar->used;
  ar->used += b;
  return p;
}

void seg_release(cm_ctx *ctx, SegArena *ar, void *p)
{
  if (!ar) ctx->pool.release(p);
}

// Status words of a look-back scan over n entries (1024-entry tiles): the
// context's (pool mode), or cleared arena words with epoch 1 (arena memory
// holds whatever the last call left there).
static void scan_excl_arena(cm_ctx *ctx, SegArena *ar, int64_t n, const int32_t *in, int32_t *out)
{
  if (!ar) {
    scan_excl_i32(ctx, n, in, out);
    return;
  }
  const int64_t tiles = n > 0 ? (n + 1023) / 1024 : 1;
  uint64_t *st = (uint64_t *)seg_alloc(ctx, ar, sizeof(uint64_t) * (size_t)tiles);
  fill_u64(ctx, st, 0ull, tiles);
  scan_excl_i32_status(ctx, n, in, out, st, 1u);
}

// Largest number of hub merge descriptors one pass can need for `total` keys:
// per hub row of len > HUB_TILE keys, at most len / HUB_TILE + 1 tiles per
// run pair plus one per pair, with fewer than total / HUB_TILE rows.
static int64_t desc_bound(int64_t total) { return 2 * (total / HUB_TILE) + 16; }

size_t segsort_i32_arena_bytes(int64_t n, int64_t length)
{
  const int64_t nb = (length + 255) / 256 > 0 ? (length + 255) / 256 : 1;
  const int64_t nh = (int64_t)SEG_NBINS * nb;
  size_t b = 0;
  b += arena_round(4 * (size_t)(length + 1));                 // segoff
  b += arena_round(8 * (size_t)std::max<int64_t>(n, 1));      // packed keys
  b += arena_round(4 * (size_t)std::max<int64_t>(n, 1));      // values
  b += arena_round(4 * (size_t)std::max<int64_t>(length, 1)); // bin lists
  b += 2 * arena_round(4 * (size_t)(nh + 1));                 // histogram, bases
  b += arena_round(8 * (size_t)((nh + 1023) / 1024 + 1));     // scan status words
  b += arena_round(16 * (size_t)(n / (HUB_TILE + 1) + 1));    // hub (start, len)
  b += arena_round(sizeof(MergeDesc) * (size_t)desc_bound(n));  // merge descriptors
  b += arena_round(8 * (size_t)std::max<int64_t>(n, 1));      // merge ping-pong keys
  b += arena_round(4 * (size_t)std::max<int64_t>(n, 1));      // and values
  return b;
}

template <typename OffT>
static void bin_impl(cm_ctx *ctx, int64_t nseg, const OffT *segoff, SegBins &b, SegArena *ar)
{
  b.list = (int32_t *)seg_alloc(ctx, ar, sizeof(int32_t) * (nseg > 0 ? nseg : 1));
  const unsigned nb = nblocks(nseg, 256);
  const int64_t nh = (int64_t)SEG_NBINS * nb;
  int32_t *hist = (int32_t *)seg_alloc(ctx, ar, sizeof(int32_t) * (nh + 1));
  int32_t *base = (int32_t *)seg_alloc(ctx, ar, sizeof(int32_t) * (nh + 1));
  launch(ctx, "seg_bin_hist", 0.0, k_bin_hist<OffT>, dim3(nb), dim3(256), 0, nseg, segoff, nb, hist);
  scan_excl_arena(ctx, ar, nh, hist, base);  // base[nh] = number of listed segments
  int64_t *d_off = ctx->d_small + 16;
  launch(ctx, "seg_bin_scatter", 0.0, k_bin_scatter<OffT>, dim3(nb), dim3(256), 0, nseg, segoff, nb,
         (const int32_t *)base, b.list, d_off);
  CM_HIP(hipMemcpyAsync(ctx->h_small + 16, d_off, sizeof(int64_t) * (SEG_NBINS + 1),
                        hipMemcpyDeviceToHost, ctx->stream));
  sync(ctx);
  for (int c = 0; c <= SEG_NBINS; ++c) b.off[c] = ctx->h_small[16 + c];
  seg_release(ctx, ar, hist);
  seg_release(ctx, ar, base);
}

void bin_segments(cm_ctx *ctx, int64_t nseg, const void *segoff, bool off64, SegBins &b, SegArena *ar)
{
  if (off64) bin_impl(ctx, nseg, (const int64_t *)segoff, b, ar);
  else bin_impl(ctx, nseg, (const int32_t *)segoff, b, ar);
}

void release_bins(cm_ctx *ctx, SegBins &b, SegArena *ar)
{
  if (b.list) seg_release(ctx, ar, b.list);
  b.list = nullptr;
}

static unsigned cap_grid(int64_t count)
{
  return (unsigned)std::min<int64_t>(count, 65535 * 8);
}

template <typename OffT>
static void sort_impl(cm_ctx *ctx, const SegBins &b, const OffT *segoff, int64_t total,
                      uint64_t *keys, uint32_t *vals, int first_bin, SegArena *ar)
{
  const int32_t *L = b.list;
  // bins from tune.seg_wave_bins up to the last wave bin go to the 128-key
  // LDS kernel instead (same result, other launch shape)
  const int wave_bins = ctx->tune.seg_wave_bins;
#define CM_WAVE_BIN(CLS, W)                                                                       \
  {                                                                                               \
    const int64_t c = CLS >= first_bin ? b.off[CLS + 1] - b.off[CLS] : 0;                         \
    if (c && CLS < wave_bins)                                                                     \
      launch(ctx, "segsort_wave", 16.0 * c * W, k_sort_wave<W, OffT>, dim3(nblocks(c * W, 256)),  \
             dim3(256), 0, L + b.off[CLS], c, segoff, keys, vals);                                \
    else if (c)                                                                                   \
      launch(ctx, "segsort_lds", 0.0, k_sort_lds<128, OffT>, dim3(cap_grid(c)), dim3(256), 0,     \
             L + b.off[CLS], c, segoff, keys, vals);                                              \
  }
  CM_WAVE_BIN(0, 2) CM_WAVE_BIN(1, 4) CM_WAVE_BIN(2, 8) CM_WAVE_BIN(3, 16) CM_WAVE_BIN(4, 32)
  CM_WAVE_BIN(5, 64)
#undef CM_WAVE_BIN
#define CM_LDS_BIN(CLS, S)                                                                        \
  {                                                                                               \
    const int64_t c = CLS >= first_bin ? b.off[CLS + 1] - b.off[CLS] : 0;                         \
    if (c) launch(ctx, "segsort_lds", 0.0, k_sort_lds<S, OffT>, dim3(cap_grid(c)), dim3(256), 0,  \
                  L + b.off[CLS], c, segoff, keys, vals);                                         \
  }
  CM_LDS_BIN(6, 128) CM_LDS_BIN(7, 256) CM_LDS_BIN(8, 512) CM_LDS_BIN(9, 1024)
  CM_LDS_BIN(10, 2048) CM_LDS_BIN(11, 4096)
#undef CM_LDS_BIN

  const int64_t nhub = b.off[13] - b.off[12];
  if (!nhub) return;
  // hub rows: fetch (start, len), build tile and merge descriptors on the host
  int64_t *d_info = (int64_t *)seg_alloc(ctx, ar, sizeof(int64_t) * 2 * nhub);
  launch(ctx, "hub_info", 0.0, k_hub_info<OffT>, dim3(nblocks(nhub, 256)), dim3(256), 0,
         L + b.off[12], nhub, segoff, d_info);
  std::vector<int64_t> info(2 * nhub);
  CM_HIP(hipMemcpyAsync(info.data(), d_info, sizeof(int64_t) * 2 * nhub, hipMemcpyDeviceToHost,
                        ctx->stream));
  sync(ctx);
  seg_release(ctx, ar, d_info);
  int64_t maxlen = 0, hubtotal = 0;
  std::vector<MergeDesc> descs;
  for (int64_t h = 0; h < nhub; ++h) {
    const int64_t st = info[2 * h], len = info[2 * h + 1];
    maxlen = std::max(maxlen, len);
    hubtotal += len;
    for (int64_t t = 0; t < len; t += HUB_TILE)
      descs.push_back(MergeDesc{st, (int32_t)std::min<int64_t>(HUB_TILE, len - t), 0, (int32_t)t, 0});
  }
  MergeDesc *d_desc = nullptr;
  size_t desc_cap = 0;
  if (ar) {  // one block of the largest size a pass can need
    desc_cap = (size_t)desc_bound(total);
    d_desc = (MergeDesc *)seg_alloc(ctx, ar, sizeof(MergeDesc) * desc_cap);
  }
  auto upload = [&](const std::vector<MergeDesc> &v) {
    if (v.size() > desc_cap) {
      CM_REQUIRE(!ar, CM_ERR_UNSUPPORTED, "segmented sort: merge descriptors exceed the scratch bound");
      if (d_desc) ctx->pool.release(d_desc);
      desc_cap = v.size();
      d_desc = (MergeDesc *)ctx->pool.alloc(sizeof(MergeDesc) * desc_cap);
    }
    CM_HIP(hipMemcpyAsync(d_desc, v.data(), sizeof(MergeDesc) * v.size(), hipMemcpyHostToDevice,
                          ctx->stream));
    sync(ctx);  // the host vector is reused
  };
  upload(descs);
  launch(ctx, "segsort_hub_tile", 0.0, k_sort_tiles, dim3(cap_grid((int64_t)descs.size())),
         dim3(256), 0, (const MergeDesc *)d_desc, (int64_t)descs.size(), keys, vals);
  uint64_t *k2 = (uint64_t *)seg_alloc(ctx, ar, sizeof(uint64_t) * (size_t)total);
  uint32_t *v2 = (uint32_t *)seg_alloc(ctx, ar, sizeof(uint32_t) * (size_t)total);
  uint64_t *sk = keys, *dk = k2;
  uint32_t *sv = vals, *dv = v2;
  for (int64_t w = HUB_TILE; w < maxlen; w *= 2) {
    descs.clear();
    for (int64_t h = 0; h < nhub; ++h) {
      const int64_t st = info[2 * h], len = info[2 * h + 1];
      for (int64_t p = 0; p < len; p += 2 * w) {
        const int64_t alen = std::min<int64_t>(w, len - p);
        const int64_t blen = std::max<int64_t>(0, std::min<int64_t>(w, len - p - alen));
        for (int64_t d0 = 0; d0 < alen + blen; d0 += HUB_TILE)
          descs.push_back(MergeDesc{st + p, (int32_t)alen, (int32_t)blen, (int32_t)d0, 0});
      }
    }
    upload(descs);
    launch(ctx, "segsort_hub_merge", 24.0 * hubtotal, k_merge, dim3(cap_grid((int64_t)descs.size())),
           dim3(256), 0, (const MergeDesc *)d_desc, (int64_t)descs.size(), (const uint64_t *)sk,
           (const uint32_t *)sv, dk, dv);
    std::swap(sk, dk);
    std::swap(sv, dv);
  }
  if (sk != keys) {  // odd number of passes: copy the hub rows back
    descs.clear();
    for (int64_t h = 0; h < nhub; ++h) {
      const int64_t st = info[2 * h], len = info[2 * h + 1];
      for (int64_t d0 = 0; d0 < len; d0 += HUB_TILE)
        descs.push_back(MergeDesc{st, (int32_t)len, 0, (int32_t)d0, 0});
    }
    upload(descs);
    launch(ctx, "segsort_hub_copy", 24.0 * hubtotal, k_merge, dim3(cap_grid((int64_t)descs.size())),
           dim3(256), 0, (const MergeDesc *)d_desc, (int64_t)descs.size(), (const uint64_t *)sk,
           (const uint32_t *)sv, keys, vals);
  }
  sync(ctx);
  seg_release(ctx, ar, k2);
  seg_release(ctx, ar, v2);
  if (d_desc) seg_release(ctx, ar, d_desc);
}

void sort_binned(cm_ctx *ctx, const SegBins &b, const void *segoff, bool off64, int64_t total,
                 uint64_t *keys, uint32_t *vals, int first_bin, SegArena *ar)
{
  if (off64) sort_impl(ctx, b, (const int64_t *)segoff, total, keys, vals, first_bin, ar);
  else sort_impl(ctx, b, (const int32_t *)segoff, total, keys, vals, first_bin, ar);
}

void segsort_u64(cm_ctx *ctx, int64_t nseg, const void *d_segoff, bool off64, int64_t total,
                 uint64_t *d_keys, uint32_t *d_vals)
{
  SegBins b;
  bin_segments(ctx, nseg, d_segoff, off64, b);
  sort_binned(ctx, b, d_segoff, off64, total, d_keys, d_vals);
  release_bins(ctx, b);
}

// ------------------------------------------- bb_segsort<int,int> interface
// Composite key: order-preserving map of the signed key in the high word,
// position within the segment in the low word.
__global__ __launch_bounds__(256) void k_pack_i32(int32_t n, const int32_t *segoff, int32_t nseg,
                                                  const int32_t *k, const int32_t *v,
                                                  uint64_t *ck, uint32_t *cv)
{
  const int32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  int lo = 0, hi = nseg;  // last segment start <= e
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (segoff[mid] <= e) lo = mid;
    else hi = mid;
  }
  const uint32_t pos = (uint32_t)(e - segoff[lo]);
  ck[e] = ((uint64_t)((uint32_t)k[e] ^ 0x80000000u) << 32) | pos;
  cv[e] = (uint32_t)v[e];
}

__global__ __launch_bounds__(256) void k_unpack_i32(int32_t n, const uint64_t *ck,
                                                    const uint32_t *cv, int32_t *k, int32_t *v)
{
  const int32_t e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  k[e] = (int32_t)((uint32_t)(ck[e] >> 32) ^ 0x80000000u);
  v[e] = (int32_t)cv[e];
}

__global__ void k_make_segoff(int32_t n, const int32_t *segs, int32_t len, int32_t *segoff)
{
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < len) segoff[i] = segs[i];
  if (i == len) segoff[i] = n;
}

void segsort_i32(cm_ctx *ctx, int32_t *keys, int32_t *vals, int32_t n, const int32_t *segs,
                 int32_t length, SegArena *ar)
{
  if (n <= 0 || length <= 0) return;
  int32_t *segoff = (int32_t *)seg_alloc(ctx, ar, sizeof(int32_t) * (length + 1));
  uint64_t *ck = (uint64_t *)seg_alloc(ctx, ar, sizeof(uint64_t) * n);
  uint32_t *cv = (uint32_t *)seg_alloc(ctx, ar, sizeof(uint32_t) * n);
  launch(ctx, "segsort_i32_prep", 0.0, k_make_segoff, dim3(nblocks(length + 1, 256)), dim3(256), 0,
         n, segs, length, segoff);
  launch(ctx, "segsort_i32_pack", 0.0, k_pack_i32, dim3(nblocks(n, 256)), dim3(256), 0, n,
         (const int32_t *)segoff, length, (const int32_t *)keys, (const int32_t *)vals, ck, cv);
  SegBins b;
  bin_segments(ctx, length, segoff, false, b, ar);
  sort_binned(ctx, b, segoff, false, n, ck, cv, 0, ar);
  release_bins(ctx, b, ar);
  launch(ctx, "segsort_i32_unpack", 0.0, k_unpack_i32, dim3(nblocks(n, 256)), dim3(256), 0, n,
         (const uint64_t *)ck, (const uint32_t *)cv, keys, vals);
  sync(ctx);
  seg_release(ctx, ar, segoff);
  seg_release(ctx, ar, ck);
  seg_release(ctx, ar, cv);
}

}  // namespace cm
