// cm_contract_hash.hip -- (c) first-seen contraction of the coarse segments
// above REG_SEG entries, by hashing instead of sorting.
//
// A coarse segment lists the rows of the members of one coarse vertex, lower
// fine id first, as the reference's contraction walks them (cu:766-808).  Its
// first-seen row keeps, for every distinct coarse neighbour, the position of
// the first occurrence and the sum of the weights; the row lists these heads
// in position order, and the head of the coarse vertex itself (edges inside
// the pair) is replaced by the last head (cu:804-808).  The sort-based path
// derives the same rows from a segmented sort; here the heads of a segment
// come from a hash table in LDS (key, first position, weight sum):
//   select  segments above REG_SEG: a chunk list (HC_CHUNK entries per chunk)
//           and items by length -- up to HC_WAVE_MAX one wave, up to
//           HC_BLOCK_MAX one block (a half-size table up to HC_BLOCK_SPLIT,
//           so that twice as many blocks fit a CU), longer ones ("hubs") one
//           block per key
//           partition of about HC_PART entries;
//   gather  per chunk of a hub: coarse neighbour ids in segment order, and
//           the chunk's entries (position, id, weight) bucketed by key
//           partition with a counting sort in LDS (wave and block items read
//           their rows directly);
//   dedup   wave / block items: table, heads ranked by position, the row
//           written compacted into a staging area and its length to cdeg;
//           hub partitions: head flag and weight sum per entry, from the
//           partition's bucket slice of every chunk of the hub;
//   hubs    count per chunk, scan per hub, compacted write per chunk;
//   emit    per chunk, once the row offsets are known: staging -> rows.
// Every item count stays on the device and fixed grids stride over it, so
// the path needs no host synchronisation.  A hub partition whose table fills
// (more than HC_HUB_SLOTS distinct keys, far beyond the expected HC_PART) sets a
// flag; the caller then redoes the level on the sort-based path.
#include "cm_common.hpp"
#include "cm_contract_hash.hpp"
#include "cm_internal.hpp"

namespace cm {

// Hub partitions: one 1024-thread block with a table of HC_HUB_SLOTS slots
// (96 KB of LDS) per partition of about HC_PART entries.  Every partition
// block reads the whole segment, so the segment is read len / HC_PART times:
// the larger the table, the fewer the passes.
constexpr int HC_HUB_LG = 13;
constexpr int HC_HUB_SLOTS = 1 << HC_HUB_LG;  // >= 2 * HC_PART
constexpr int HC_HUB_THREADS = 1024;
constexpr int HC_HUB_LIST = 6144;  // entries of the partition listed for the second pass
constexpr int HC_SELECT = 16;       // segments per thread in the select pass
constexpr int EMIT_BATCH = 8;       // entries per lane per step of the chunk copy passes

__device__ __forceinline__ void wave_sync()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ int ceil_log2(int x)
{
  int l = 0;
  while ((1 << l) < x) ++l;
  return l;
}

// Fibonacci hashing of a coarse id into a table of 1 << lg slots.
__device__ __forceinline__ uint32_t hc_slot0(int32_t key, int lg)
{
  return ((uint32_t)key * 0x9E3779B1u) >> (32 - lg);
}

// Key partition of a hub entry: a second multiplicative hash scaled to P.
__device__ __forceinline__ int hc_part(int32_t key, int P)
{
  return (int)(((uint64_t)((uint32_t)key * 0x85EBCA77u) * (uint32_t)P) >> 32);
}

// Slot of `key` (keys >= 0, -1 = empty), inserting it if absent; -1 when
// the table is full.  Linear probing with compare-and-swap in LDS.
__device__ __forceinline__ int hc_insert(int32_t *tk, int lg, int32_t key)
{
  const uint32_t mask = (1u << lg) - 1;
  uint32_t h = hc_slot0(key, lg);
  for (uint32_t probe = 0; probe <= mask; ++probe) {
    const int32_t old = atomicCAS(&tk[h], -1, key);
    if (old == -1 || old == key) return (int)h;
    h = (h + 1) & mask;
  }
  return -1;
}

// Slot of a key already inserted (the table no longer changes); -1 if absent.
__device__ __forceinline__ int hc_find(const int32_t *tk, int lg, int32_t key)
{
  const uint32_t mask = (1u << lg) - 1;
  uint32_t h = hc_slot0(key, lg);
  for (uint32_t probe = 0; probe <= mask; ++probe) {
    const int32_t k = tk[h];
    if (k == key) return (int)h;
    if (k == -1) return -1;
    h = (h + 1) & mask;
  }
  return -1;
}

// Entry i of segment c (row(v) then row(u), v = rep[c]): its coarse
// neighbour id and weight.  `at` holds the segment's row starts and the
// length of row(v).
struct SegRows {
  int64_t b0, b1;
  int32_t d0;
};

template <typename OffT>
__device__ __forceinline__ SegRows seg_rows(int32_t c, const int32_t *__restrict__ rep,
                                            const int32_t *__restrict__ match, const OffT *__restrict__ xadj)
{
  const int32_t v = rep[c], u = match[v];
  SegRows r;
  r.b0 = (int64_t)xadj[v];
  r.d0 = (int32_t)((int64_t)xadj[v + 1] - r.b0);
  r.b1 = (int64_t)xadj[u];  // used only when u != v (then the segment is longer than row(v))
  return r;
}

// ---------------------------------------------------------------- select
template <typename OffT>
__global__ __launch_bounds__(256) void k_hc_select(const int64_t *__restrict__ ncp, const OffT *__restrict__ segoff,
                                                   HcLists hl)
{
  const int64_t nc = *ncp;
  const int64_t c0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * HC_SELECT;
  // the HC_SELECT + 1 offsets load at once (clamped to nc)
  int64_t so[HC_SELECT + 1];
#pragma unroll
  for (int t = 0; t <= HC_SELECT; ++t) so[t] = (int64_t)segoff[c0 + t < nc ? c0 + t : nc];
  unsigned cnt[HC_N_COUNTERS] = {}, at[HC_N_COUNTERS];
#pragma unroll
  for (int t = 0; t < HC_SELECT; ++t)
    if (c0 + t < nc) hc_count(so[t + 1] - so[t], cnt);
  hc_reserve<4>(cnt, at, hl.ctr);
#pragma unroll
  for (int t = 0; t < HC_SELECT; ++t)
    if (c0 + t < nc) hc_write((int32_t)(c0 + t), so[t + 1] - so[t], at, hl);
}

// ---------------------------------------------------------------- gather
// One block per hub chunk (HC_CHUNK entries, 8 per thread).  Every entry's
// coarse neighbour id goes to cval in segment order.  For hubs of at most
// HC_PMAX partitions the chunk's entries are also counting-sorted by key
// partition into bpos/bkey/bw (at the chunk's own segment positions; bpos
// holds the position inside the chunk, which has HC_CHUNK <= 2^16 entries), with
// the partition starts in coff[q * (HC_PMAX + 1) + p] (coff[.. + P] = the
// chunk's length); the order inside a partition's slice is arbitrary (the
// partition pass takes minima and sums).  Longer hubs keep their weights in
// segment order (wv) for the whole-segment pass.  With unit weights
// (adjwgt == nullptr) neither weight copy exists.
template <typename OffT>
__global__ __launch_bounds__(256) void k_hc_gather(const SegChunk *__restrict__ hchunks, const unsigned *ctr,
                                                   const OffT *__restrict__ segoff, const int32_t *__restrict__ rep,
                                                   const int32_t *__restrict__ match, const OffT *__restrict__ xadj,
                                                   const int32_t *__restrict__ adjncy,
                                                   const int32_t *__restrict__ adjwgt,
                                                   const int32_t *__restrict__ cmap, int32_t *__restrict__ cval,
                                                   int32_t *__restrict__ wv, uint16_t *__restrict__ bpos,
                                                   int32_t *__restrict__ bkey, int32_t *__restrict__ bw,
                                                   uint16_t *__restrict__ coff)
{
  constexpr int PER = HC_CHUNK / 256;
  __shared__ unsigned s_cnt[HC_PMAX + 1];
  __shared__ unsigned s_sc[4];
  const unsigned nq = ctr[HC_N_HCHUNK];
  const int tid = threadIdx.x;
  for (unsigned q = blockIdx.x; q < nq; q += gridDim.x) {
    const SegChunk ch = hchunks[q];
    const int64_t s0 = (int64_t)segoff[ch.c];
    const int64_t len = (int64_t)segoff[ch.c + 1] - s0;
    const int P = (int)((len + HC_PART - 1) / HC_PART);
    const bool bucket = P <= HC_PMAX;
    const int32_t v = rep[ch.c], u = match[v];
    const int64_t b0 = (int64_t)xadj[v], d0 = (int64_t)xadj[v + 1] - b0;
    const int64_t b1 = u != v ? (int64_t)xadj[u] : 0;
    const int64_t i0 = (int64_t)ch.k * HC_CHUNK;
    const int cn = (int)(len - i0 < HC_CHUNK ? len - i0 : HC_CHUNK);
    // clamped, unconditional loads: all issue before the first wait
    int32_t a[PER], w[PER], cv[PER];
#pragma unroll
    for (int t = 0; t < PER; ++t) {
      const int i = t * 256 + tid, ii = i < cn ? i : cn - 1;
      const int64_t e = i0 + ii;
      const int64_t j = e < d0 ? b0 + e : b1 + (e - d0);
      a[t] = adjncy[j];
      w[t] = edge_w(adjwgt, j);
    }
#pragma unroll
    for (int t = 0; t < PER; ++t) cv[t] = cmap[a[t]];
#pragma unroll
    for (int t = 0; t < PER; ++t) {
      const int i = t * 256 + tid;
      if (i < cn) {
        cval[s0 + i0 + i] = cv[t];
        if (!bucket && wv) wv[s0 + i0 + i] = w[t];
      }
    }
    if (!bucket) continue;  // uniform over the block
    for (int p = tid; p <= P; p += 256) s_cnt[p] = 0u;
    __syncthreads();
    int pp[PER];
    unsigned rk[PER];
#pragma unroll
    for (int t = 0; t < PER; ++t) {
      const int i = t * 256 + tid;
      pp[t] = hc_part(cv[t], P);
      rk[t] = i < cn ? atomicAdd(&s_cnt[pp[t]], 1u) : 0u;
    }
    __syncthreads();
    unsigned tot;
    const unsigned ex = block_excl_scan<unsigned, 4>(tid < P ? s_cnt[tid] : 0u, tot, s_sc);
    uint16_t *co = coff + (size_t)q * (HC_PMAX + 1);
    if (tid < P) {
      co[tid] = (uint16_t)ex;
      s_cnt[tid] = ex;
    }
    if (tid == 0) co[P] = (uint16_t)cn;
    __syncthreads();
#pragma unroll
    for (int t = 0; t < PER; ++t) {
      const int i = t * 256 + tid;
      if (i < cn) {
        const int64_t o = s0 + i0 + (int64_t)(s_cnt[pp[t]] + rk[t]);
        bpos[o] = (uint16_t)i;
        bkey[o] = cv[t];
        if (bw) bw[o] = w[t];  // none for unit weights
      }
    }
    __syncthreads();  // s_cnt is reused by the next chunk
  }
}

// ------------------------------------------------------ dedup: one wave
// Segments of up to 2^(LG-1) entries (above REG_SEG), one per wave, a table
// of up to 2^LG slots per wave; *count items.
template <typename OffT, int LG>
__global__ __launch_bounds__(256) void k_hc_dedup_wave(const int32_t *__restrict__ items, const unsigned *count,
                                                       const OffT *__restrict__ segoff,
                                                       const int32_t *__restrict__ rep,
                                                       const int32_t *__restrict__ match,
                                                       const OffT *__restrict__ xadj,
                                                       const int32_t *__restrict__ adjncy,
                                                       const int32_t *__restrict__ adjwgt,
                                                       const int32_t *__restrict__ cmap, int32_t *__restrict__ tc,
                                                       int32_t *__restrict__ tw, int32_t *__restrict__ cdeg,
                                                       unsigned *ovf)
{
  constexpr int SLOTS = 1 << LG;
  constexpr int PER = SLOTS / 2 / CM_WAVE;
  __shared__ int32_t s_key[4][SLOTS], s_pos[4][SLOTS], s_sum[4][SLOTS];
  const int w = threadIdx.x >> 6, l = lane();
  int32_t *tk = s_key[w], *tp = s_pos[w], *ts = s_sum[w];
  const unsigned n = *count;
  const int64_t nw = (int64_t)gridDim.x * 4;
  for (int64_t q = (int64_t)blockIdx.x * 4 + w; q < n; q += nw) {
    const int32_t c = items[q];
    const int64_t s0 = (int64_t)segoff[c];
    const int len = (int)((int64_t)segoff[c + 1] - s0);
    int lg = ceil_log2(2 * len);
    lg = lg < 6 ? 6 : lg;
    for (int i = l; i < (1 << lg); i += CM_WAVE) {
      tk[i] = -1;
      tp[i] = 0x7fffffff;
      ts[i] = 0;
    }
    wave_sync();
    const SegRows sr = seg_rows<OffT>(c, rep, match, xadj);
    int32_t cu[PER], sl[PER], wt[PER], a[PER];
#pragma unroll
    for (int t = 0; t < PER; ++t) {  // clamped, unconditional loads
      const int i = l + t * CM_WAVE < len ? l + t * CM_WAVE : len - 1;
      const int64_t j = i < sr.d0 ? sr.b0 + i : sr.b1 + (i - sr.d0);
      a[t] = adjncy[j];
      wt[t] = edge_w(adjwgt, j);
    }
#pragma unroll
    for (int t = 0; t < PER; ++t) {
      const int i = l + t * CM_WAVE;
      const int32_t x = cmap[a[t]];
      cu[t] = i < len ? x : -1;
      sl[t] = -1;
      if (i < len) {
        sl[t] = hc_insert(tk, lg, cu[t]);
        atomicMin(&tp[sl[t]], i);
        atomicAdd(&ts[sl[t]], wt[t]);
      }
    }
    wave_sync();
    int running = 0, self_rank = -1, rk[PER];
    bool hd[PER];
#pragma unroll
    for (int t = 0; t < PER; ++t) {
      const int i = l + t * CM_WAVE;
      hd[t] = i < len && tp[sl[t]] == i;
      const uint64_t mk = __ballot(hd[t]);
      rk[t] = running + mask_rank(mk);
      const uint64_t sm = __ballot(hd[t] && cu[t] == c);
      if (sm) self_rank = __shfl(rk[t], __ffsll((unsigned long long)sm) - 1, CM_WAVE);
      running += (int)__popcll(mk);
    }
#pragma unroll
    for (int t = 0; t < PER; ++t) {
      if (hd[t] && cu[t] != c) {
        int r = rk[t];
        if (self_rank >= 0 && r == running - 1) r = self_rank;
        hc_stage(tc, tw, s0 + r, cu[t], ts[sl[t]], ovf);
      }
    }
    if (l == 0) cdeg[c] = running - (self_rank >= 0 ? 1 : 0);
    wave_sync();
  }
}

// ----------------------------------------------------- dedup: one block
// Segments of up to 2^(LG-1) entries (above HC_WAVE_MAX), one per block, a
// table of 2^LG slots; *count items.
template <typename OffT, int LG>
__global__ __launch_bounds__(256) void k_hc_dedup_block(const int32_t *__restrict__ items, const unsigned *count,
                                                        const OffT *__restrict__ segoff,
                                                        const int32_t *__restrict__ rep,
                                                        const int32_t *__restrict__ match,
                                                        const OffT *__restrict__ xadj,
                                                        const int32_t *__restrict__ adjncy,
                                                        const int32_t *__restrict__ adjwgt,
                                                        const int32_t *__restrict__ cmap, int32_t *__restrict__ tc,
                                                        int32_t *__restrict__ tw, int32_t *__restrict__ cdeg,
                                                        unsigned *ovf)
{
  constexpr int SLOTS = 1 << LG;
  constexpr int PER = SLOTS / 2 / 256;
  __shared__ int32_t tk[SLOTS], tp[SLOTS], ts[SLOTS];
  __shared__ int sh[4];
  __shared__ int s_self;
  const unsigned n = *count;
  for (int64_t q = blockIdx.x; q < n; q += gridDim.x) {
    const int32_t c = items[q];
    const int64_t s0 = (int64_t)segoff[c];
    const int len = (int)((int64_t)segoff[c + 1] - s0);
    const int lg = ceil_log2(2 * len);
    for (int i = threadIdx.x; i < (1 << lg); i += 256) {
      tk[i] = -1;
      tp[i] = 0x7fffffff;
      ts[i] = 0;
    }
    if (threadIdx.x == 0) s_self = -1;
    __syncthreads();
    const SegRows sr = seg_rows<OffT>(c, rep, match, xadj);
    int32_t cu[PER], sl[PER], wt[PER], a[PER];
#pragma unroll
    for (int t = 0; t < PER; ++t) {  // clamped, unconditional loads
      const int i = (int)threadIdx.x + t * 256 < len ? (int)threadIdx.x + t * 256 : len - 1;
      const int64_t j = i < sr.d0 ? sr.b0 + i : sr.b1 + (i - sr.d0);
      a[t] = adjncy[j];
      wt[t] = edge_w(adjwgt, j);
    }
#pragma unroll
    for (int t = 0; t < PER; ++t) {
      const int i = threadIdx.x + t * 256;
      const int32_t x = cmap[a[t]];
      cu[t] = i < len ? x : -1;
      sl[t] = -1;
      if (i < len) {
        sl[t] = hc_insert(tk, lg, cu[t]);
        atomicMin(&tp[sl[t]], i);
        atomicAdd(&ts[sl[t]], wt[t]);
      }
    }
    __syncthreads();
    int running = 0, rk[PER];
    bool hd[PER];
    const int steps = (len + 255) / 256;  // block-uniform
#pragma unroll
    for (int t = 0; t < PER; ++t) {
      hd[t] = false;
      rk[t] = 0;
      if (t < steps) {
        const int i = threadIdx.x + t * 256;
        hd[t] = i < len && tp[sl[t]] == i;
        int tot;
        rk[t] = running + block256_excl_scan<int>(hd[t] ? 1 : 0, tot, sh);
        if (hd[t] && cu[t] == c) s_self = rk[t];
        running += tot;
      }
    }
    __syncthreads();
    const int self_rank = s_self;
#pragma unroll
    for (int t = 0; t < PER; ++t) {
      if (hd[t] && cu[t] != c) {
        int r = rk[t];
        if (self_rank >= 0 && r == running - 1) r = self_rank;
        hc_stage(tc, tw, s0 + r, cu[t], ts[sl[t]], ovf);
      }
    }
    if (threadIdx.x == 0) cdeg[c] = running - (self_rank >= 0 ? 1 : 0);
    __syncthreads();
  }
}

// Head flag and weight sum of a hub entry.  Unit weights (hw == nullptr):
// the head byte holds the sum itself (at most 4 without repeated edges; a
// larger one sets the overflow word and the level is redone).
__device__ __forceinline__ void hub_mark(uint8_t *head, int32_t *hw, int64_t at, bool h, int32_t sum,
                                         unsigned *overflow)
{
  if (hw) {
    head[at] = h ? 1 : 0;
    hw[at] = h ? sum : 0;
  } else {
    if (h && sum > 4) *overflow = 1u;
    head[at] = h ? (uint8_t)(sum > 4 ? 4 : sum) : 0;
  }
}

// ------------------------------------------------- dedup: hub partitions
// One block per key partition of a hub segment: the block tables the keys of
// its partition; each entry of the partition gets its head flag and, for
// heads, the weight sum.  Bucketed hubs (P <= HC_PMAX) give the block its
// partition's slice of every chunk; longer hubs make it read the whole
// segment.  The first pass also lists the partition's entries (position <<
// 12 | slot) in LDS, so the second pass walks that list instead (lists that
// overflow, and segments of 2^19 entries or more, read the slices or the
// segment again).
template <typename OffT>
__global__ __launch_bounds__(HC_HUB_THREADS) void k_hc_dedup_hub(const HubPart *__restrict__ parts,
                                                                 const unsigned *ctr,
                                                                 const OffT *__restrict__ segoff,
                                                                 const int32_t *__restrict__ cval,
                                                                 const int32_t *__restrict__ wv,
                                                                 const uint16_t *__restrict__ bpos,
                                                                 const int32_t *__restrict__ bkey,
                                                                 const int32_t *__restrict__ bw,
                                                                 const uint16_t *__restrict__ coff,
                                                                 uint8_t *__restrict__ head,
                                                                 int32_t *__restrict__ hw, unsigned *overflow)
{
  constexpr int lg = HC_HUB_LG;
  constexpr int SCAN = 8;
  constexpr int T = HC_HUB_THREADS;
  constexpr int NWV = T / CM_WAVE;
  __shared__ int32_t tk[HC_HUB_SLOTS], tp[HC_HUB_SLOTS], ts[HC_HUB_SLOTS];
  __shared__ uint32_t lst[HC_HUB_LIST];
  // bucketed hubs: the partition's slice of each chunk (start, and the
  // running total of the slice lengths before it)
  constexpr int MAXCH = HC_PMAX * HC_PART / HC_CHUNK + 1;
  __shared__ uint16_t s_sb[MAXCH];
  __shared__ uint32_t s_so[MAXCH + 1];
  __shared__ uint32_t s_sc[NWV];
  __shared__ int s_full;
  __shared__ unsigned s_cnt;
  static_assert(MAXCH <= T, "one chunk per thread in the slice scan");
  const unsigned n = ctr[HC_N_PART];
  const int l = lane(), wv_id = threadIdx.x / CM_WAVE;
  for (int64_t q = blockIdx.x; q < n; q += gridDim.x) {
    const HubPart hp = parts[q];
    const int64_t s0 = (int64_t)segoff[hp.c];
    const int64_t len = (int64_t)segoff[hp.c + 1] - s0;
    const bool bucket = hp.P <= HC_PMAX;
    const int nch = (int)((len + HC_CHUNK - 1) / HC_CHUNK);
    for (int i = threadIdx.x; i < HC_HUB_SLOTS; i += T) {
      tk[i] = -1;
      tp[i] = 0x7fffffff;
      ts[i] = 0;
    }
    if (threadIdx.x == 0) {
      s_full = 0;
      s_cnt = 0;
    }
    __syncthreads();
    // list entries pack (position << lg) | slot into 32 bits
    const bool listable = len < (1ll << (32 - lg));
    // one entry of the partition: table it and list it (called by whole
    // waves; lanes without an entry pass ok = false)
    auto take = [&](bool ok, int32_t key, int32_t pos, int32_t wt) {
      int s = -1;
      if (ok) {
        s = hc_insert(tk, lg, key);
        if (s < 0) {
          s_full = 1;
        } else {
          atomicMin(&tp[s], pos);
          atomicAdd(&ts[s], wt);
        }
      }
      // one LDS counter update per wave
      const uint64_t mk = __ballot(s >= 0);
      unsigned base = 0;
      if (mk && l == 0) base = atomicAdd(&s_cnt, (unsigned)__popcll(mk));
      base = __shfl(base, 0, CM_WAVE);
      const unsigned k = base + (unsigned)mask_rank(mk);
      if (s >= 0 && k < (unsigned)HC_HUB_LIST) lst[k] = ((uint32_t)pos << lg) | (uint32_t)s;
    };
    if (bucket) {
      // the slices laid end to end, so that all threads share them whatever
      // their number and lengths (a hub of a few chunks has slices of up to
      // HC_CHUNK entries)
      uint32_t sl = 0;
      if (threadIdx.x < nch) {
        const uint16_t *co = coff + (size_t)(hp.hfirst + threadIdx.x) * (HC_PMAX + 1);
        const uint16_t b = co[hp.p];
        s_sb[threadIdx.x] = b;
        sl = (uint32_t)co[hp.p + 1] - b;
      }
      uint32_t total;
      const uint32_t ex = block_excl_scan<uint32_t, NWV>(sl, total, s_sc);
      if (threadIdx.x < nch) s_so[threadIdx.x] = ex;
      if (threadIdx.x == 0) s_so[nch] = total;
      __syncthreads();
      for (uint32_t f0 = 0; f0 < total; f0 += T) {
        const uint32_t f = f0 + threadIdx.x;
        const bool ok = f < total;
        // the chunk whose slice holds entry f: the last j with s_so[j] <= f
        int lo = 0, hi = nch - 1;
        while (lo < hi) {
          const int mid = (lo + hi + 1) >> 1;
          if (s_so[mid] <= f) lo = mid;
          else hi = mid - 1;
        }
        const int64_t o = ok ? s0 + (int64_t)lo * HC_CHUNK + s_sb[lo] + (f - s_so[lo]) : s0;
        take(ok, bkey[o], lo * HC_CHUNK + (int32_t)bpos[o], bw ? bw[o] : 1);
      }
    } else {
      for (int64_t ib = threadIdx.x; ib < len; ib += SCAN * T) {
        int32_t cu[SCAN];
#pragma unroll
        for (int t = 0; t < SCAN; ++t) {
          const int64_t i = ib + (int64_t)t * T;
          cu[t] = i < len ? cval[s0 + i] : -1;
        }
#pragma unroll
        for (int t = 0; t < SCAN; ++t) {
          const int64_t i = ib + (int64_t)t * T;
          const bool ok = cu[t] >= 0 && hc_part(cu[t], hp.P) == hp.p;
          take(ok, cu[t], (int32_t)i, ok ? (wv ? wv[s0 + i] : 1) : 0);
        }
      }
    }
    __syncthreads();
    if (s_full && threadIdx.x == 0) atomicOr(overflow, 1u);
    if (listable && s_cnt <= (unsigned)HC_HUB_LIST) {
      for (unsigned k = threadIdx.x; k < s_cnt; k += T) {
        const uint32_t e = lst[k];
        const int32_t i = (int32_t)(e >> lg);
        const int s = (int)(e & (uint32_t)(HC_HUB_SLOTS - 1));
        const bool h = tp[s] == i;
        hub_mark(head, hw, s0 + i, h, ts[s], overflow);
      }
    } else if (bucket) {
      for (int j = wv_id; j < nch; j += NWV) {
        const uint16_t *co = coff + (size_t)(hp.hfirst + j) * (HC_PMAX + 1);
        const int b = co[hp.p], e = co[hp.p + 1];
        const int64_t cb = s0 + (int64_t)j * HC_CHUNK;
        for (int i = b + l; i < e; i += CM_WAVE) {
          const int32_t key = bkey[cb + i], pos = j * HC_CHUNK + (int32_t)bpos[cb + i];
          const int s = hc_find(tk, lg, key);
          const bool h = s >= 0 && tp[s] == pos;
          hub_mark(head, hw, s0 + pos, h, h ? ts[s] : 0, overflow);
        }
      }
    } else {
      for (int64_t i = threadIdx.x; i < len; i += T) {
        const int32_t cu = cval[s0 + i];
        if (hc_part(cu, hp.P) != hp.p) continue;
        const int s = hc_find(tk, lg, cu);
        const bool h = s >= 0 && tp[s] == (int32_t)i;
        hub_mark(head, hw, s0 + i, h, h ? ts[s] : 0, overflow);
      }
    }
    __syncthreads();
  }
}

// Hub chunks (the dense hub chunk list): heads per chunk and the rank of the
// self head inside it.
template <typename OffT>
__global__ __launch_bounds__(256) void k_hc_hub_count(const SegChunk *__restrict__ hchunks, const unsigned *ctr,
                                                      const OffT *__restrict__ segoff,
                                                      const uint8_t *__restrict__ head,
                                                      const int32_t *__restrict__ cval, int32_t *__restrict__ ccnt,
                                                      int32_t *__restrict__ cself)
{
  const unsigned nq = ctr[HC_N_HCHUNK];
  const int l = lane();
  const int64_t nw = (int64_t)gridDim.x * (blockDim.x / CM_WAVE);
  for (int64_t q = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / CM_WAVE; q < nq; q += nw) {
    const SegChunk ch = hchunks[q];
    const int64_t s0 = (int64_t)segoff[ch.c];
    const int64_t len = (int64_t)segoff[ch.c + 1] - s0;
    const int64_t i0 = (int64_t)ch.k * HC_CHUNK, i1 = len < i0 + HC_CHUNK ? len : i0 + HC_CHUNK;
    int cnt = 0, self = -1;
    for (int64_t ib = i0; ib < i1; ib += EMIT_BATCH * CM_WAVE) {
      bool hh[EMIT_BATCH];
      int32_t cv[EMIT_BATCH];
#pragma unroll
      for (int t = 0; t < EMIT_BATCH; ++t) {  // clamped loads, issued together
        const int64_t i = ib + (int64_t)t * CM_WAVE + l, ii = i < i1 ? i : i1 - 1;
        const uint8_t hb = head[s0 + ii];
        hh[t] = i < i1 && hb;
        cv[t] = cval[s0 + ii];
      }
#pragma unroll
      for (int t = 0; t < EMIT_BATCH; ++t) {
        const uint64_t mk = __ballot(hh[t]);
        const uint64_t sm = __ballot(hh[t] && cv[t] == ch.c);
        if (sm) self = cnt + (int)__popcll(mk & ((1ull << (__ffsll((unsigned long long)sm) - 1)) - 1));
        cnt += (int)__popcll(mk);
      }
    }
    if (l == 0) {
      ccnt[q] = cnt;
      cself[q] = self;
    }
  }
}

// One wave per hub: chunk offsets, row length, rank of the self head (all
// indexed by the dense hub chunk list).
template <typename OffT>
__global__ __launch_bounds__(256) void k_hc_hub_scan(const int32_t *__restrict__ hubs, const unsigned *ctr,
                                                     const SegChunk *__restrict__ hchunks,
                                                     const int32_t *__restrict__ ccnt,
                                                     const int32_t *__restrict__ cself, int32_t *__restrict__ coff,
                                                     int32_t *__restrict__ hself, int32_t *__restrict__ htot,
                                                     int32_t *__restrict__ cdeg)
{
  const unsigned nh = ctr[HC_N_HUB];
  const int l = lane();
  const int64_t nw = (int64_t)gridDim.x * (blockDim.x / CM_WAVE);
  for (int64_t h = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / CM_WAVE; h < nh; h += nw) {
    const int32_t first = hubs[h];
    const SegChunk c0 = hchunks[first];
    int carry = 0, self = -1;
    for (int k0 = 0; k0 < c0.nch; k0 += CM_WAVE) {
      const int k = k0 + l;
      const int x = k < c0.nch ? ccnt[first + k] : 0;
      const int inc = wave_incl_scan<int>(x);
      const int off = carry + inc - x;
      if (k < c0.nch) {
        coff[first + k] = off;
        const int sr = cself[first + k];
        if (sr >= 0) self = off + sr;
      }
      carry += __shfl(inc, CM_WAVE - 1, CM_WAVE);
    }
    self = wave_max<int>(self);
    if (l == 0) {
      hself[first] = self;
      htot[first] = carry;
      cdeg[c0.c] = carry - (self >= 0 ? 1 : 0);
    }
  }
}

// Hub chunks: heads written compacted in position order, the self head
// dropped and replaced by the last head.
template <typename OffT>
__global__ __launch_bounds__(256) void k_hc_hub_write(const SegChunk *__restrict__ hchunks, const unsigned *ctr,
                                                      const OffT *__restrict__ segoff,
                                                      const uint8_t *__restrict__ head,
                                                      const int32_t *__restrict__ cval, const int32_t *__restrict__ hw,
                                                      const int32_t *__restrict__ coff,
                                                      const int32_t *__restrict__ hself,
                                                      const int32_t *__restrict__ htot, int32_t *__restrict__ tc,
                                                      int32_t *__restrict__ tw, unsigned *ovf)
{
  const unsigned nq = ctr[HC_N_HCHUNK];
  const int l = lane();
  const int64_t nw = (int64_t)gridDim.x * (blockDim.x / CM_WAVE);
  for (int64_t q = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / CM_WAVE; q < nq; q += nw) {
    const SegChunk ch = hchunks[q];
    const int64_t s0 = (int64_t)segoff[ch.c];
    const int64_t len = (int64_t)segoff[ch.c + 1] - s0;
    const int self = hself[ch.first], tot = htot[ch.first];
    const int64_t i0 = (int64_t)ch.k * HC_CHUNK, i1 = len < i0 + HC_CHUNK ? len : i0 + HC_CHUNK;
    int r0 = coff[q];
    for (int64_t ib = i0; ib < i1; ib += EMIT_BATCH * CM_WAVE) {
      bool hh[EMIT_BATCH];
      int32_t cv[EMIT_BATCH], hv[EMIT_BATCH];
#pragma unroll
      for (int t = 0; t < EMIT_BATCH; ++t) {  // clamped loads, issued together
        const int64_t i = ib + (int64_t)t * CM_WAVE + l, ii = i < i1 ? i : i1 - 1;
        const uint8_t hb = head[s0 + ii];
        hh[t] = i < i1 && hb;
        cv[t] = cval[s0 + ii];
        hv[t] = hw ? hw[s0 + ii] : (int32_t)hb;  // unit weights: the head byte holds the sum
      }
#pragma unroll
      for (int t = 0; t < EMIT_BATCH; ++t) {
        const uint64_t mk = __ballot(hh[t]);
        if (hh[t] && cv[t] != ch.c) {
          int r = r0 + mask_rank(mk);
          if (self >= 0 && r == tot - 1) r = self;
          hc_stage(tc, tw, s0 + r, cv[t], hv[t], ovf);
        }
        r0 += (int)__popcll(mk);
      }
    }
  }
}

// ------------------------------------------------------------------ emit
// PACKED: the staging holds one word per entry (tw unused).
template <typename OffT, bool PACKED>
__global__ __launch_bounds__(256) void k_hc_emit(const SegChunk *__restrict__ chunks, const unsigned *ctr,
                                                 const OffT *__restrict__ segoff, const int32_t *__restrict__ cdeg,
                                                 const OffT *__restrict__ cxadj, const int32_t *__restrict__ tc,
                                                 const int32_t *__restrict__ tw, int32_t *__restrict__ cadjncy,
                                                 int32_t *__restrict__ cadjwgt)
{
  const unsigned nq = ctr[HC_N_CHUNK];
  const int l = lane();
  const int64_t nw = (int64_t)gridDim.x * (blockDim.x / CM_WAVE);
  for (int64_t q = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / CM_WAVE; q < nq; q += nw) {
    const SegChunk ch = chunks[q];
    const int64_t s0 = (int64_t)segoff[ch.c], o0 = (int64_t)cxadj[ch.c];
    const int64_t d = cdeg[ch.c];
    const int64_t i0 = (int64_t)ch.k * HC_CHUNK, i1 = d < i0 + HC_CHUNK ? d : i0 + HC_CHUNK;
    // EMIT_BATCH entries per lane per step: all loads issue before the stores
    for (int64_t ib = i0 + l; ib < i1; ib += EMIT_BATCH * CM_WAVE) {
      int32_t a[EMIT_BATCH], w[EMIT_BATCH];
#pragma unroll
      for (int t = 0; t < EMIT_BATCH; ++t) {
        const int64_t i = ib + (int64_t)t * CM_WAVE, ii = i < i1 ? i : i1 - 1;
        a[t] = tc[s0 + ii];
        if constexpr (!PACKED) w[t] = tw[s0 + ii];
      }
#pragma unroll
      for (int t = 0; t < EMIT_BATCH; ++t) {
        const int64_t i = ib + (int64_t)t * CM_WAVE;
        if (i < i1) {
          if constexpr (PACKED) {
            cadjncy[o0 + i] = hc_unpack_id(a[t]);
            cadjwgt[o0 + i] = hc_unpack_w(a[t]);
          } else {
            cadjncy[o0 + i] = a[t];
            cadjwgt[o0 + i] = w[t];
          }
        }
      }
    }
  }
}

// ------------------------------------------------------------- host side
// bounds of the item counts for nc coarse vertices (segments above REG_SEG
// hold at most m entries)
struct HcBounds {
  int64_t nseg, nchunk, nwavel, nblk, nblkl, nhub, npart, nhchunk;
  HcBounds(int64_t nc, int64_t m)
  {
    nseg = std::min<int64_t>(nc, m / (REG_SEG + 1)) + 1;
    nchunk = m / HC_CHUNK + nseg;
    nwavel = std::min<int64_t>(nc, m / (HC_WAVE_SPLIT + 1)) + 1;
    nblk = std::min<int64_t>(nc, m / (HC_WAVE_MAX + 1)) + 1;
    nblkl = std::min<int64_t>(nc, m / (HC_BLOCK_SPLIT + 1)) + 1;
    nhub = m / (HC_BLOCK_MAX + 1) + 1;
    npart = m / HC_PART + nhub;
    nhchunk = m / HC_CHUNK + nhub;
  }
};

bool hash_needed(const Level &L) { return L.maxdeg < 0 || 2 * L.maxdeg > REG_SEG; }

void hash_alloc(cm_ctx *ctx, const Level &L, int64_t nc, HashContract &hc)
{
  const int64_t m = L.m;
  const HcBounds b(nc, m);
  auto al = [&](size_t bytes) {
    void *p = ctx->pool.alloc(bytes ? bytes : 1);
    hc.blocks.push_back(p);
    return p;
  };
  hc.nc_bound = nc;
  hc.lists.chunks = (SegChunk *)al(sizeof(SegChunk) * b.nchunk);
  hc.lists.wave = (int32_t *)al(4 * b.nseg);
  hc.lists.wavel = (int32_t *)al(4 * b.nwavel);
  hc.lists.block = (int32_t *)al(4 * b.nblk);
  hc.lists.blockl = (int32_t *)al(4 * b.nblkl);
  hc.lists.parts = (HubPart *)al(sizeof(HubPart) * b.npart);
  hc.lists.hubs = (int32_t *)al(4 * b.nhub);
  hc.lists.hchunks = (SegChunk *)al(sizeof(SegChunk) * b.nhchunk);
  static_assert(HASH_CTR_SLOT + HC_CTR_WORDS <= CTR_SLOTS, "hash counters");
  hc.lists.ctr = ctx->d_ctr + HASH_CTR_SLOT;
  hc.tc = (int32_t *)al(4 * m);
  // unit weights (no weight array) and coarse ids below HC_PACK_MAX_ID:
  // packed staging, one word per entry
  const bool unit = !L.adjwgt || L.adjwgt == ctx->ones;  // the shared array of ones is never written
  hc.tw = !unit || nc >= HC_PACK_MAX_ID ? (int32_t *)al(4 * m) : nullptr;
}

template <typename OffT>
void hash_prepare_impl(cm_ctx *ctx, const Level &L, const int32_t *match, const int32_t *cmap, int64_t nc,
                       const int64_t *d_nc, const int32_t *rep, const OffT *segoff, int32_t *cdeg,
                       int64_t seg_max, HashContract &hc)
{
  hc.seg_max = seg_max;
  if (seg_max < 0) seg_max = INT64_MAX;
  const int64_t m = L.m;
  const OffT *xadj = (const OffT *)L.xadj;
  if (!hc.lists.chunks) hash_alloc(ctx, L, nc, hc);
  nc = hc.nc_bound;
  const HcBounds b(nc, m);
  const int64_t nseg = b.nseg, nchunk = b.nchunk, nwavel = b.nwavel, nblk = b.nblk, nblkl = b.nblkl,
                nhub = b.nhub, npart = b.npart;
  auto al = [&](size_t bytes) {
    void *p = ctx->pool.alloc(bytes ? bytes : 1);
    hc.blocks.push_back(p);
    return p;
  };
  const HcLists &hl = hc.lists;
  unsigned *ctr = hl.ctr;
  if (!hc.selected) {
    if (!ctx->hash_ctr_zero) fill_u32(ctx, ctr, 0u, HC_CTR_WORDS);
    ctx->hash_ctr_zero = false;
    launch(ctx, "hc_select", 0.0, k_hc_select<OffT>, dim3(nblocks(nc, 256 * HC_SELECT)), dim3(256), 0, d_nc,
           segoff, hl);
  }
  const int32_t *wave_items = hl.wave, *wavel_items = hl.wavel, *block_items = hl.block, *blockl_items = hl.blockl,
                *hubs = hl.hubs;
  const HubPart *parts = hl.parts;
  const unsigned gchunk = (unsigned)std::min<int64_t>(4096, (nchunk + 3) / 4);
  hc.gchunk = gchunk;
  static_assert(HC_CHUNK <= 1 << 16, "chunk positions in 16 bits");
  static_assert(2 * HC_WAVE_SPLIT == 1 << 7 && 2 * HC_WAVE_MAX == 1 << 9, "wave table sizes");
  // the wave passes stride over their device-side item counts: a small
  // level (whose counts are far below the bounds) dispatches fewer idle
  // blocks, a large one up to 4096
  const int64_t gwave = std::min<int64_t>(4096, std::max<int64_t>(256, m / 2048));
  launch(ctx, "hc_dedup_wave", 0.0, k_hc_dedup_wave<OffT, 7>,
         dim3((unsigned)std::min<int64_t>(gwave, (nseg + 3) / 4)), dim3(256), 0, (const int32_t *)wave_items,
         (const unsigned *)ctr + HC_N_WAVE, segoff, rep, match, xadj, L.adjncy, L.adjwgt, cmap, hc.tc, hc.tw,
         cdeg, ctr + HC_OVERFLOW);
  if (seg_max > HC_WAVE_SPLIT)
    launch(ctx, "hc_dedup_wave_l", 0.0, k_hc_dedup_wave<OffT, 9>,
           dim3((unsigned)std::min<int64_t>(gwave, (nwavel + 3) / 4)), dim3(256), 0, (const int32_t *)wavel_items,
           (const unsigned *)ctr + HC_N_WAVEL, segoff, rep, match, xadj, L.adjncy, L.adjwgt, cmap, hc.tc, hc.tw,
           cdeg, ctr + HC_OVERFLOW);
  static_assert(2 * HC_BLOCK_SPLIT == 1 << 11 && 2 * HC_BLOCK_MAX == 1 << 12, "block table sizes");
  if (seg_max > HC_WAVE_MAX)
    launch(ctx, "hc_dedup_block", 0.0, k_hc_dedup_block<OffT, 11>, dim3((unsigned)std::min<int64_t>(4096, nblk)),
           dim3(256), 0, (const int32_t *)block_items, (const unsigned *)ctr + HC_N_BLOCK, segoff, rep, match,
           xadj, L.adjncy, L.adjwgt, cmap, hc.tc, hc.tw, cdeg, ctr + HC_OVERFLOW);
  if (seg_max > HC_BLOCK_SPLIT)
    launch(ctx, "hc_dedup_block_l", 0.0, k_hc_dedup_block<OffT, 12>,
           dim3((unsigned)std::min<int64_t>(2048, nblkl)), dim3(256), 0, (const int32_t *)blockl_items,
           (const unsigned *)ctr + HC_N_BLOCKL, segoff, rep, match, xadj, L.adjncy, L.adjwgt, cmap, hc.tc, hc.tw,
           cdeg, ctr + HC_OVERFLOW);
  if (m > HC_BLOCK_MAX && seg_max > HC_BLOCK_MAX) {  // a hub needs more than HC_BLOCK_MAX entries
    // unit weights (adjwgt == nullptr): no weight copies; the dedup counts 1 per entry
    int32_t *cval = (int32_t *)al(4 * m), *wv = L.adjwgt ? (int32_t *)al(4 * m) : nullptr;
    uint16_t *bpos = (uint16_t *)al(2 * m);
    int32_t *bkey = (int32_t *)al(4 * m);
    int32_t *bw = L.adjwgt ? (int32_t *)al(4 * m) : nullptr;
    uint16_t *coff_b = (uint16_t *)al(sizeof(uint16_t) * (HC_PMAX + 1) * b.nhchunk);
    launch(ctx, "hc_gather", 0.0, k_hc_gather<OffT>, dim3((unsigned)std::min<int64_t>(8192, b.nhchunk)), dim3(256),
           0, (const SegChunk *)hl.hchunks, (const unsigned *)ctr, segoff, rep, match, xadj, L.adjncy, L.adjwgt, cmap,
           cval, wv, bpos, bkey, bw, coff_b);
    uint8_t *head = (uint8_t *)al(m);
    int32_t *hw = L.adjwgt ? (int32_t *)al(4 * m) : nullptr;  // unit weights: sums in the head bytes
    // per hub chunk (indexed by the dense hub chunk list)
    const int64_t nhc = b.nhchunk;
    int32_t *ccnt = (int32_t *)al(4 * nhc), *cself = (int32_t *)al(4 * nhc), *coff = (int32_t *)al(4 * nhc);
    int32_t *hself = (int32_t *)al(4 * nhc), *htot = (int32_t *)al(4 * nhc);
    const unsigned ghc = (unsigned)std::min<int64_t>(4096, (nhc + 3) / 4);
    launch(ctx, "hc_dedup_hub", 0.0, k_hc_dedup_hub<OffT>, dim3((unsigned)std::min<int64_t>(1024, npart)),
           dim3(HC_HUB_THREADS), 0, (const HubPart *)parts, (const unsigned *)ctr, segoff, (const int32_t *)cval,
           (const int32_t *)wv, (const uint16_t *)bpos, (const int32_t *)bkey, (const int32_t *)bw,
           (const uint16_t *)coff_b, head, hw, ctr + HC_OVERFLOW);
    launch(ctx, "hc_hub_count", 0.0, k_hc_hub_count<OffT>, dim3(ghc), dim3(256), 0, (const SegChunk *)hl.hchunks,
           (const unsigned *)ctr, segoff, (const uint8_t *)head, (const int32_t *)cval, ccnt, cself);
    launch(ctx, "hc_hub_scan", 0.0, k_hc_hub_scan<OffT>, dim3((unsigned)std::min<int64_t>(1024, (nhub + 3) / 4)),
           dim3(256), 0, (const int32_t *)hubs, (const unsigned *)ctr, (const SegChunk *)hl.hchunks, (const int32_t *)ccnt,
           (const int32_t *)cself, coff, hself, htot, cdeg);
    launch(ctx, "hc_hub_write", 0.0, k_hc_hub_write<OffT>, dim3(ghc), dim3(256), 0, (const SegChunk *)hl.hchunks,
           (const unsigned *)ctr, segoff, (const uint8_t *)head, (const int32_t *)cval, (const int32_t *)hw,
           (const int32_t *)coff, (const int32_t *)hself, (const int32_t *)htot, hc.tc, hc.tw,
           ctr + HC_OVERFLOW);
  }
}

template <typename OffT>
void hash_emit_impl(cm_ctx *ctx, const OffT *segoff, const int32_t *cdeg, const OffT *cxadj, int32_t *cadjncy,
                    int32_t *cadjwgt, const HashContract &hc)
{
  if (hc.seg_max < 0 || hc.seg_max > TILE_COPY_SEG)
    launch(ctx, "hc_emit", 0.0, hc.tw ? k_hc_emit<OffT, false> : k_hc_emit<OffT, true>, dim3(hc.gchunk), dim3(256),
           0, (const SegChunk *)hc.lists.chunks, (const unsigned *)hc.lists.ctr, segoff, cdeg, cxadj,
           (const int32_t *)hc.tc, (const int32_t *)hc.tw, cadjncy, cadjwgt);
}

const unsigned *hash_overflow_flag(const HashContract &hc)
{
  return hc.seg_max < 0 || hc.seg_max > HC_BLOCK_MAX || !hc.tw ? hc.lists.ctr + HC_OVERFLOW : nullptr;
}

template void hash_prepare_impl<int32_t>(cm_ctx *, const Level &, const int32_t *, const int32_t *, int64_t,
                                         const int64_t *, const int32_t *, const int32_t *, int32_t *, int64_t,
                                         HashContract &);
template void hash_prepare_impl<int64_t>(cm_ctx *, const Level &, const int32_t *, const int32_t *, int64_t,
                                         const int64_t *, const int32_t *, const int64_t *, int32_t *, int64_t,
                                         HashContract &);
template void hash_emit_impl<int32_t>(cm_ctx *, const int32_t *, const int32_t *, const int32_t *, int32_t *,
                                      int32_t *, const HashContract &);
template void hash_emit_impl<int64_t>(cm_ctx *, const int64_t *, const int32_t *, const int64_t *, int32_t *,
                                      int32_t *, const HashContract &);

void hash_release(cm_ctx *ctx, HashContract &hc)
{
  for (void *p : hc.blocks) ctx->pool.release(p);
  hc.blocks.clear();
}

}  // namespace cm
