// cm_match.hip -- (a) deterministic seeded handshake heavy-edge matching.
//
// Replaces the reference's matching steps (cu:1699-1747: init cu:1411-1419,
// per-row weight sort cu:1723, racy cuda_shem2 cu:1563-1595, conflict cleanup
// cu:1600-1630).  Synchronous rounds over the round-start state:
//   propose: an unmatched v keeps last round's proposal if that vertex is
//            still unmatched (it is then still v's best), otherwise scans its
//            row for the maximum eligible edge -- unmatched neighbour u with
//            vwgt[v]+vwgt[u] <= maxvwgt -- under the strict total order
//            (adjwgt, hash(seed, min, max), min, max); no candidate => dead.
//   accept:  dead => self;  mutual proposals => matched.
// The heaviest remaining eligible edge is always mutual, so every round makes
// progress and the loop ends with a maximal matching.  Ordering only by
// vertex ids and the seed makes the result independent of launch shape and
// of a graph's position inside a batch.  The fine CSR is never reordered
// (the reference sorts every row in place, cu:1723).
// Launch r does the accepts of round r-1 and the proposals of round r.  On
// large levels with long rows those are two kernels (k_accept, then the
// round kernel reading complete taken bits); elsewhere one fused kernel.
#include "cm_common.hpp"
#include "cm_internal.hpp"

namespace cm {

constexpr int LONG_DEG = LONG_ROW;  // levels with longer rows list their rows by class
constexpr int LONG_CHUNK = 2048;    // entries per wave of a hub row
// On levels with rows longer than LONG_DEG, every row above SHORT_DEG
// entries is listed by class and scanned by a group of lanes: in a
// thread-per-row wave one long row keeps all 64 lanes waiting through its
// dependent steps.  Class k takes rows of (class_lo(k), class_hi(k)] entries,
// class_lanes(k) lanes per row: classes 0..2 scan a row in one step of
// SCAN_BATCH entries per lane, CM_WAVE / class_lanes(k) rows at a time;
// class 3 ("mid" rows, up to LONG_CHUNK entries) scans one row per wave.
// Levels without long rows (meshes) keep every row on the thread-per-row
// path and need no lists.
constexpr int SHORT_DEG = 8;
constexpr int NCLASS = 4;
constexpr int MID_CLASS = 3;
__host__ __device__ constexpr int class_lanes(int k) { return k == 0 ? 4 : k == 1 ? 16 : k == 2 ? 32 : CM_WAVE; }
__host__ __device__ constexpr int class_hi(int k) { return k == MID_CLASS ? LONG_CHUNK : 4 * class_lanes(k); }
__host__ __device__ constexpr int class_lo(int k) { return k == 0 ? SHORT_DEG : class_hi(k - 1); }
static_assert(class_hi(0) == 16 && class_hi(1) == 64 && class_hi(2) == 128, "row classes");
static_assert(class_hi(1) == LONG_DEG, "classes 0..1 cover the rows up to LONG_DEG");

// The row lists of the classes (vertex ids).
struct RowLists {
  int32_t *rows[NCLASS];
};
constexpr int SELECT_PER_THREAD = 16;
constexpr int MID_GROUP = 16;  // rows per wave of the lane-group classes
constexpr int WAVE_GROUP = 16;  // rows per wave of the whole-wave (mid) class
constexpr int ROUND_BATCH = MATCH_BATCH;  // launches queued between host checks
// Row entries per scan step on levels without long rows (8 was slower on
// the grid: a short row then loads more clamped duplicates).
constexpr int MESH_BATCH = 4;
// Levels with long rows and at least this many vertices run their accepts in
// a launch of their own and keep taken bits (a 4 MB match[] is the L2 of one
// XCD; its bit array is 128 KB).
constexpr int64_t BITS_MIN_N = 1 << 20;
constexpr int TK_MATCH = 0, TK_SPLIT = 1;

// flags[2r] = 1 if a vertex matched or died in round r, flags[2r+1] = 1 if
// any vertex is still unmatched after round r.  Round r's accept runs in
// launch r+1, so launch r (r >= 2) runs only if round r-2 was live.
__device__ __forceinline__ bool round_live(const unsigned *flags, int r)
{
  return flags[2 * r] != 0 && flags[2 * r + 1] != 0;
}

// Per-vertex matching parameters.  A single graph uses maxv0/seed0 and
// base 0; a batch (disjoint union of graphs) looks up the owning graph of v
// so that every graph keeps its own weight limit, seed and local vertex ids.
__device__ __forceinline__ void owner_params(const MatchOwners &o, int64_t v, int32_t &maxv,
                                             uint64_t &seed, int64_t &base)
{
  if (o.gid) {
    const int32_t g = o.gid[v];
    maxv = o.maxv[g];
    seed = o.seed[g];
    base = o.vbase[g];
  } else {
    maxv = o.maxv0;
    seed = o.seed0;
    base = 0;
  }
}

__device__ __forceinline__ bool key_better(uint64_t a1, uint64_t a2, uint64_t b1, uint64_t b2)
{
  return a1 > b1 || (a1 == b1 && a2 > b2);
}

// Also clears the hash contraction's item counters (hzero, 16 words) and the
// taken bits (nbw words, at most n).
__global__ __launch_bounds__(256) void k_match_init(int64_t n, int32_t *match, int32_t *prop, unsigned *flags,
                                                    int64_t nflags, unsigned *hzero, uint64_t *bits, int64_t nbw)
{
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v < 16) hzero[v] = 0u;
  if (v < nbw) bits[v] = 0u;
  if (v < n) {
    match[v] = -1;
    prop[v] = -1;
    prop[n + v] = -1;  // both halves: an entry never read back is still a valid value
  }
  for (int64_t i = v; i < nflags; i += (int64_t)gridDim.x * blockDim.x) flags[i] = 0u;
}

// Rows of at most LONG_CHUNK entries are listed by vertex in their class
// (above); longer rows ("hubs") are cut into chunks of LONG_CHUNK entries,
// one wave per chunk.  Chunk descriptor: the row, its first chunk, its chunk
// count and the chunk's index within the row.
struct LongChunk {
  int32_t v, first, nch, k;
};

// count[0] = hub chunks, count[1 + k] = rows of class k.  The order of the
// lists does not matter (the reduction over a row's chunks is exact).
template <typename OffT>
__global__ __launch_bounds__(256) void k_select_long(int64_t n, const OffT *xadj, LongChunk *chunks, RowLists rl,
                                                     unsigned *count)
{
  // SELECT_PER_THREAD consecutive vertices per thread and one update of each
  // counter per block, so that blocks do not queue on the counters.  Every
  // per-class array is indexed by unrolled loop counters only.
  constexpr int NC = NCLASS + 1;  // [0] hub chunks, [1 + k] class k
  __shared__ unsigned s_w[NC][4], s_b[NC];
  const int64_t v0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * SELECT_PER_THREAD;
  unsigned nch[SELECT_PER_THREAD], cnt[NC] = {};
  int cls[SELECT_PER_THREAD];  // -1: thread-per-row or hub
#pragma unroll
  for (int t = 0; t < SELECT_PER_THREAD; ++t) {
    const int64_t v = v0 + t;
    const int64_t d = v < n ? (int64_t)(xadj[v + 1] - xadj[v]) : 0;
    nch[t] = d > LONG_CHUNK ? (unsigned)((d + LONG_CHUNK - 1) / LONG_CHUNK) : 0u;
    cls[t] = -1;
#pragma unroll
    for (int k = NCLASS - 1; k >= 0; --k)
      if (d > class_lo(k) && d <= class_hi(k)) cls[t] = k;
    cnt[0] += nch[t];
#pragma unroll
    for (int k = 0; k < NCLASS; ++k) cnt[1 + k] += cls[t] == k ? 1u : 0u;
  }
  unsigned inc[NC];
  const int w = threadIdx.x >> 6;
#pragma unroll
  for (int i = 0; i < NC; ++i) {
    inc[i] = wave_incl_scan<unsigned>(cnt[i]);
    if (lane() == CM_WAVE - 1) s_w[i][w] = inc[i];
  }
  __syncthreads();
  if (threadIdx.x < NC) {
    const int i = threadIdx.x;
    const unsigned tot = s_w[i][0] + s_w[i][1] + s_w[i][2] + s_w[i][3];
    s_b[i] = tot ? atomicAdd(&count[i], tot) : 0u;
  }
  __syncthreads();
  unsigned at[NC];
#pragma unroll
  for (int i = 0; i < NC; ++i) {
    at[i] = s_b[i] + inc[i] - cnt[i];
    for (int j = 0; j < w; ++j) at[i] += s_w[i][j];
  }
#pragma unroll
  for (int t = 0; t < SELECT_PER_THREAD; ++t) {
    const int32_t v = (int32_t)(v0 + t);
#pragma unroll
    for (int k = 0; k < NCLASS; ++k)
      if (cls[t] == k) rl.rows[k][at[1 + k]++] = v;
    if (nch[t] > 0) {
      const unsigned hb = at[0];
      for (unsigned q = 0; q < nch[t]; ++q) chunks[hb + q] = LongChunk{v, (int32_t)hb, (int32_t)nch[t], (int32_t)q};
      at[0] += nch[t];
    }
  }
}

// The frontier.  After launch tune.compact_round of a large level most vertices
// are matched, yet every later launch would still visit all n vertices, all
// class rows and all hub chunks to find that out (its time then goes to
// dispatch and one load per item).  k_compact lists again what is still
// unmatched after that launch: the vertices (for the accepts and the
// thread-per-row proposals), the class rows, and the hub chunks of unmatched
// rows (each row's chunks kept together, `first` re-pointed).  A later launch
// that skips a matched vertex skips only work that returns at once, so the
// rounds, flags and matching are those of the full launches.  The lists are
// in no particular order (a launch's result does not depend on its order).
// The round after which the lists are built and the smallest level that
// builds them are launch-shape parameters of the context (tune.compact_round,
// tune.compact_min_n; smaller levels' rounds cost about a launch).
// Vertex regions: compaction block b lists the unmatched vertices of
// [b * VREGION, (b + 1) * VREGION) at slist[b * VREGION ..] with their count
// in vcount[b], and block b of a list launch takes region b.  No atomics: a
// counter shared by many blocks serialises its updates.
constexpr int VREGION = 1024;
constexpr int VREGION_PER_THREAD = VREGION / 256;
// The class lists and the hub chunks: LIST_BLOCKS blocks per list, each
// over a contiguous range of it, with one atomic per block and list.
constexpr int LIST_BLOCKS = 64;

// Exclusive position of this thread's c entries in the list counted by
// *count: one atomic per block.  Block-uniform call (256 threads).
__device__ __forceinline__ unsigned block_append(unsigned c, unsigned *count, unsigned *s_w, unsigned &s_b)
{
  unsigned tot;
  const unsigned before = block256_excl_scan<unsigned>(c, tot, s_w);
  if (threadIdx.x == 0) s_b = tot ? atomicAdd(count, tot) : 0u;
  __syncthreads();
  const unsigned at = s_b + before;
  __syncthreads();  // s_b is reused by the next call
  return at;
}

// Entry i of list L (0: hub chunks, 1 + k: rows of class k) as the number
// of slots it keeps: a class row of an unmatched vertex keeps 1, the first
// chunk of an unmatched hub row keeps the row's chunks.
__device__ __forceinline__ unsigned compact_keep(int L, int64_t i, const int32_t *match, const LongChunk *chunks,
                                                 const RowLists &rl, LongChunk &c)
{
  if (L == 0) {
    c = chunks[i];
    return c.k == 0 && match[c.v] == -1 ? (unsigned)c.nch : 0u;
  }
  int32_t v = 0;
#pragma unroll
  for (int k = 0; k < NCLASS; ++k)
    if (L == 1 + k) v = rl.rows[k][i];
  c.v = v;
  return match[v] == -1 ? 1u : 0u;
}

// Blocks [0, gv): the vertex regions.  Blocks gv + L * LIST_BLOCKS + j: range
// j of list L.  cnt2: [0] hub chunks, [1 + k] rows of class k (zero before
// the launch).
__global__ __launch_bounds__(256) void k_compact(int64_t n, const int32_t *match, const unsigned *flags, int r0,
                                                 const LongChunk *chunks, RowLists rl, const unsigned *cnt,
                                                 LongChunk *chunks2, RowLists rl2, unsigned *cnt2, int32_t *slist,
                                                 unsigned *vcount, unsigned gv)
{
  // launch r0 + 1 and all later ones return at once if round r0 - 1 did
  // nothing: no list is read
  if (!round_live(flags, r0 - 1)) return;
  __shared__ unsigned s_w[4], s_b;
  if (blockIdx.x < gv) {
    const int64_t v0 = (int64_t)blockIdx.x * VREGION + threadIdx.x * VREGION_PER_THREAD;
    int32_t mv[VREGION_PER_THREAD];
    unsigned c = 0;
#pragma unroll
    for (int t = 0; t < VREGION_PER_THREAD; ++t) {
      mv[t] = v0 + t < n ? match[v0 + t] : 0;
      c += mv[t] == -1 ? 1u : 0u;
    }
    unsigned tot;
    unsigned at = block256_excl_scan<unsigned>(c, tot, s_w);
    int32_t *out = slist + (int64_t)blockIdx.x * VREGION;
#pragma unroll
    for (int t = 0; t < VREGION_PER_THREAD; ++t)
      if (mv[t] == -1) out[at++] = (int32_t)(v0 + t);
    if (threadIdx.x == 0) vcount[blockIdx.x] = tot;
    return;
  }
  const int L = (int)((blockIdx.x - gv) / LIST_BLOCKS), j = (int)((blockIdx.x - gv) % LIST_BLOCKS);
  const int64_t nl = cnt[L];
  const int64_t per = (nl + LIST_BLOCKS - 1) / LIST_BLOCKS;
  const int64_t lo = std::min<int64_t>(nl, j * per), hi = std::min<int64_t>(nl, lo + per);
  // two passes over the range: count, one atomic, then write (the second
  // pass reads again what the first brought into the caches)
  unsigned c = 0;
  for (int64_t i = lo + threadIdx.x; i < hi; i += blockDim.x) {
    LongChunk ch;
    c += compact_keep(L, i, match, chunks, rl, ch);
  }
  unsigned at = block_append(c, cnt2 + L, s_w, s_b);
  for (int64_t i = lo + threadIdx.x; i < hi; i += blockDim.x) {
    LongChunk ch;
    const unsigned keep = compact_keep(L, i, match, chunks, rl, ch);
    if (!keep) continue;
    if (L == 0) {
      for (int q = 0; q < ch.nch; ++q) chunks2[at + q] = LongChunk{ch.v, (int32_t)at, ch.nch, q};
    } else {
#pragma unroll
      for (int k = 0; k < NCLASS; ++k)
        if (L == 1 + k) rl2.rows[k][at] = ch.v;
    }
    at += keep;
  }
}

// Wave-uniform copies of lane i's values (i uniform).
__device__ __forceinline__ int32_t bcast(int32_t x, int i) { return __builtin_amdgcn_readlane(x, i); }
__device__ __forceinline__ int64_t bcast(int64_t x, int i)
{
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int32_t)(uint32_t)x, i);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int32_t)(uint32_t)((uint64_t)x >> 32), i);
  return (int64_t)(((uint64_t)hi << 32) | lo);
}
__device__ __forceinline__ uint64_t bcast(uint64_t x, int i) { return (uint64_t)bcast((int64_t)x, i); }

// Accepted in round r-1 (from its proposals alone): dead, or mutual.
__device__ __forceinline__ bool accepted_prev(const int32_t *prev, int32_t u)
{
  const int32_t p = prev[u];
  return p == -2 || (p >= 0 && prev[p] == u);
}

// Best eligible neighbour of v among vertices unmatched after round r-1, by
// the strict order (adjwgt, hash(seed,min,max), min, max); -2 if none.
// Lanes [l0, l0+step) stride the row; the caller reduces across lanes.
// TK (how a neighbour's state is read): TK_MATCH, match[u] plus the proposals
// of round r-1 (accepted_prev: two dependent gathers per candidate that would
// beat the current best); TK_SPLIT, the taken bits alone (bit u: match[u] !=
// -1; they stay in the L2 where match[], 32 times larger, would not),
// complete because k_accept did round r-1's accepts in the launch before.
template <typename OffT, int SCAN_BATCH = 4, int TK = TK_MATCH>
__device__ __forceinline__ void scan_row(int64_t v, OffT b, OffT e, int l0, int step,
                                         const int32_t *__restrict__ adjncy,
                                         const int32_t *__restrict__ adjwgt,
                                         const int32_t *__restrict__ vwgt, int32_t vw, int32_t maxv,
                                         uint64_t seed, int64_t base, const int32_t *match,
                                         const uint32_t *bits, const int32_t *prev, bool have_prev, bool check_w,
                                         bool wunit,
                                         int32_t &best, uint64_t &bk1, uint64_t &bk2)
{
  // SCAN_BATCH entries per lane per step: their row loads and the gathers
  // that depend on them are issued together, so a lane waits once per batch
  // Loads go through clamped indices (the last entry of the range, or v
  // itself) and are unconditional: conditional loads compile to a branch and
  // a wait each, which serialises the batch.
  for (OffT j0 = b + l0; j0 < e; j0 += (OffT)SCAN_BATCH * step) {
    int32_t u[SCAN_BATCH], mu[SCAN_BATCH], wu[SCAN_BATCH];
    uint32_t w[SCAN_BATCH];
#pragma unroll
    for (int t = 0; t < SCAN_BATCH; ++t) {
      const OffT j = j0 + (OffT)t * step;
      const OffT jj = j < e ? j : e - 1;
      const int32_t a = adjncy[jj];
      // unit weights: every lane reads the first word (one cache line)
      // instead of streaming a second array of ones
      const uint32_t x = (uint32_t)adjwgt[wunit ? 0 : jj];
      u[t] = j < e ? a : -1;
      w[t] = j < e ? x : 0u;
    }
    const uint32_t wmin = best != -2 ? (uint32_t)(bk1 >> 32) : 0u;
#pragma unroll
    for (int t = 0; t < SCAN_BATCH; ++t) {
      const bool look = u[t] >= 0 && u[t] != (int32_t)v && w[t] >= wmin;
      const int32_t uu = look ? u[t] : (int32_t)v;
      // round 0 runs before any accept, so every neighbour is unmatched: all
      // lanes read one shared word instead of gathering
      if constexpr (TK == TK_SPLIT)
        mu[t] = (bits[have_prev ? (uint32_t)uu >> 5 : 0u] >> (uu & 31)) & 1u ? 0 : -1;
      else
        mu[t] = match[have_prev ? uu : 0];
      wu[t] = check_w ? vwgt[uu] : 0;  // without check_w no pair can exceed maxv
      if (!look) u[t] = -1;
    }
#pragma unroll
    for (int t = 0; t < SCAN_BATCH; ++t) {
      if (u[t] < 0 || mu[t] != -1 || vw + wu[t] > maxv) continue;
      if (best != -2 && w[t] < (uint32_t)(bk1 >> 32)) continue;
      const uint32_t a = (uint32_t)(min((int64_t)u[t], v) - base), bb = (uint32_t)(max((int64_t)u[t], v) - base);
      const uint64_t k1 = ((uint64_t)w[t] << 32) | edge_hash(seed, a, bb);
      const uint64_t k2 = ((uint64_t)a << 32) | bb;
      // the two dependent gathers of accepted_prev only for edges that would
      // beat the current best (about ln(degree) of them for random keys)
      if ((best == -2 || key_better(k1, k2, bk1, bk2)) &&
          !(TK != TK_SPLIT && have_prev && accepted_prev(prev, u[t]))) {
        best = u[t];
        bk1 = k1;
        bk2 = k2;
      }
    }
  }
}

// Wave-wide maximum of (best, bk1, bk2) under key_better; -2 = none.
__device__ __forceinline__ void wave_best(int32_t &best, uint64_t &bk1, uint64_t &bk2)
{
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) {
    const int32_t ob = __shfl_xor(best, d, CM_WAVE);
    const uint64_t o1 = __shfl_xor(bk1, d, CM_WAVE);
    const uint64_t o2 = __shfl_xor(bk2, d, CM_WAVE);
    if (ob != -2 && (best == -2 || key_better(o1, o2, bk1, bk2))) {
      best = ob;
      bk1 = o1;
      bk2 = o2;
    }
  }
}

// Per-chunk partial results of hub rows.  scanned[first chunk of a row] =
// the launch (r + 1) in which the row's chunks were scanned; k_hub_reduce
// reduces them after that launch.
struct LongPartial {
  int32_t *best;
  uint64_t *k1, *k2;
  unsigned *scanned;
};

// One chunk of a hub row, scanned by the whole wave.  The chunk's best goes
// to part[q] with plain stores; the next kernel on the stream reduces the
// row's chunks (a hand-off inside the launch would need an L2 writeback and
// invalidate per chunk for its release/acquire pair).
template <typename OffT, int TK>
__device__ __forceinline__ void hub_chunk(int64_t q, const LongChunk *chunks, const OffT *__restrict__ xadj,
                                          const int32_t *__restrict__ adjncy, const int32_t *__restrict__ adjwgt,
                                          const int32_t *__restrict__ vwgt, const MatchOwners &own,
                                          const int32_t *match, const uint32_t *bits, const int32_t *prev,
                                          int32_t *cur,
                                          const LongPartial &part, bool have_prev, int64_t maxvw,
                                          unsigned epoch)
{
  const int l = lane();
  const LongChunk c = chunks[q];
  const int32_t v = c.v;
  // loads by dependency level, as in the short path.  A vertex matched
  // before this launch wrote no proposal in round r-1, so its prev entry
  // is stale and must not be followed: test match first.
  // TK_SPLIT: match[] is complete, so accepted_prev is never needed
  constexpr bool ACC = TK != TK_SPLIT;
  const int32_t mv = match[v];
  const int32_t p = have_prev ? prev[v] : -1;
  if (mv != -1) return;
  const int32_t pc = p >= 0 ? p : v;
  const int32_t pp = ACC && have_prev ? prev[pc] : -1;
  const int32_t mp = match[pc];
  const OffT xb = xadj[v], xe = xadj[v + 1];
  const int32_t vw = vwgt[v];
  const int32_t ppp = ACC && have_prev ? prev[pp >= 0 ? pp : pc] : -1;
  const bool v_acc = ACC && have_prev && (p == -2 || (p >= 0 && pp == v));  // accepted_prev(prev, v)
  if (v_acc) return;
  const bool p_acc = pp == -2 || (pp >= 0 && ppp == p);
  if (p >= 0 && mp == -1 && !p_acc) {
    if (l == 0 && c.k == 0) cur[v] = p;
    return;
  }
  int32_t maxv;
  uint64_t seed;
  int64_t base;
  owner_params(own, v, maxv, seed, base);
  int32_t best = -2;
  uint64_t bk1 = 0, bk2 = 0;
  const OffT b = xb + (OffT)c.k * LONG_CHUNK, e0 = xe;
  const OffT e = e0 - b > (OffT)LONG_CHUNK ? b + (OffT)LONG_CHUNK : e0;
  scan_row<OffT, 4, TK>(v, b, e, l, CM_WAVE, adjncy, adjwgt, vwgt, vw, maxv, seed, base, match, bits, prev,
                          have_prev, maxvw < 0 || (int64_t)vw + maxvw > maxv, own.wunit, best, bk1, bk2);
  wave_best(best, bk1, bk2);
  if (l == 0) {
    part.best[q] = best;
    part.k1[q] = bk1;
    part.k2[q] = bk2;
    if (c.k == 0) part.scanned[q] = epoch;
  }
}

// After launch r: cur[v] of every hub row scanned in it = the best of its
// chunks' partial results.  One wave per row, found at its first chunk.
__global__ __launch_bounds__(256) void k_hub_reduce(const LongChunk *chunks, const unsigned *nchunks,
                                                    LongPartial part, int32_t *cur, const unsigned *flags, int r)
{
  if (r >= 2 && !round_live(flags, r - 2)) return;  // launch r did nothing
  const unsigned nh = nchunks[0];
  const int l = lane();
  const int64_t nw = (int64_t)gridDim.x * (blockDim.x / CM_WAVE);
  for (int64_t q = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / CM_WAVE; q < nh; q += nw) {
    const LongChunk c = chunks[q];
    if (c.k != 0 || part.scanned[q] != (unsigned)r + 1u) continue;
    int32_t best = -2;
    uint64_t bk1 = 0, bk2 = 0;
    for (int i = l; i < c.nch; i += CM_WAVE) {
      const int32_t ob = part.best[q + i];
      const uint64_t o1 = part.k1[q + i], o2 = part.k2[q + i];
      if (ob != -2 && (best == -2 || key_better(o1, o2, bk1, bk2))) {
        best = ob;
        bk1 = o1;
        bk2 = o2;
      }
    }
    wave_best(best, bk1, bk2);
    if (l == 0) cur[c.v] = best;
  }
}

// Mid rows mids[m0 .. m0+WAVE_GROUP): lane i loads row i's vertex state (the
// dependency levels of the short path, for all rows of the group at once),
// then the wave scans the rows that need it one after another, with each
// row's parameters read from its lane.  A wave thus waits for the state
// loads once per group instead of once per row.
template <typename OffT, int TK>
__device__ __forceinline__ void mid_group(int64_t m0, unsigned nm, const int32_t *mids,
                                          const OffT *__restrict__ xadj, const int32_t *__restrict__ adjncy,
                                          const int32_t *__restrict__ adjwgt, const int32_t *__restrict__ vwgt,
                                          const MatchOwners &own, const int32_t *match, const uint32_t *bits,
                                          const int32_t *prev,
                                          int32_t *cur, bool have_prev, int64_t maxvw)
{
  const int l = lane();
  const int64_t mi = m0 + l;
  const bool has = l < WAVE_GROUP && mi < (int64_t)nm;
  const int32_t v = mids[has ? mi : m0];
  const int32_t mv = match[v];
  const int32_t p = have_prev ? prev[v] : -1;
  const OffT xb = xadj[v], xe = xadj[v + 1];
  const int32_t vw = vwgt[v];
  int32_t maxv;
  uint64_t seed;
  int64_t base;
  owner_params(own, v, maxv, seed, base);
  // v's own prev entry is stale once v is matched: it is not followed
  const int32_t pc = mv == -1 && p >= 0 ? p : v;
  constexpr bool ACC = TK != TK_SPLIT;  // TK_SPLIT: match[] is complete
  const int32_t pp = ACC && have_prev ? prev[pc] : -1;
  const int32_t mp = match[pc];
  const int32_t ppp = ACC && have_prev ? prev[pp >= 0 ? pp : pc] : -1;
  const bool v_acc = ACC && have_prev && (p == -2 || (p >= 0 && pp == v));  // accepted_prev(prev, v)
  const bool p_acc = pp == -2 || (pp >= 0 && ppp == p);
  const bool live = has && mv == -1 && !v_acc;
  const bool reuse = p >= 0 && mp == -1 && !p_acc;
  if (live && reuse) cur[v] = p;  // still unmatched, so still the best
  const int32_t cw = maxvw < 0 || (int64_t)vw + maxvw > maxv ? 1 : 0;
  uint64_t todo = __ballot(live && !reuse);
  while (todo) {
    const int i = __ffsll((unsigned long long)todo) - 1;
    todo &= todo - 1;
    const int32_t vi = bcast(v, i);
    const OffT bi = (OffT)bcast((int64_t)xb, i), ei = (OffT)bcast((int64_t)xe, i);
    int32_t best = -2;
    uint64_t bk1 = 0, bk2 = 0;
    scan_row<OffT, 4, TK>(vi, bi, ei, l, CM_WAVE, adjncy, adjwgt, vwgt, bcast(vw, i), bcast(maxv, i),
                            bcast(seed, i), bcast(base, i), match, bits, prev, have_prev, bcast(cw, i) != 0,
                            own.wunit, best, bk1, bk2);
    wave_best(best, bk1, bk2);
    if (l == 0) cur[vi] = best;
  }
}

// Best of lanes within each aligned group of SUB lanes.
template <int SUB>
__device__ __forceinline__ void group_best(int32_t &best, uint64_t &bk1, uint64_t &bk2)
{
#pragma unroll
  for (int d = SUB / 2; d > 0; d >>= 1) {
    const int32_t ob = __shfl_xor(best, d, CM_WAVE);
    const uint64_t o1 = __shfl_xor(bk1, d, CM_WAVE);
    const uint64_t o2 = __shfl_xor(bk2, d, CM_WAVE);
    if (ob != -2 && (best == -2 || key_better(o1, o2, bk1, bk2))) {
      best = ob;
      bk1 = o1;
      bk2 = o2;
    }
  }
}

// Rows smids[m0 .. m0+MID_GROUP) of at most 4 * SUB entries: the vertex
// state as in mid_group (lane i for row i), then the rows to scan go
// CM_WAVE / SUB at a time, one per group of SUB lanes (such a row is one step
// of SCAN_BATCH entries per lane).
template <int SUB, typename OffT, int TK>
__device__ __forceinline__ void small_group(int64_t m0, unsigned ns, const int32_t *smids,
                                            const OffT *__restrict__ xadj, const int32_t *__restrict__ adjncy,
                                            const int32_t *__restrict__ adjwgt, const int32_t *__restrict__ vwgt,
                                            const MatchOwners &own, const int32_t *match, const uint32_t *bits,
                                            const int32_t *prev,
                                            int32_t *cur, bool have_prev, int64_t maxvw)
{
  const int l = lane();
  const int64_t mi = m0 + l;
  const bool has = l < MID_GROUP && mi < (int64_t)ns;
  const int32_t v = smids[has ? mi : m0];
  const int32_t mv = match[v];
  const int32_t p = have_prev ? prev[v] : -1;
  const OffT xb = xadj[v], xe = xadj[v + 1];
  const int32_t vw = vwgt[v];
  int32_t maxv;
  uint64_t seed;
  int64_t base;
  owner_params(own, v, maxv, seed, base);
  const int32_t pc = mv == -1 && p >= 0 ? p : v;
  constexpr bool ACC = TK != TK_SPLIT;  // TK_SPLIT: match[] is complete
  const int32_t pp = ACC && have_prev ? prev[pc] : -1;
  const int32_t mp = match[pc];
  const int32_t ppp = ACC && have_prev ? prev[pp >= 0 ? pp : pc] : -1;
  const bool v_acc = ACC && have_prev && (p == -2 || (p >= 0 && pp == v));  // accepted_prev(prev, v)
  const bool p_acc = pp == -2 || (pp >= 0 && ppp == p);
  const bool live = has && mv == -1 && !v_acc;
  const bool reuse = p >= 0 && mp == -1 && !p_acc;
  if (live && reuse) cur[v] = p;  // still unmatched, so still the best
  const int32_t cw = maxvw < 0 || (int64_t)vw + maxvw > maxv ? 1 : 0;
  uint64_t todo = __ballot(live && !reuse);
  const int g = l / SUB, gl = l % SUB;
  while (todo) {
    // group g takes the g-th row still to do (none if fewer remain)
    uint64_t t = todo;
    for (int k = 0; k < g; ++k) t &= t - 1;
    const int i = t ? __ffsll((unsigned long long)t) - 1 : -1;
    for (int k = 0; k < CM_WAVE / SUB; ++k) todo &= todo - 1;
    const int src = i >= 0 ? i : 0;
    const int32_t vi = __shfl(v, src, CM_WAVE);
    const int64_t bi = __shfl((int64_t)xb, src, CM_WAVE), ei = __shfl((int64_t)xe, src, CM_WAVE);
    const int32_t vwi = __shfl(vw, src, CM_WAVE), maxvi = __shfl(maxv, src, CM_WAVE);
    const uint64_t seedi = __shfl(seed, src, CM_WAVE);
    const int64_t basei = __shfl(base, src, CM_WAVE);
    const int32_t cwi = __shfl(cw, src, CM_WAVE);
    int32_t best = -2;
    uint64_t bk1 = 0, bk2 = 0;
    if (i >= 0)
      scan_row<OffT, 4, TK>(vi, (OffT)bi, (OffT)ei, gl, SUB, adjncy, adjwgt, vwgt, vwi, maxvi, seedi, basei,
                              match, bits, prev, have_prev, cwi != 0, own.wunit, best, bk1, bk2);
    group_best<SUB>(best, bk1, bk2);
    if (i >= 0 && gl == 0) cur[vi] = best;
  }
}

// At most 7 waves per SIMD: the register bound keeps the lane-group paths
// from taking the kernel down to 6 waves, which the latency-bound row scans
// feel (R-MAT rounds -1.5%).
// Launch r of the handshake: accept of round r-1 (from prev = proposals of
// round r-1) and proposals of round r (into cur), over the state after round
// r-1.  match[] is written only for the thread's own vertex; readers treat a
// neighbour as taken if match[u] != -1 or it was accepted in round r-1, which
// gives the same answer whether or not the neighbour's own write has landed.
// One vertex of the thread-per-row path: the accept of round r-1 for every
// vertex, the proposal of round r for rows of at most short_deg entries.
// changed / active are or-ed with this vertex's status.
// An unmatched vertex, for launch 0 of a level whose match array that
// launch initialises itself: its scans read this word instead of match[0].
__device__ const int32_t unmatched_word = -1;

// TK_SPLIT: no accept here (k_accept did it), and match[] is complete.
template <typename OffT, bool INIT = false, int SB = 4, int TK = TK_MATCH>
__device__ __forceinline__ void short_row(int64_t v, int64_t n, const OffT *__restrict__ xadj,
                                          const int32_t *__restrict__ adjncy, const int32_t *__restrict__ adjwgt,
                                          const int32_t *__restrict__ vwgt, const MatchOwners &own, int32_t *match,
                                          const uint32_t *bits, const int32_t *prev, int32_t *cur, bool have_prev,
                                          int propose, int64_t maxvw, int short_deg, bool &changed, bool &active)
{
  // INIT (launch 0 of a level without long rows, which has no k_match_init):
  // v is unmatched; this thread writes match[v] and the other proposal
  // half's entry (the prev of launch 0) as k_match_init would
  if (INIT && v < n) {
    match[v] = -1;
    const_cast<int32_t *>(prev)[v] = -1;
  }
  // Loads in three dependency levels (by v; by p = prev[v]; by prev[p]),
  // each through clamped indices so that a level issues as one batch.
  const int64_t vc = v < n ? v : n - 1;
  const int32_t mv = INIT ? -1 : match[vc];
  const int32_t p = have_prev ? prev[vc] : -1;
  if (v < n && mv == -1) {
    constexpr bool ACC = TK != TK_SPLIT;
    const int32_t pc = p >= 0 ? p : (int32_t)v;
    const int32_t pp = ACC && have_prev ? prev[pc] : -1;  // prev[p]
    const int32_t mp = match[pc];
    const OffT b = xadj[v], e = xadj[v + 1];
    const int32_t vw = vwgt[v];
    const int32_t ppp = ACC && have_prev ? prev[pp >= 0 ? pp : pc] : -1;  // prev[prev[p]]
    bool done = false;
    if (ACC && have_prev) {
      if (p == -2) {
        match[v] = (int32_t)v;
        done = true;
      } else if (p >= 0 && pp == (int32_t)v) {
        match[v] = p;
        done = true;
      }
    }
    changed |= done;
    active |= !done;
    if (!done && propose && (int64_t)(e - b) <= short_deg) {
      // p accepted in round r-1 (from the proposals alone): dead or mutual
      const bool p_acc = pp == -2 || (pp >= 0 && ppp == p);
      if (p >= 0 && mp == -1 && !p_acc) {
        cur[v] = p;  // still unmatched, so still the best
      } else {
        int32_t maxv;
        uint64_t seed;
        int64_t base;
        owner_params(own, v, maxv, seed, base);
        int32_t best = -2;
        uint64_t bk1 = 0, bk2 = 0;
        scan_row<OffT, SB, TK>(v, b, e, 0, 1, adjncy, adjwgt, vwgt, vw, maxv, seed, base,
                                 INIT ? &unmatched_word : match, bits, prev, have_prev,
                                 maxvw < 0 || (int64_t)vw + maxvw > maxv, own.wunit, best, bk1, bk2);
        cur[v] = best;
      }
    }
  }
}

// Blocks [0, nshort) take one vertex per thread (rows <= short_deg and all
// accepts), or (LIST) take vertex region blockIdx.x of the frontier:
// slist[blockIdx.x * VREGION ..] with vcount[blockIdx.x] entries;
// the remaining blocks take the chunks of hub rows and the row groups of the
// classes, one wave per item.
// MESH: a level without rows above LONG_DEG entries (every row on the
// thread-per-row path, no long-row lists): the long-row paths are compiled
// out, which leaves the kernel without SGPR spills (the full round kernel
// spills 34 to VGPR lanes); grid -3.6%.
// TK_SPLIT: the accepts (and taken bits and round flags) are k_accept's,
// launched just before; this launch only proposes, and not at all when round
// r-1 left nothing to do.
template <typename OffT, bool LIST, bool INIT = false, bool MESH = false, int TK = TK_MATCH>
__device__ __forceinline__ void round_body(int64_t n, const OffT *__restrict__ xadj,
                                           const int32_t *__restrict__ adjncy, const int32_t *__restrict__ adjwgt,
                                           const int32_t *__restrict__ vwgt, const MatchOwners &own, int32_t *match,
                                           const uint32_t *bits, const int32_t *prev, int32_t *cur,
                                           const LongChunk *chunks,
                                           const RowLists &rl, const unsigned *nchunks, const LongPartial &part,
                                           unsigned *flags, int r, unsigned nshort, int propose, int64_t maxvw,
                                           int short_deg, const int32_t *slist, const unsigned *vcount,
                                           int64_t nflags = 0, unsigned *hzero = nullptr)
{
  if (r >= 2 && !round_live(flags, r - 2)) return;
  if (TK == TK_SPLIT && r >= 1 && !round_live(flags, r - 1)) return;
  const bool have_prev = r >= 1;
  if (INIT) {  // the rest of k_match_init: the round flags and the hash counters
    const int64_t gt = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gt < 16) hzero[gt] = 0u;
    for (int64_t i = gt; i < nflags; i += (int64_t)gridDim.x * blockDim.x) flags[i] = 0u;
  }
  if (blockIdx.x < nshort) {
    __shared__ int s_changed, s_active;
    if (threadIdx.x == 0) {
      s_changed = 0;
      s_active = 0;
    }
    __syncthreads();
    bool changed = false, active = false;
    if constexpr (LIST) {  // the frontier: only vertices still unmatched are listed
      const unsigned cnt = vcount[blockIdx.x];
      const int32_t *list = slist + (int64_t)blockIdx.x * VREGION;
      for (unsigned i = threadIdx.x; i < cnt; i += blockDim.x)
        short_row<OffT, false, 4, TK>(list[i], n, xadj, adjncy, adjwgt, vwgt, own, match, bits, prev, cur,
                                      have_prev, propose, maxvw, short_deg, changed, active);
    } else {
      short_row<OffT, INIT, MESH ? MESH_BATCH : 4, TK>((int64_t)blockIdx.x * blockDim.x + threadIdx.x, n, xadj,
                                                       adjncy, adjwgt, vwgt, own, match, bits, prev, cur, have_prev,
                                                       propose, maxvw, short_deg, changed, active);
    }
    if (TK != TK_SPLIT && have_prev) {
      if (__ballot(changed) && lane() == 0) s_changed = 1;
      if (__ballot(active) && lane() == 0) s_active = 1;
      __syncthreads();
      // round status words are flags, read only by later launches: one
      // plain store of 1 per block.  Stores of one value from many blocks
      // merge in the L2s and reach memory at the end of the launch; an atomic
      // (or a check-then-store through the memory-side coherence point) would
      // add a round trip to the end of every block.
      if (threadIdx.x == 0) {
        unsigned *f = flags + 2 * (r - 1);
        if (s_changed) f[0] = 1u;
        if (s_active) f[1] = 1u;
      }
    }
    return;
  }
  if constexpr (INIT || MESH) return;  // the level has no long rows
  if (!propose) return;
  // Long rows.  Work items in order: the hub chunks (one wave each), then
  // the row groups of each class in class order (one wave per group).
  // Whether a row is scanned depends only on state fixed before this launch
  // (see above), so all chunks of a row agree.
  const unsigned nh = nchunks[0];
  int64_t gend[NCLASS];
  int64_t items = nh;
#pragma unroll
  for (int k = 0; k < NCLASS; ++k) {
    const int rows = k == MID_CLASS ? WAVE_GROUP : MID_GROUP;
    items += ((int64_t)nchunks[1 + k] + rows - 1) / rows;
    gend[k] = items;
  }
  const int64_t nw = (int64_t)(gridDim.x - nshort) * (blockDim.x / CM_WAVE);
  for (int64_t it = ((int64_t)(blockIdx.x - nshort) * blockDim.x + threadIdx.x) / CM_WAVE; it < items; it += nw) {
    if (it < nh)
      hub_chunk<OffT, TK>(it, chunks, xadj, adjncy, adjwgt, vwgt, own, match, bits, prev, cur, part, have_prev,
                            maxvw, (unsigned)r + 1u);
    else if (it < gend[0])
      small_group<class_lanes(0), OffT, TK>((it - nh) * MID_GROUP, nchunks[1], rl.rows[0], xadj, adjncy, adjwgt,
                                              vwgt, own, match, bits, prev, cur, have_prev, maxvw);
    else if (it < gend[1])
      small_group<class_lanes(1), OffT, TK>((it - gend[0]) * MID_GROUP, nchunks[2], rl.rows[1], xadj, adjncy,
                                              adjwgt, vwgt, own, match, bits, prev, cur, have_prev, maxvw);
    else if (it < gend[2])
      small_group<class_lanes(2), OffT, TK>((it - gend[1]) * MID_GROUP, nchunks[3], rl.rows[2], xadj, adjncy,
                                              adjwgt, vwgt, own, match, bits, prev, cur, have_prev, maxvw);
    else
      mid_group<OffT, TK>((it - gend[2]) * WAVE_GROUP, nchunks[1 + MID_CLASS], rl.rows[MID_CLASS], xadj, adjncy,
                            adjwgt, vwgt, own, match, bits, prev, cur, have_prev, maxvw);
  }
}

// Launch r of the handshake over all vertices and the level's long-row lists.
template <typename OffT, int TK>
__global__ __launch_bounds__(256, 7) void k_round(int64_t n, const OffT *__restrict__ xadj,
                                               const int32_t *__restrict__ adjncy,
                                               const int32_t *__restrict__ adjwgt,
                                               const int32_t *__restrict__ vwgt, MatchOwners own,
                                               int32_t *match, const uint32_t *bits, const int32_t *prev, int32_t *cur,
                                               const LongChunk *chunks, RowLists rl, const unsigned *nchunks,
                                               LongPartial part, unsigned *flags, int r, unsigned nshort,
                                               int propose, int64_t maxvw, int short_deg, const int64_t *np)
{
  if (np) n = *np;  // queued ahead: the level's count was not yet on the host
  round_body<OffT, false, false, false, TK>(n, xadj, adjncy, adjwgt, vwgt, own, match, bits, prev, cur, chunks, rl,
                                              nchunks, part, flags, r, nshort, propose, maxvw, short_deg, nullptr,
                                              nullptr);
}

// Launch r of a level without long rows (thread-per-row path only).
template <typename OffT>
__global__ __launch_bounds__(256, 7) void k_round_mesh(int64_t n, const OffT *__restrict__ xadj,
                                                    const int32_t *__restrict__ adjncy,
                                                    const int32_t *__restrict__ adjwgt,
                                                    const int32_t *__restrict__ vwgt, MatchOwners own,
                                                    int32_t *match, const int32_t *prev, int32_t *cur,
                                                    const LongChunk *chunks, RowLists rl, const unsigned *nchunks,
                                                    LongPartial part, unsigned *flags, int r, unsigned nshort,
                                                    int propose, int64_t maxvw, int short_deg, const int64_t *np)
{
  if (np) n = *np;  // queued ahead: the level's count was not yet on the host
  round_body<OffT, false, false, true>(n, xadj, adjncy, adjwgt, vwgt, own, match, nullptr, prev, cur, chunks, rl,
                                       nchunks, part, flags, r, nshort, propose, maxvw, short_deg, nullptr, nullptr);
}

// Launch 0 (round 0 proposals only: no accept, no prev), compiled for r = 0.
// INIT: a level without long rows, whose launch 0 also does the work of
// k_match_init (the level's first launch then starts the rounds at once).
template <typename OffT, bool INIT>
__global__ __launch_bounds__(256, 7) void k_round0(int64_t n, const OffT *__restrict__ xadj,
                                                const int32_t *__restrict__ adjncy,
                                                const int32_t *__restrict__ adjwgt,
                                                const int32_t *__restrict__ vwgt, MatchOwners own, int32_t *match,
                                                int32_t *prev, int32_t *cur, const LongChunk *chunks, RowLists rl,
                                                const unsigned *nchunks, LongPartial part, unsigned *flags,
                                                unsigned nshort, int64_t maxvw, int short_deg, int64_t nflags,
                                                unsigned *hzero, const int64_t *np)
{
  if (np) n = *np;  // queued ahead: the level's count was not yet on the host
  round_body<OffT, false, INIT, INIT>(n, xadj, adjncy, adjwgt, vwgt, own, match, nullptr, prev, cur, chunks, rl,
                                      nchunks, part, flags, 0, nshort, 1, maxvw, short_deg, nullptr, nullptr, nflags,
                                      hzero);
}

// Launch r over the frontier (vertex regions and compacted lists).
template <typename OffT, int TK>
__global__ __launch_bounds__(256, 7) void k_round_list(int64_t n, const OffT *__restrict__ xadj,
                                                    const int32_t *__restrict__ adjncy,
                                                    const int32_t *__restrict__ adjwgt,
                                                    const int32_t *__restrict__ vwgt, MatchOwners own,
                                                    int32_t *match, const uint32_t *bits, const int32_t *prev,
                                                    int32_t *cur,
                                                    const LongChunk *chunks, RowLists rl, const unsigned *nchunks,
                                                    LongPartial part, unsigned *flags, int r, unsigned nshort,
                                                    int propose, int64_t maxvw, int short_deg,
                                                    const int32_t *slist, const unsigned *vcount)
{
  round_body<OffT, true, false, false, TK>(n, xadj, adjncy, adjwgt, vwgt, own, match, bits, prev, cur, chunks, rl,
                                             nchunks, part, flags, r, nshort, propose, maxvw, short_deg, slist,
                                             vcount);
}

// The accepts of round r-1 on a TK_SPLIT level (launch r's first kernel):
// every vertex still unmatched (LIST: each listed vertex of region
// blockIdx.x) takes its mutual partner, or itself if it proposed to no one
// (-2).  It writes the taken bits, and round r-1's flags as round_body does
// on the other levels.
template <bool LIST>
__global__ __launch_bounds__(256) void k_accept(int64_t n, int32_t *match, uint32_t *bits, const int32_t *prev,
                                                unsigned *flags, int r, const int32_t *slist, const unsigned *vcount)
{
  if (r >= 2 && !round_live(flags, r - 2)) return;
  __shared__ int s_changed, s_active;
  __shared__ uint32_t s_bits[VREGION / 32];
  if (threadIdx.x == 0) {
    s_changed = 0;
    s_active = 0;
  }
  if (LIST && threadIdx.x < VREGION / 32) s_bits[threadIdx.x] = 0u;
  __syncthreads();
  bool changed = false, active = false;
  // v < n; returns whether v is matched after the accept
  auto accept = [&](int64_t v) -> bool {
    const int32_t mv = match[v];
    const int32_t p = prev[v];
    if (mv != -1) return true;
    const int32_t pp = prev[p >= 0 ? (int64_t)p : v];
    const int32_t m = p == -2 ? (int32_t)v : (p >= 0 && pp == (int32_t)v ? p : -1);
    if (m != -1) match[v] = m;
    changed |= m != -1;
    active |= m == -1;
    return m != -1;
  };
  if constexpr (LIST) {
    const unsigned cnt = vcount[blockIdx.x];
    const int32_t *list = slist + (int64_t)blockIdx.x * VREGION;
    for (unsigned i = threadIdx.x; i < cnt; i += blockDim.x) {
      const int32_t v = list[i];  // region blockIdx.x: [blockIdx.x * VREGION, + VREGION)
      if (accept(v)) atomicOr(&s_bits[(v & (VREGION - 1)) >> 5], 1u << (v & 31));
    }
  } else {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t mk = __ballot(v < n && accept(v));
    const int64_t w0 = v - lane();  // the wave's first vertex, a multiple of 64
    if (lane() == 0 && w0 < n) reinterpret_cast<uint64_t *>(bits)[w0 >> 6] = mk;
  }
  if (__ballot(changed) && lane() == 0) s_changed = 1;
  if (__ballot(active) && lane() == 0) s_active = 1;
  __syncthreads();
  if (LIST && threadIdx.x < VREGION / 32 && s_bits[threadIdx.x])
    bits[(int64_t)blockIdx.x * (VREGION / 32) + threadIdx.x] |= s_bits[threadIdx.x];  // the region's own words
  if (threadIdx.x == 0) {
    unsigned *f = flags + 2 * (r - 1);
    if (s_changed) f[0] = 1u;
    if (s_active) f[1] = 1u;
  }
}

// Small levels without long rows: every round in one workgroup, the state
// in LDS (16-bit ids: -1 unmatched / no proposal, -2 dead).  It runs what
// launches 0 .. R of the kernels above would (R = the round budget): launch
// r's accepts of round r-1 and proposals of round r, the same round flags,
// the same stop (a round that matched nothing or left nothing unmatched).
// A level still live after R rounds leaves the state of launch R (match[]
// after round R-1's accepts, round R's proposals in prop half R & 1) for
// the ordinary launches to continue from launch R + 1.  Also does launch 0's
// initialisation of a level without long rows (flags, hash counters).
constexpr int LDS_MATCH_THREADS = 1024;
// Up to 4096 vertices (Delaunay 4720's levels after the first: 0.457 ->
// 0.413 ms per coarsening).  One workgroup scans its vertices' rows one after
// another, so larger levels are faster as launches of the round kernel (grid
// 725^2 with this form up to 8192 or 16384 vertices: 0.94 -> 0.97 or 1.03 ms).
constexpr int64_t LDS_MATCH_MAX_N = 4096;
constexpr int LDS_ROUNDS = 32;  // the round budget of one k_match_lds

template <typename OffT>
__device__ __forceinline__ int32_t lds_scan(int32_t v, const OffT *__restrict__ xadj,
                                            const int32_t *__restrict__ adjncy,
                                            const int32_t *__restrict__ adjwgt,
                                            const int32_t *__restrict__ vwgt, const MatchOwners &own,
                                            const int16_t *s_m)
{
  int32_t maxv;
  uint64_t seed;
  int64_t base;
  owner_params(own, v, maxv, seed, base);
  const int32_t vw = vwgt[v];
  // 32-bit entry indices: a level of at most LDS_MATCH_MAX_N vertices holds
  // fewer than 2^31 entries.  (With int64 offsets a 64-bit index here gave
  // wrong proposals in the rescans of rounds >= 1, and only there; the 32-bit
  // form matches the oracle and the int32 build.)
  const int b = (int)xadj[v], e = (int)xadj[v + 1];
  int32_t best = -2;
  uint64_t bk1 = 0, bk2 = 0;
  for (int j0 = b; j0 < e; j0 += 4) {
    int32_t u[4], wu[4];
    uint32_t w[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {  // clamped, unconditional loads: one wait per batch
      const int jj = j0 + t < e ? j0 + t : e - 1;
      u[t] = adjncy[jj];
      w[t] = (uint32_t)adjwgt[jj];
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const bool in = j0 + t < e && u[t] != v && s_m[u[t]] == -1;
      wu[t] = vwgt[in ? u[t] : v];
      if (!in) u[t] = -1;
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      if (u[t] < 0 || vw + wu[t] > maxv) continue;
      const uint32_t a = (uint32_t)(min(u[t], v) - base), bb = (uint32_t)(max(u[t], v) - base);
      const uint64_t k1 = ((uint64_t)w[t] << 32) | edge_hash(seed, a, bb);
      const uint64_t k2 = ((uint64_t)a << 32) | bb;
      if (best == -2 || key_better(k1, k2, bk1, bk2)) {
        best = u[t];
        bk1 = k1;
        bk2 = k2;
      }
    }
  }
  return best;
}

template <typename OffT>
__global__ __launch_bounds__(LDS_MATCH_THREADS) void k_match_lds(int64_t n, const OffT *__restrict__ xadj,
                                                                 const int32_t *__restrict__ adjncy,
                                                                 const int32_t *__restrict__ adjwgt,
                                                                 const int32_t *__restrict__ vwgt, MatchOwners own,
                                                                 int32_t *match, int32_t *prop, int64_t half,
                                                                 unsigned *flags, int64_t nflags, unsigned *hzero,
                                                                 int rounds, int max_rounds, const int64_t *np)
{
  __shared__ int s_changed, s_active;
  __shared__ int16_t s_state[2 * LDS_MATCH_MAX_N];  // match, then proposals
  int16_t *const s_m = s_state, *const s_p = s_state + LDS_MATCH_MAX_N;
  if (np) n = *np;  // queued ahead: the level's count was not yet on the host
  const int nv = (int)n;
  const int t = threadIdx.x;
  if (t < 16) hzero[t] = 0u;
  for (int64_t i = t; i < nflags; i += LDS_MATCH_THREADS) flags[i] = 0u;
  for (int v = t; v < nv; v += LDS_MATCH_THREADS) s_m[v] = -1;
  __syncthreads();
  // launch 0: round 0's proposals
  for (int v = t; v < nv; v += LDS_MATCH_THREADS) s_p[v] = (int16_t)lds_scan<OffT>(v, xadj, adjncy, adjwgt, vwgt, own, s_m);
  __syncthreads();
  bool live = true;
  int r = 1;
  for (; r <= rounds; ++r) {
    // launch r: round r-1's accepts (reading proposals only)
    if (t == 0) {
      s_changed = 0;
      s_active = 0;
    }
    __syncthreads();
    bool changed = false, active = false;
    for (int v = t; v < nv; v += LDS_MATCH_THREADS) {
      if (s_m[v] != -1) continue;
      const int p = s_p[v];
      const int q = p == -2 ? v : (p >= 0 && s_p[p] == v ? p : -1);
      if (q != -1) s_m[v] = (int16_t)q;
      changed |= q != -1;
      active |= q == -1;
    }
    if (__ballot(changed) && lane() == 0) s_changed = 1;
    if (__ballot(active) && lane() == 0) s_active = 1;
    __syncthreads();
    live = s_changed && s_active;
    if (t == 0) {
      flags[2 * (r - 1)] = s_changed ? 1u : 0u;
      flags[2 * (r - 1) + 1] = s_active ? 1u : 0u;
    }
    if (!live || r == max_rounds) break;  // launch max_rounds only accepts
    // round r's proposals over the state after round r-1: a proposal is kept
    // while its target is unmatched (it is then still the best)
    for (int v = t; v < nv; v += LDS_MATCH_THREADS) {
      if (s_m[v] != -1) continue;
      const int p = s_p[v];
      if (!(p >= 0 && s_m[p] == -1)) s_p[v] = (int16_t)lds_scan<OffT>(v, xadj, adjncy, adjwgt, vwgt, own, s_m);
    }
    __syncthreads();
  }
  // the working match; a level still live continues from launch rounds + 1
  for (int v = t; v < nv; v += LDS_MATCH_THREADS) {
    match[v] = s_m[v];
    if (live && r > rounds) prop[(int64_t)(rounds & 1) * half + v] = s_p[v];
  }
}

// out[v] = match[v], with vertices still unmatched (-1) matched to themselves;
// out may be match itself.
__global__ __launch_bounds__(256) void k_match_finish(int64_t n, const int32_t *match, int32_t *out)
{
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v < n) {
    const int32_t u = match[v];
    out[v] = u == -1 ? (int32_t)v : u;
  }
}

// Device state of one matching between host calls: the round launches can be
// queued in batches without reading anything back, and resumed if the flags
// show that the last batch did not finish.
struct MatchRun {
  const Level *L = nullptr;
  Level own_level;             // the level of a run queued ahead (L points here)
  const int64_t *d_n = nullptr;  // queued ahead: launches read the count here
  int64_t half = 0;            // stride of the two proposal halves (the n it began with)
  MatchOwners own;
  int max_rounds = 0;
  int32_t *match = nullptr;
  int32_t *prop = nullptr;
  uint32_t *bits = nullptr;     // taken bits (large levels with long rows)
  LongChunk *chunks = nullptr;  // hub chunks
  RowLists rl{};                // rows by class
  int short_deg = LONG_DEG;     // longest row on the thread-per-row path
  LongPartial part{};
  unsigned *flags = nullptr, *nchunks = nullptr;
  unsigned g = 0, glong = 0;
  unsigned ghub = 0;  // blocks of k_hub_reduce (0: the level has no hub row)
  bool fused_init = false;  // launch 0 initialises match, proposals and flags
  bool lds = false;         // launches 0 .. LDS_ROUNDS are one k_match_lds
  // the frontier lists from launch compact_round + 1 on (large levels only)
  bool compact = false;
  int compact_round = 0;
  LongChunk *chunks2 = nullptr;
  RowLists rl2{};
  int32_t *slist = nullptr;
  unsigned *vcount = nullptr;    // vertices listed per region
  unsigned *nchunks2 = nullptr;  // [0] hub chunks, [1 + k] class k
  unsigned gs = 0;               // short blocks of the list launches: one per vertex region
  int r = 0;  // launches queued so far
  std::vector<void *> blocks;
};

template <typename OffT>
static void match_start(cm_ctx *ctx, MatchRun &mr)
{
  const Level &L = *mr.L;
  const int64_t n = L.n;
  const OffT *xadj = (const OffT *)L.xadj;
  auto al = [&](size_t bytes) {
    void *p = ctx->pool.alloc(bytes ? bytes : 1);
    mr.blocks.push_back(p);
    return p;
  };
  mr.prop = (int32_t *)al(sizeof(int32_t) * 2 * n);  // ping-pong
  mr.half = n;
  // the shared unit array is never written, so a level that points to it has
  // unit edge weights
  mr.own.wunit = L.m > 0 && L.adjwgt == ctx->ones;
  // chunks: sum over rows longer than LONG_CHUNK of ceil(d / LONG_CHUNK)
  const bool has_long = L.maxdeg < 0 || L.maxdeg > LONG_DEG;
  // hub chunks: at most m / LONG_CHUNK full ones plus one partial per hub
  // row (each hub row holds more than LONG_CHUNK entries)
  const int64_t cap = has_long ? 2 * (L.m / LONG_CHUNK) + 1 : 1;
  mr.short_deg = has_long ? SHORT_DEG : LONG_DEG;
  mr.chunks = (LongChunk *)al(sizeof(LongChunk) * cap);
  int64_t groups = 0;  // wave items of the class lists, at most
  for (int k = 0; k < NCLASS; ++k) {
    const int64_t ck = has_long ? std::min<int64_t>(n, L.m / (class_lo(k) + 1)) + 1 : 1;
    mr.rl.rows[k] = (int32_t *)al(sizeof(int32_t) * ck);
    const int rows = k == MID_CLASS ? WAVE_GROUP : MID_GROUP;
    groups += (ck + rows - 1) / rows;
  }
  mr.part.best = (int32_t *)al(sizeof(int32_t) * cap);
  mr.part.k1 = (uint64_t *)al(sizeof(uint64_t) * cap);
  mr.part.k2 = (uint64_t *)al(sizeof(uint64_t) * cap);
  mr.part.scanned = (unsigned *)al(sizeof(unsigned) * cap);
  mr.flags = (unsigned *)al(sizeof(unsigned) * (2 * (size_t)mr.max_rounds + 16));
  mr.nchunks = mr.flags + 2 * (size_t)mr.max_rounds;  // [0] hub chunks, [1 + k] rows of class k
  mr.nchunks2 = mr.nchunks + 8;
  static_assert(NCLASS + 1 <= 8, "list counts");
  mr.compact_round = ctx->tune.compact_round;
  mr.compact = !mr.d_n && n >= ctx->tune.compact_min_n && mr.max_rounds > mr.compact_round + 1;
  if (mr.compact) {
    mr.slist = (int32_t *)al(sizeof(int32_t) * (size_t)nblocks(n, VREGION) * VREGION);
    mr.vcount = (unsigned *)al(sizeof(unsigned) * nblocks(n, VREGION));
    mr.chunks2 = (LongChunk *)al(sizeof(LongChunk) * cap);
    for (int k = 0; k < NCLASS; ++k) {
      const int64_t ck = has_long ? std::min<int64_t>(n, L.m / (class_lo(k) + 1)) + 1 : 1;
      mr.rl2.rows[k] = (int32_t *)al(sizeof(int32_t) * ck);
    }
  }
  const bool has_hub = L.maxdeg < 0 || L.maxdeg > LONG_CHUNK;
  if (has_hub) fill_u32(ctx, mr.part.scanned, 0u, cap);  // no launch number is 0
  mr.g = nblocks(n, 256);
  // a level without long rows has launch 0 initialise the state itself (its
  // thread-per-row blocks cover every vertex); the long-row lists are
  // counted before launch 0, into words that k_match_init clears
  mr.fused_init = !has_long;
  mr.lds = mr.fused_init && n <= LDS_MATCH_MAX_N && !mr.compact;
  const int64_t nbw = !mr.fused_init && n >= BITS_MIN_N ? (n + 63) / 64 : 0;
  if (nbw) mr.bits = (uint32_t *)al(sizeof(uint64_t) * nbw);
  if (!mr.fused_init)
    launch(ctx, "match_init", 8.0 * n, k_match_init, dim3(mr.g), dim3(256), 0, n, mr.match, mr.prop, mr.flags,
           2 * (int64_t)mr.max_rounds + 16, ctx->d_ctr + HASH_CTR_SLOT, (uint64_t *)mr.bits, nbw);
  ctx->hash_ctr_zero = true;  // by k_match_init or by launch 0
  // extra blocks whose waves stride over the long-row chunks (none when the
  // level is known to have no long row)
  if (has_long) {
    launch(ctx, "match_select_long", (double)sizeof(OffT) * (n + 1), k_select_long<OffT>,
           dim3(nblocks(n, 256 * SELECT_PER_THREAD)), dim3(256), 0, n, xadj, mr.chunks, mr.rl, mr.nchunks);
    const int64_t items = cap + groups;
    mr.glong = (unsigned)std::min<int64_t>(8192, std::max<int64_t>(16, (items + 3) / 4));  // a wave each
    if (has_hub) mr.ghub = (unsigned)std::min<int64_t>(4096, (cap + 3) / 4);
  }
  mr.gs = nblocks(n, VREGION);
}

// Queues up to `count` more launches (launch r: accept of round r-1,
// proposals of round r; launch max_rounds only accepts).
template <typename OffT>
static void match_launch_direct(cm_ctx *ctx, MatchRun &mr, int count)
{
  const Level &L = *mr.L;
  const int64_t n = L.n;
  const double row_bytes = (double)sizeof(OffT) * (n + 1) + 8.0 * L.m + 8.0 * n;
  if (mr.lds && mr.r == 0 && count > 0) {  // launches 0 .. LDS_ROUNDS in one
    const int rounds = std::min(LDS_ROUNDS, mr.max_rounds);
    launch(ctx, "match_lds", row_bytes, k_match_lds<OffT>, dim3(1), dim3(LDS_MATCH_THREADS), 0, n,
           (const OffT *)L.xadj, L.adjncy, L.adjwgt, L.vwgt, mr.own, mr.match, mr.prop, mr.half, mr.flags,
           2 * (int64_t)mr.max_rounds + 16, ctx->d_ctr + HASH_CTR_SLOT, rounds, mr.max_rounds, mr.d_n);
    mr.r = rounds + 1;
    return;
  }
  for (int i = 0; i < count && mr.r <= mr.max_rounds; ++i, ++mr.r) {
    const int r = mr.r;
    int32_t *cur = mr.prop + (size_t)(r & 1) * mr.half;
    const int32_t *prv = mr.prop + (size_t)((r + 1) & 1) * mr.half;
    const bool lst = mr.compact && r > mr.compact_round;
    const LongChunk *chunks = lst ? mr.chunks2 : mr.chunks;
    const unsigned *counts = lst ? mr.nchunks2 : mr.nchunks;
    const unsigned gshort = lst ? mr.gs : mr.g, glong = mr.glong;
    const int tk = mr.bits ? TK_SPLIT : TK_MATCH;
    if (tk == TK_SPLIT && r >= 1)
      launch(ctx, "match_accept", 0.0, lst ? k_accept<true> : k_accept<false>, dim3(gshort), dim3(256), 0, n,
             mr.match, mr.bits, prv, mr.flags, r, (const int32_t *)mr.slist, (const unsigned *)mr.vcount);
    if (tk == TK_SPLIT && r >= 1 && r == mr.max_rounds) {
      // the last launch only accepts: k_accept was all of it
    } else if (lst)
      launch(ctx, "match_round", 0.0,
             tk == TK_SPLIT ? k_round_list<OffT, TK_SPLIT> : k_round_list<OffT, TK_MATCH>,
             dim3(gshort + glong), dim3(256), 0, n, (const OffT *)L.xadj, L.adjncy, L.adjwgt, L.vwgt, mr.own,
             mr.match, mr.bits, prv, cur, chunks, mr.rl2, counts, mr.part, mr.flags, r, gshort,
             r < mr.max_rounds ? 1 : 0, L.maxvw, mr.short_deg, (const int32_t *)mr.slist,
             (const unsigned *)mr.vcount);
    else if (r == 0 && r < mr.max_rounds)
      launch(ctx, "match_round", row_bytes, mr.fused_init ? k_round0<OffT, true> : k_round0<OffT, false>,
             dim3(gshort + glong), dim3(256), 0, n, (const OffT *)L.xadj, L.adjncy, L.adjwgt, L.vwgt, mr.own,
             mr.match, const_cast<int32_t *>(prv), cur, chunks, mr.rl, counts, mr.part, mr.flags, gshort, L.maxvw,
             mr.short_deg, 2 * (int64_t)mr.max_rounds + 16, ctx->d_ctr + HASH_CTR_SLOT, mr.d_n);
    else if (mr.fused_init)
      launch(ctx, "match_round", r == 0 ? row_bytes : 0.0, k_round_mesh<OffT>, dim3(gshort + glong), dim3(256), 0, n,
             (const OffT *)L.xadj, L.adjncy, L.adjwgt, L.vwgt, mr.own, mr.match, prv, cur, chunks, mr.rl, counts,
             mr.part, mr.flags, r, gshort, r < mr.max_rounds ? 1 : 0, L.maxvw, mr.short_deg, mr.d_n);
    else
      launch(ctx, "match_round", r == 0 ? row_bytes : 0.0,
             tk == TK_SPLIT ? k_round<OffT, TK_SPLIT> : k_round<OffT, TK_MATCH>,
             dim3(gshort + glong), dim3(256), 0, n, (const OffT *)L.xadj, L.adjncy, L.adjwgt, L.vwgt, mr.own,
             mr.match, mr.bits, prv, cur, chunks, mr.rl, counts, mr.part, mr.flags, r, gshort,
             r < mr.max_rounds ? 1 : 0, L.maxvw, mr.short_deg, mr.d_n);
    if (mr.ghub && r < mr.max_rounds)
      launch(ctx, "match_hub_reduce", 0.0, k_hub_reduce, dim3(mr.ghub), dim3(256), 0, chunks, counts, mr.part,
             cur, (const unsigned *)mr.flags, r);
    if (mr.compact && r == mr.compact_round)
      launch(ctx, "match_compact", 4.0 * n, k_compact, dim3(mr.gs + (1 + NCLASS) * LIST_BLOCKS), dim3(256), 0, n,
             (const int32_t *)mr.match, (const unsigned *)mr.flags, r, (const LongChunk *)mr.chunks, mr.rl,
             (const unsigned *)mr.nchunks, mr.chunks2, mr.rl2, mr.nchunks2, mr.slist, mr.vcount, mr.gs);
  }
}

// Every value the launches of match_launch_direct(count) take, as bytes: two
// batches with equal keys launch the same kernels with the same arguments.
template <typename OffT>
static std::string batch_key(const MatchRun &mr, int count)
{
  const Level &L = *mr.L;
  std::string k;
  auto put = [&](const auto &v) { k.append((const char *)&v, sizeof(v)); };
  put(sizeof(OffT));
  put(L.n);
  put(L.m);
  put(L.xadj);
  put(L.adjncy);
  put(L.adjwgt);
  put(L.vwgt);
  put(L.maxvw);
  put(mr.own);
  put(mr.max_rounds);
  put(mr.match);
  put(mr.prop);
  put(mr.bits);
  put(mr.chunks);
  put(mr.rl);
  put(mr.short_deg);
  put(mr.part);
  put(mr.flags);
  put(mr.nchunks);
  put(mr.g);
  put(mr.fused_init);
  put(mr.d_n);
  put(mr.half);
  put(mr.glong);
  put(mr.ghub);
  put(mr.compact);
  put(mr.compact_round);
  put(mr.chunks2);
  put(mr.rl2);
  put(mr.slist);
  put(mr.vcount);
  put(mr.nchunks2);
  put(mr.gs);
  put(mr.r);
  put(count);
  return k;
}

// A batch of round launches: directly, or (tune.graph_rounds, no per-kernel
// timing) as a replay of the hipGraph captured the first time a batch with
// the same key was queued.  A replay runs the same kernels with the same
// arguments over the buffers' current contents, and a graph of dependent
// nodes starts each kernel sooner than separate launches do.
constexpr size_t GRAPH_CACHE_MAX = 256;

template <typename OffT>
static void match_launch(cm_ctx *ctx, MatchRun &mr, int count)
{
  if (!ctx->tune.graph_rounds || ctx->timing || count < 2) {
    match_launch_direct<OffT>(ctx, mr, count);
    return;
  }
  const int r0 = mr.r;
  const std::string key = batch_key<OffT>(mr, count);
  auto it = ctx->graphs.find(key);
  if (it == ctx->graphs.end()) {
    if (ctx->graphs.size() >= GRAPH_CACHE_MAX) {
      sync(ctx);  // no replay may still be queued
      for (auto &kv : ctx->graphs) (void)hipGraphExecDestroy(kv.second);
      ctx->graphs.clear();
    }
    hipGraph_t g = nullptr;
    CM_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    try {
      match_launch_direct<OffT>(ctx, mr, count);
    } catch (...) {
      (void)hipStreamEndCapture(ctx->stream, &g);
      if (g) (void)hipGraphDestroy(g);
      mr.r = r0;
      throw;
    }
    CM_HIP(hipStreamEndCapture(ctx->stream, &g));
    hipGraphExec_t ge = nullptr;
    const hipError_t e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) throw Error(CM_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    it = ctx->graphs.emplace(key, ge).first;
    ctx->stats["round_graph_captures"].count += 1;
  } else {
    ctx->stats["round_graph_replays"].count += 1;
    // the launches are in the graph: advance the run as they would have
    for (int i = 0; i < count && mr.r <= mr.max_rounds; ++i) ++mr.r;
  }
  CM_HIP(hipGraphLaunch(it->second, ctx->stream));
}

// Queues the flags of every accepted round (0 .. r-2: round r-2 is accepted
// by launch r-1) into the pinned host buffer.
static void match_queue_flags(cm_ctx *ctx, MatchRun &mr)
{
  const size_t words = 2 * (size_t)std::max(mr.r - 1, 1);
  ensure_host_flags(ctx, words);
  CM_HIP(hipMemcpyAsync(ctx->h_flags, mr.flags, sizeof(unsigned) * words, hipMemcpyDeviceToHost, ctx->stream));
}

// After a synchronisation that follows match_queue_flags: did the queued
// launches finish the matching?
static bool match_complete(cm_ctx *ctx, const MatchRun &mr)
{
  if (mr.r > mr.max_rounds) return true;
  if (mr.r < 2) return false;
  const unsigned *h = ctx->h_flags;
  return h[2 * (mr.r - 2)] == 0 || h[2 * (mr.r - 2) + 1] == 0;
}

// Rounds that did work: round 0, then every round whose predecessor changed
// something and left active vertices.
static int match_rounds(cm_ctx *ctx, const MatchRun &mr)
{
  const int acc = std::min(mr.r - 1, mr.max_rounds);  // accepted rounds 0 .. acc-1
  const unsigned *h = ctx->h_flags;
  int rounds = 1;
  while (rounds < acc && h[2 * (rounds - 1)] != 0 && h[2 * (rounds - 1) + 1] != 0) ++rounds;
  return rounds;
}

MatchRun *match_begin(cm_ctx *ctx, const Level &L, const MatchOwners &own, int max_rounds, int32_t *d_match)
{
  MatchRun *mr = new MatchRun();
  mr->L = &L;
  mr->own = own;
  mr->max_rounds = max_rounds > 0 ? max_rounds : 4096;
  mr->match = d_match;
  if (L.n > 0) {
    try {
      if (L.off64) match_start<int64_t>(ctx, *mr);
      else match_start<int32_t>(ctx, *mr);
    } catch (...) {
      match_end(ctx, mr);
      throw;
    }
  }
  return mr;
}

MatchRun *match_begin_ahead(cm_ctx *ctx, const Level &bound, const MatchOwners &own, int max_rounds,
                            int32_t *d_match, const int64_t *d_n, int count)
{
  CM_REQUIRE(bound.n > 0 && bound.maxdeg >= 0 && bound.maxdeg <= LONG_DEG, CM_ERR_INVALID,
             "a matching queued ahead needs a level without long rows");
  MatchRun *mr = new MatchRun();
  mr->own_level = bound;
  mr->L = &mr->own_level;
  mr->own = own;
  mr->max_rounds = max_rounds > 0 ? max_rounds : 4096;
  mr->match = d_match;
  mr->d_n = d_n;
  // the counters are cleared by launch 0, behind the work already queued:
  // until match_adopt the context keeps its own account of them
  const bool ctr_zero = ctx->hash_ctr_zero;
  try {
    if (bound.off64) {
      match_start<int64_t>(ctx, *mr);
      match_launch_direct<int64_t>(ctx, *mr, count);
    } else {
      match_start<int32_t>(ctx, *mr);
      match_launch_direct<int32_t>(ctx, *mr, count);
    }
  } catch (...) {
    ctx->hash_ctr_zero = ctr_zero;
    match_end(ctx, mr);
    throw;
  }
  ctx->hash_ctr_zero = ctr_zero;
  return mr;
}

void match_adopt(cm_ctx *ctx, MatchRun *mr, const Level &L)
{
  Level &own = mr->own_level;
  own.n = L.n;
  own.m = L.m;
  own.maxdeg = L.maxdeg;
  own.maxvw = L.maxvw;
  mr->d_n = nullptr;
  mr->g = nblocks(L.n, 256);
  // the flags of the queued launches come with the next read-back
  const size_t words = 2 * (size_t)std::max(mr->r - 1, 1);
  ensure_host_flags(ctx, words);
  ctx->rb_flags = mr->flags;
  ctx->rb_nflags = (int64_t)words;
  ctx->hash_ctr_zero = true;  // launch 0 cleared them after every earlier use
}

void match_queue(cm_ctx *ctx, MatchRun *mr, int count, bool copy_flags)
{
  if (mr->L->n == 0) return;
  if (mr->L->off64) match_launch<int64_t>(ctx, *mr, count);
  else match_launch<int32_t>(ctx, *mr, count);
  if (copy_flags) {
    match_queue_flags(ctx, *mr);
  } else {  // the next read-back (readback_begin) copies them
    const size_t words = 2 * (size_t)std::max(mr->r - 1, 1);
    ensure_host_flags(ctx, words);
    ctx->rb_flags = mr->flags;
    ctx->rb_nflags = (int64_t)words;
  }
}

bool match_done(cm_ctx *ctx, MatchRun *mr)
{
  return mr->L->n == 0 || match_complete(ctx, *mr);
}

int match_finish(cm_ctx *ctx, MatchRun *mr)
{
  for (;;) {
    if (match_done(ctx, mr)) break;
    match_queue(ctx, mr, ROUND_BATCH);
    sync(ctx);
  }
  return mr->L->n == 0 ? 0 : match_rounds(ctx, *mr);
}

void match_copy_finished(cm_ctx *ctx, MatchRun *mr, int32_t *out)
{
  const int64_t n = mr->L->n;
  if (n > 0)
    launch(ctx, "match_finish", 8.0 * n, k_match_finish, dim3(nblocks(n, 256)), dim3(256), 0, n,
           (const int32_t *)mr->match, out);
}

void match_end(cm_ctx *ctx, MatchRun *mr)
{
  for (void *p : mr->blocks) ctx->pool.release(p);
  delete mr;
}

int match_handshake(cm_ctx *ctx, const Level &L, const MatchOwners &own, int max_rounds,
                    int32_t *d_match, bool finish)
{
  MatchRun *mr = match_begin(ctx, L, own, max_rounds, d_match);
  int rounds = 0;
  try {
    if (L.n > 0) {
      match_queue(ctx, mr, ROUND_BATCH);
      sync(ctx);
    }
    rounds = match_finish(ctx, mr);
    if (finish) match_copy_finished(ctx, mr, d_match);
  } catch (...) {
    match_end(ctx, mr);
    throw;
  }
  match_end(ctx, mr);
  return rounds;
}

int match_handshake(cm_ctx *ctx, const Level &L, int32_t maxvwgt, uint64_t seed, int max_rounds,
                    int32_t *d_match, bool finish)
{
  MatchOwners own{};
  own.maxv0 = maxvwgt;
  own.seed0 = seed;
  return match_handshake(ctx, L, own, max_rounds, d_match, finish);
}

}  // namespace cm
