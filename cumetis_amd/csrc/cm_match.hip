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
#include "cm_common.hpp"
#include "cm_internal.hpp"

namespace cm {

constexpr int LONG_DEG = 64;  // rows longer than this are scanned by a whole wave
constexpr int ROUND_BATCH = 9;  // launches queued between host checks

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

__global__ __launch_bounds__(256) void k_match_init(int64_t n, int32_t *match, int32_t *prop)
{
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v < n) {
    match[v] = -1;
    prop[v] = -1;
  }
}

template <typename OffT>
__global__ __launch_bounds__(256) void k_select_long(int64_t n, const OffT *xadj, int32_t *list,
                                                     unsigned *count)
{
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool is_long = v < n && (int64_t)(xadj[v + 1] - xadj[v]) > LONG_DEG;
  const uint64_t mk = __ballot(is_long);
  if (!mk) return;
  const int leader = __ffsll((unsigned long long)mk) - 1;
  unsigned base = 0;
  if (lane() == leader) base = atomicAdd(count, (unsigned)__popcll(mk));
  base = __shfl(base, leader, CM_WAVE);
  if (is_long) list[base + mask_rank(mk)] = (int32_t)v;
}

// Accepted in round r-1 (from its proposals alone): dead, or mutual.
__device__ __forceinline__ bool accepted_prev(const int32_t *prev, int32_t u)
{
  const int32_t p = prev[u];
  return p == -2 || (p >= 0 && prev[p] == u);
}

// Best eligible neighbour of v among vertices unmatched after round r-1, by
// the strict order (adjwgt, hash(seed,min,max), min, max); -2 if none.
// Lanes [l0, l0+step) stride the row; the caller reduces across lanes.
template <typename OffT>
__device__ __forceinline__ void scan_row(int64_t v, OffT b, OffT e, int l0, int step,
                                         const int32_t *__restrict__ adjncy,
                                         const int32_t *__restrict__ adjwgt,
                                         const int32_t *__restrict__ vwgt, int32_t vw, int32_t maxv,
                                         uint64_t seed, int64_t base, const int32_t *match,
                                         const int32_t *prev, bool have_prev, int32_t &best,
                                         uint64_t &bk1, uint64_t &bk2)
{
  for (OffT j = b + l0; j < e; j += step) {
    const int32_t u = adjncy[j];
    const uint32_t w = (uint32_t)adjwgt[j];
    if (best != -2 && w < (uint32_t)(bk1 >> 32)) continue;
    if (u == (int32_t)v || match[u] != -1 || vw + vwgt[u] > maxv) continue;
    const uint32_t a = (uint32_t)(min((int64_t)u, v) - base), bb = (uint32_t)(max((int64_t)u, v) - base);
    const uint64_t k1 = ((uint64_t)w << 32) | edge_hash(seed, a, bb);
    const uint64_t k2 = ((uint64_t)a << 32) | bb;
    // the two dependent gathers of accepted_prev only for edges that would
    // beat the current best (about ln(degree) of them for random keys)
    if ((best == -2 || key_better(k1, k2, bk1, bk2)) && !(have_prev && accepted_prev(prev, u))) {
      best = u;
      bk1 = k1;
      bk2 = k2;
    }
  }
}

// Launch r of the handshake: accept of round r-1 (from prev = proposals of
// round r-1) and proposals of round r (into cur), over the state after round
// r-1.  match[] is written only for the thread's own vertex; readers treat a
// neighbour as taken if match[u] != -1 or it was accepted in round r-1, which
// gives the same answer whether or not the neighbour's own write has landed.
// Blocks [0, nshort) take one vertex per thread (rows <= LONG_DEG and all
// accepts); the remaining blocks take long rows from `list`, one wave each.
template <typename OffT>
__global__ __launch_bounds__(256) void k_round(int64_t n, const OffT *__restrict__ xadj,
                                               const int32_t *__restrict__ adjncy,
                                               const int32_t *__restrict__ adjwgt,
                                               const int32_t *__restrict__ vwgt, MatchOwners own,
                                               int32_t *match, const int32_t *prev, int32_t *cur,
                                               const int32_t *list, const unsigned *list_n,
                                               unsigned *flags, int r, unsigned nshort, int propose)
{
  if (r >= 2 && !round_live(flags, r - 2)) return;
  const bool have_prev = r >= 1;
  if (blockIdx.x < nshort) {
    __shared__ int s_changed, s_active;
    if (threadIdx.x == 0) {
      s_changed = 0;
      s_active = 0;
    }
    __syncthreads();
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool changed = false, active = false;
    if (v < n && match[v] == -1) {
      bool done = false;
      if (have_prev) {
        const int32_t p = prev[v];
        if (p == -2) {
          match[v] = (int32_t)v;
          done = true;
        } else if (p >= 0 && prev[p] == (int32_t)v) {
          match[v] = p;
          done = true;
        }
      }
      changed = done;
      active = !done;
      if (!done && propose) {
        const OffT b = xadj[v], e = xadj[v + 1];
        if ((int64_t)(e - b) <= LONG_DEG) {
          const int32_t p = have_prev ? prev[v] : -1;
          if (p >= 0 && match[p] == -1 && !accepted_prev(prev, p)) {
            cur[v] = p;  // still unmatched, so still the best
          } else {
            int32_t maxv;
            uint64_t seed;
            int64_t base;
            owner_params(own, v, maxv, seed, base);
            int32_t best = -2;
            uint64_t bk1 = 0, bk2 = 0;
            scan_row<OffT>(v, b, e, 0, 1, adjncy, adjwgt, vwgt, vwgt[v], maxv, seed, base, match, prev,
                           have_prev, best, bk1, bk2);
            cur[v] = best;
          }
        }
      }
    }
    if (have_prev) {
      if (__ballot(changed) && lane() == 0) s_changed = 1;
      if (__ballot(active) && lane() == 0) s_active = 1;
      __syncthreads();
      // round status words are flags: one plain store per block, and only
      // while still clear, so blocks do not serialise on one address
      if (threadIdx.x == 0) {
        unsigned *f = flags + 2 * (r - 1);
        if (s_changed && __hip_atomic_load(&f[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0)
          __hip_atomic_store(&f[0], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (s_active && __hip_atomic_load(&f[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0)
          __hip_atomic_store(&f[1], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
    return;
  }
  if (!propose) return;
  // long rows: one wave per vertex
  const unsigned nl = *list_n;
  const int l = lane();
  const int64_t nw = (int64_t)(gridDim.x - nshort) * (blockDim.x / CM_WAVE);
  for (int64_t wi = ((int64_t)(blockIdx.x - nshort) * blockDim.x + threadIdx.x) / CM_WAVE; wi < nl; wi += nw) {
    const int32_t v = list[wi];
    if (match[v] != -1 || (have_prev && accepted_prev(prev, v))) continue;
    const int32_t p = have_prev ? prev[v] : -1;
    if (p >= 0 && match[p] == -1 && !accepted_prev(prev, p)) {
      if (l == 0) cur[v] = p;
      continue;
    }
    int32_t maxv;
    uint64_t seed;
    int64_t base;
    owner_params(own, v, maxv, seed, base);
    int32_t best = -2;
    uint64_t bk1 = 0, bk2 = 0;
    scan_row<OffT>(v, xadj[v], xadj[v + 1], l, CM_WAVE, adjncy, adjwgt, vwgt, vwgt[v], maxv, seed, base,
                   match, prev, have_prev, best, bk1, bk2);
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
    if (l == 0) cur[v] = best;
  }
}

__global__ __launch_bounds__(256) void k_match_finish(int64_t n, int32_t *match)
{
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v < n && match[v] == -1) match[v] = (int32_t)v;
}

template <typename OffT>
static int match_impl(cm_ctx *ctx, const Level &L, const MatchOwners &own, int max_rounds,
                      int32_t *match)
{
  const int64_t n = L.n;
  if (n == 0) return 0;
  const OffT *xadj = (const OffT *)L.xadj;
  int32_t *prop = (int32_t *)ctx->pool.alloc(sizeof(int32_t) * 2 * n);  // ping-pong
  int32_t *list = (int32_t *)ctx->pool.alloc(sizeof(int32_t) * n);
  unsigned *flags = (unsigned *)ctx->pool.alloc(sizeof(unsigned) * (2 * (size_t)max_rounds + 2));
  unsigned *list_n = flags + 2 * (size_t)max_rounds;
  fill_u32(ctx, flags, 0u, 2 * (int64_t)max_rounds + 2);
  const unsigned g = nblocks(n, 256);
  const double row_bytes = (double)sizeof(OffT) * (n + 1) + 8.0 * L.m + 8.0 * n;
  launch(ctx, "match_init", 8.0 * n, k_match_init, dim3(g), dim3(256), 0, n, match, prop);
  launch(ctx, "match_select_long", (double)sizeof(OffT) * (n + 1), k_select_long<OffT>, dim3(g),
         dim3(256), 0, n, xadj, list, list_n);
  const unsigned glong = 256;  // extra blocks whose waves stride over the long rows
  // launch r: accept of round r-1, proposals of round r (launch max_rounds
  // only accepts)
  int r = 0;
  std::vector<unsigned> h(2);
  for (;;) {
    for (int i = 0; i < ROUND_BATCH && r <= max_rounds; ++i, ++r) {
      int32_t *cur = prop + (size_t)(r & 1) * n;
      const int32_t *prv = prop + (size_t)((r + 1) & 1) * n;
      launch(ctx, "match_round", r == 0 ? row_bytes : 0.0, k_round<OffT>, dim3(g + glong), dim3(256), 0, n,
             xadj, L.adjncy, L.adjwgt, L.vwgt, own, match, prv, cur, (const int32_t *)list,
             (const unsigned *)list_n, flags, r, g, r < max_rounds ? 1 : 0);
    }
    if (r < 2) continue;
    // status of the last accepted round (r-2: accepted by launch r-1)
    CM_HIP(hipMemcpyAsync(h.data(), flags + 2 * (r - 2), sizeof(unsigned) * 2, hipMemcpyDeviceToHost,
                          ctx->stream));
    sync(ctx);
    if (h[0] == 0 || h[1] == 0 || r > max_rounds) break;
  }
  // rounds that did work: round 0, then every round whose predecessor changed
  // something and left active vertices
  const int acc = std::min(r - 1, max_rounds);  // accepted rounds 0 .. acc-1
  std::vector<unsigned> all(2 * (size_t)std::max(acc, 1));
  if (acc > 0)
    CM_HIP(hipMemcpyAsync(all.data(), flags, sizeof(unsigned) * 2 * acc, hipMemcpyDeviceToHost, ctx->stream));
  launch(ctx, "match_finish", 4.0 * n, k_match_finish, dim3(g), dim3(256), 0, n, match);
  sync(ctx);
  int rounds = 1;
  while (rounds < acc && all[2 * (rounds - 1)] != 0 && all[2 * (rounds - 1) + 1] != 0) ++rounds;
  ctx->pool.release(prop);
  ctx->pool.release(list);
  ctx->pool.release(flags);
  return rounds;
}

int match_handshake(cm_ctx *ctx, const Level &L, const MatchOwners &own, int max_rounds,
                    int32_t *d_match)
{
  if (max_rounds <= 0) max_rounds = 4096;
  return L.off64 ? match_impl<int64_t>(ctx, L, own, max_rounds, d_match)
                 : match_impl<int32_t>(ctx, L, own, max_rounds, d_match);
}

int match_handshake(cm_ctx *ctx, const Level &L, int32_t maxvwgt, uint64_t seed, int max_rounds,
                    int32_t *d_match)
{
  MatchOwners own{};
  own.maxv0 = maxvwgt;
  own.seed0 = seed;
  return match_handshake(ctx, L, own, max_rounds, d_match);
}

}  // namespace cm
