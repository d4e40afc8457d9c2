// cm_refine.hip -- one k-way refinement pass on the GPU with a deterministic
// move schedule (the reference's pass is cu:3244-3515 and races: its moves
// check balance and then update part weights with separate atomics,
// cu:3439-3443).
//   1. candidates: vertices whose external edge weight exceeds the internal
//      weight (cu:3244-3267);
//   2. destination: per candidate, the connectivity to every part is summed
//      in an LDS row owned by one wave; the destination is the first part
//      with the strictly largest connectivity, or the own part when none
//      beats it (cu:3272-3309);
//   3. candidates are sorted by (destination, vertex) with the segmented
//      sort (the reference sorts by destination, cu:3387);
//   4. one wave walks the sorted list on the device and accepts the moves
//      that keep both parts inside [int(w/1.03), int(w*1.03)] of their
//      target (cu:3401-3488): first the candidates with from <= to, then
//      those with from > to, destinations ascending, each destination's
//      candidates by vertex (k_accept_moves).
#include <vector>

#include "cm_common.hpp"
#include "cm_internal.hpp"
#include "cm_refine.hpp"

namespace cm {

constexpr int REFINE_MAX_PARTS = 2048;
constexpr int CAND_LONG = 256;  // k_refine_candidates: longer rows are summed by a wave

// xadj values index adjncy/adjwgt directly; neighbour ids index where_nb
// (for a batch slice these are the slice pointers shifted back by its bases)
template <typename OffT>
__global__ __launch_bounds__(256) void k_refine_candidates(int64_t n, const OffT *__restrict__ xadj,
                                                           const int32_t *__restrict__ adjncy,
                                                           const int32_t *__restrict__ adjwgt,
                                                           const int32_t *__restrict__ where,
                                                           const int32_t *__restrict__ where_nb,
                                                           int32_t *list, unsigned *count)
{
  __shared__ unsigned s_cnt, s_base;
  __shared__ unsigned s_w[4];
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  bool cand = false;
  int me = 0;
  OffT b = 0, e = 0;
  if (v < n) {
    me = where[v];
    b = xadj[v];
    e = xadj[v + 1];
  }
  // rows longer than CAND_LONG (R-MAT's hubs) are summed by the whole wave,
  // one after another; the others by their own lane
  const bool lng = v < n && e - b > (OffT)CAND_LONG;
  if (v < n && !lng) {
    int64_t in = 0, ex = 0;
    for (OffT j = b; j < e; ++j) {
      if (where_nb[adjncy[j]] == me) in += adjwgt[j];
      else ex += adjwgt[j];
    }
    cand = ex > in;
  }
  for (uint64_t lm = __ballot(lng); lm; lm &= lm - 1) {
    const int src = __ffsll((unsigned long long)lm) - 1;
    const OffT rb = __shfl(b, src, CM_WAVE), re = __shfl(e, src, CM_WAVE);
    const int rme = __shfl(me, src, CM_WAVE);
    int64_t in = 0, ex = 0;
    for (OffT j = rb + (OffT)lane(); j < re; j += CM_WAVE) {
      if (where_nb[adjncy[j]] == rme) in += adjwgt[j];
      else ex += adjwgt[j];
    }
    in = wave_sum<int64_t>(in);
    ex = wave_sum<int64_t>(ex);
    if (lane() == src) cand = ex > in;
  }
  const uint64_t mk = __ballot(cand);
  const int w = threadIdx.x >> 6;
  if (lane() == 0) s_w[w] = (unsigned)__popcll(mk);
  __syncthreads();
  if (threadIdx.x == 0) {
    s_cnt = s_w[0] + s_w[1] + s_w[2] + s_w[3];
    s_base = s_cnt ? atomicAdd(count, s_cnt) : 0u;
  }
  __syncthreads();
  if (cand) {
    unsigned before = 0;
    for (int i = 0; i < w; ++i) before += s_w[i];
    list[s_base + before + mask_rank(mk)] = (int32_t)v;
  }
}

// One wave per candidate: LDS connectivity row of nparts entries.
template <typename OffT>
__global__ __launch_bounds__(256) void k_refine_dest(const int32_t *list, const unsigned *count,
                                                     const OffT *__restrict__ xadj,
                                                     const int32_t *__restrict__ adjncy,
                                                     const int32_t *__restrict__ adjwgt,
                                                     const int32_t *__restrict__ where,
                                                     const int32_t *__restrict__ where_nb,
                                                     const int32_t *__restrict__ vwgt, int nparts,
                                                     uint64_t *key, uint32_t *idx, int32_t *from,
                                                     int32_t *vw)
{
  extern __shared__ int32_t conn_all[];
  const int w = threadIdx.x >> 6, l = lane();
  int32_t *conn = conn_all + w * nparts;
  const unsigned nc = *count;
  const int64_t nw = (int64_t)gridDim.x * 4;
  for (int p = l; p < nparts; p += CM_WAVE) conn[p] = 0;
  for (int64_t q = (int64_t)blockIdx.x * 4 + w; q < nc; q += nw) {
    const int32_t v = list[q];
    const int me = where[v];
    for (OffT j = xadj[v] + l; j < xadj[v + 1]; j += CM_WAVE) atomicAdd(&conn[where_nb[adjncy[j]]], adjwgt[j]);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    int32_t best = -1, bp = nparts;
    for (int p = l; p < nparts; p += CM_WAVE) {
      const int32_t c = conn[p];
      if (c > best) {
        best = c;
        bp = p;
      }
    }
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) {
      const int32_t ob = __shfl_xor(best, d, CM_WAVE);
      const int32_t op = __shfl_xor(bp, d, CM_WAVE);
      if (ob > best || (ob == best && op < bp)) {
        best = ob;
        bp = op;
      }
    }
    const int32_t own = conn[me];
    const int to = best > own ? bp : me;
    __builtin_amdgcn_wave_barrier();
    for (OffT j = xadj[v] + l; j < xadj[v + 1]; j += CM_WAVE) conn[where_nb[adjncy[j]]] = 0;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (l == 0) {
      // bit 63: the second pass of the walk (from > to), whose candidates
      // then sort after all of the first pass's
      key[q] = ((uint64_t)(me > to) << 63) | ((uint64_t)(uint32_t)to << 32) | (uint32_t)v;
      idx[q] = (uint32_t)q;
      from[q] = me;
      vw[q] = vwgt[v];
    }
  }
}

// The serial move walk of one refinement pass, done by one wave.  The
// candidates, sorted by (destination, vertex), are taken 64 at a time: lane i
// checks its move against the part weights as they would be if every earlier
// candidate of the chunk was accepted (their weights summed over the earlier
// lanes).  Up to the first lane that fails, that assumption holds, so those
// lanes are accepted and the first failing lane is rejected exactly as in
// the serial walk; the next chunk starts after it.  A chunk with no
// rejection advances by 64.  Candidates whose destination is already too
// heavy for them are rejected without stopping the chunk (see below).
// The serial walk's two passes (from <= to, then from > to; disjoint
// candidates, each in (destination, vertex) order) are one sweep: the sort
// key's top bit puts the second pass's candidates after the first's.  Moves
// within one part (no connectivity beats the own part) leave the weights
// unchanged and are not applied.  pw_io: part weights in and out; out[0]:
// moves applied.
// from_s, vw_s: source part and vertex weight of candidate q in sorted order
// (k_sorted_gather).  The next window's loads are issued before the current
// window's checks, for the full advance that most windows make.
__global__ __launch_bounds__(CM_WAVE) void k_accept_moves(const uint64_t *__restrict__ key,
                                                          const int32_t *__restrict__ from_s,
                                                          const int32_t *__restrict__ vw_s, unsigned nc, int nparts,
                                                          int64_t maxw, int64_t minw, int64_t *pw_io,
                                                          int32_t *where, unsigned *out)
{
  extern __shared__ unsigned long long pw[];  // nparts part weights (two's complement)
  __shared__ int s_to[CM_WAVE], s_me[CM_WAVE];
  __shared__ long long s_w[CM_WAVE];
  const int l = lane();
  for (int p = l; p < nparts; p += CM_WAVE) pw[p] = (unsigned long long)pw_io[p];
  __syncthreads();
  unsigned applied = 0;
  struct Cand {
    uint64_t k;
    int me, w;
  };
  auto fetch = [&](unsigned q) {
    Cand c{0, 0, 0};
    if (q < nc) {
      c.k = key[q];
      c.me = from_s[q];
      c.w = vw_s[q];
    }
    return c;
  };
  {
    unsigned q0 = 0;
    Cand cur = fetch((unsigned)l);
    while (q0 < nc) {
      const unsigned q = q0 + (unsigned)l;
      const Cand nxt = fetch(q0 + CM_WAVE + (unsigned)l);  // (used after a full advance)
      bool has = q < nc;
      const int to = (int)((cur.k >> 32) & 0x7FFFFFFFu), v = (int)(uint32_t)cur.k, me = cur.me;
      const long long w = cur.w;
      s_to[l] = has ? to : -1;
      s_me[l] = has ? me : -1;
      __syncthreads();
      // A candidate whose destination is already too heavy for it is rejected
      // whenever the walk reaches it, unless a candidate before it in this
      // window moves a vertex out of that part first (the candidates before
      // the window are in pw already): it then neither applies nor stops the
      // window.  Without this, parts at their bound (R-MAT, k = 256) made
      // every rejection a window of its own.
      // The same holds for a candidate whose source part is already too
      // light to give the vertex up, unless a candidate before it in this
      // window moves a vertex into that part first.
      bool lighten = false, heavier = false;
      for (int j = 0; j < l; ++j) {
        lighten |= s_me[j] == to && s_to[j] != to;
        heavier |= s_to[j] == me && s_me[j] != me;
      }
      if (has && ((!lighten && (long long)pw[to] + w > maxw) || (!heavier && (long long)pw[me] - w < minw))) {
        has = false;
        s_to[l] = -1;
        s_me[l] = -1;
      }
      s_w[l] = has ? w : 0;
      __syncthreads();
      long long dto = 0, dme = 0;
      for (int j = 0; j < l; ++j) {
        const int tj = s_to[j], mj = s_me[j];
        const long long wj = s_w[j];
        dto += (tj == to ? wj : 0) - (mj == to ? wj : 0);
        dme += (tj == me ? wj : 0) - (mj == me ? wj : 0);
      }
      const long long a = (long long)pw[to] + dto + w, b = (long long)pw[me] + dme - w;
      const bool ok = a >= minw && a <= maxw && b >= minw && b <= maxw;
      const uint64_t fail = __ballot(has && !ok);
      const int f = fail ? __ffsll((unsigned long long)fail) - 1 : CM_WAVE;
      const bool apply = has && l < f && me != to;
      __syncthreads();  // every lane has read the weights
      if (apply) {
        atomicAdd(&pw[to], (unsigned long long)w);
        atomicAdd(&pw[me], (unsigned long long)(-w));
        where[v] = to;
      }
      applied += (unsigned)__popcll(__ballot(apply));
      __syncthreads();
      if (f < CM_WAVE) {
        q0 += (unsigned)f + 1u;
        cur = fetch(q0 + (unsigned)l);
      } else {
        q0 += (unsigned)CM_WAVE;
        cur = nxt;
      }
    }
  }
  for (int p = l; p < nparts; p += CM_WAVE) pw_io[p] = (int64_t)pw[p];
  if (l == 0) out[0] = applied;
}

// Source part and vertex weight of the sorted candidates, in sorted order.
__global__ __launch_bounds__(256) void k_sorted_gather(unsigned nc, const uint32_t *__restrict__ idx,
                                                       const int32_t *__restrict__ from,
                                                       const int32_t *__restrict__ vw, int32_t *__restrict__ from_s,
                                                       int32_t *__restrict__ vw_s)
{
  const unsigned q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < nc) {
    const uint32_t i = idx[q];
    from_s[q] = from[i];
    vw_s[q] = vw[i];
  }
}

__global__ __launch_bounds__(256) void k_apply_moves(int64_t nm, const int32_t *moves, int32_t *where)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nm) where[moves[2 * i]] = moves[2 * i + 1];
}

// Balancing: vertices of overweight parts, one wave each.  Destination = the
// adjacent part with room and the largest connectivity (lowest id on ties),
// else `fallback` (the lightest part).  Key orders by gain, then vertex.
template <typename OffT>
__global__ __launch_bounds__(256) void k_balance_dest(int64_t n, const OffT *__restrict__ xadj,
                                                      const int32_t *__restrict__ adjncy,
                                                      const int32_t *__restrict__ adjwgt,
                                                      const int32_t *__restrict__ where,
                                                      const int32_t *__restrict__ where_nb,
                                                      const int32_t *__restrict__ vwgt, int nparts,
                                                      const int32_t *__restrict__ over,
                                                      const int32_t *__restrict__ room, int fallback,
                                                      uint64_t *key, int32_t *to_out, unsigned *count)
{
  extern __shared__ int32_t conn_all[];
  const int w = threadIdx.x >> 6, l = lane();
  int32_t *conn = conn_all + w * nparts;
  for (int p = l; p < nparts; p += CM_WAVE) conn[p] = 0;
  const int64_t nw = (int64_t)gridDim.x * 4;
  for (int64_t v = (int64_t)blockIdx.x * 4 + w; v < n; v += nw) {
    const int me = where[v];
    if (!over[me]) continue;
    for (OffT j = xadj[v] + l; j < xadj[v + 1]; j += CM_WAVE) atomicAdd(&conn[where_nb[adjncy[j]]], adjwgt[j]);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    int32_t best = -1, bp = nparts;
    for (int p = l; p < nparts; p += CM_WAVE) {
      const int32_t c = conn[p];
      if (room[p] && c > 0 && c > best) {
        best = c;
        bp = p;
      }
    }
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) {
      const int32_t ob = __shfl_xor(best, d, CM_WAVE);
      const int32_t op = __shfl_xor(bp, d, CM_WAVE);
      if (ob > best || (ob == best && op < bp)) {
        best = ob;
        bp = op;
      }
    }
    const int32_t own = conn[me];
    const int to = bp < nparts ? bp : fallback;
    const int64_t gain = (int64_t)(bp < nparts ? best : 0) - own;
    __builtin_amdgcn_wave_barrier();
    for (OffT j = xadj[v] + l; j < xadj[v + 1]; j += CM_WAVE) conn[where_nb[adjncy[j]]] = 0;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (l == 0 && to != me) {
      const unsigned q = atomicAdd(count, 1u);
      const int64_t g = gain > 0x7fffffffll ? 0x7fffffffll : (gain < -0x7fffffffll ? -0x7fffffffll : gain);
      const uint32_t gk = (uint32_t)(0x80000000ll - g);
      key[q] = ((uint64_t)gk << 32) | (uint32_t)v;
      to_out[(int64_t)v] = to;
    }
  }
}

template <typename OffT>
static int balance_impl(cm_ctx *ctx, const Level &L, int nparts, int32_t *where, std::vector<int64_t> &pw,
                        int64_t tvwgt)
{
  const int64_t n = L.n;
  const float tp = (float)(1.0 / nparts);
  const int64_t maxw = (int64_t)(tp * (float)tvwgt * 1.03f);  // the refinement's upper bound
  int moved_total = 0;
  std::vector<int32_t> hvw;
  for (int iter = 0; iter < 4; ++iter) {
    std::vector<int32_t> over(nparts, 0), room(nparts, 0);
    bool any = false;
    int fallback = 0;
    for (int p = 0; p < nparts; ++p) {
      over[p] = pw[p] > maxw;
      room[p] = pw[p] < maxw;
      any |= over[p] != 0;
      if (pw[p] < pw[fallback]) fallback = p;
    }
    if (!any) break;
    const OffT *xadj = (const OffT *)L.xadj;
    int32_t *d_over = (int32_t *)ctx->pool.alloc(4 * nparts), *d_room = (int32_t *)ctx->pool.alloc(4 * nparts);
    uint64_t *key = (uint64_t *)ctx->pool.alloc(8 * (n > 0 ? n : 1));
    int32_t *to = (int32_t *)ctx->pool.alloc(4 * (n > 0 ? n : 1));
    unsigned *count = ctx->d_ctr + 5;
    CM_HIP(hipMemcpyAsync(d_over, over.data(), 4 * nparts, hipMemcpyHostToDevice, ctx->stream));
    CM_HIP(hipMemcpyAsync(d_room, room.data(), 4 * nparts, hipMemcpyHostToDevice, ctx->stream));
    fill_u32(ctx, count, 0u, 1);
    const unsigned grid = (unsigned)std::min<int64_t>((n + 3) / 4, 16384);
    launch(ctx, "balance_dest", 0.0, k_balance_dest<OffT>, dim3(grid), dim3(256), sizeof(int32_t) * 4 * nparts, n,
           xadj, L.adjncy - L.ebase, L.adjwgt - L.ebase, (const int32_t *)where, (const int32_t *)(where - L.vbase),
           (const int32_t *)L.vwgt, nparts, (const int32_t *)d_over, (const int32_t *)d_room, fallback, key, to,
           count);
    unsigned hc = 0;
    CM_HIP(hipMemcpyAsync(&hc, count, 4, hipMemcpyDeviceToHost, ctx->stream));
    sync(ctx);
    int moved = 0;
    if (hc > 0) {
      int64_t *segoff = (int64_t *)ctx->pool.alloc(16);
      uint32_t *idx = (uint32_t *)ctx->pool.alloc(4 * hc);
      int64_t hseg[2] = {0, (int64_t)hc};
      CM_HIP(hipMemcpyAsync(segoff, hseg, sizeof(hseg), hipMemcpyHostToDevice, ctx->stream));
      fill_u32(ctx, idx, 0u, hc);
      segsort_u64(ctx, 1, segoff, true, hc, key, idx);  // gain descending, then vertex
      std::vector<uint64_t> hk(hc);
      std::vector<int32_t> hto(n), hwh(n);
      if (hvw.empty()) {
        hvw.resize(n);
        CM_HIP(hipMemcpyAsync(hvw.data(), L.vwgt, 4 * n, hipMemcpyDeviceToHost, ctx->stream));
      }
      CM_HIP(hipMemcpyAsync(hk.data(), key, 8 * hc, hipMemcpyDeviceToHost, ctx->stream));
      CM_HIP(hipMemcpyAsync(hto.data(), to, 4 * n, hipMemcpyDeviceToHost, ctx->stream));
      CM_HIP(hipMemcpyAsync(hwh.data(), where, 4 * n, hipMemcpyDeviceToHost, ctx->stream));
      sync(ctx);
      std::vector<int32_t> moves;
      for (unsigned q = 0; q < hc; ++q) {
        const int v = (int)(uint32_t)hk[q];
        const int from = hwh[v], dst = hto[v];
        if (pw[from] > maxw && pw[dst] + hvw[v] <= maxw) {
          pw[from] -= hvw[v];
          pw[dst] += hvw[v];
          moves.push_back(v);
          moves.push_back(dst);
        }
      }
      moved = (int)moves.size() / 2;
      if (moved) {
        int32_t *dm = (int32_t *)ctx->pool.alloc(4 * moves.size());
        CM_HIP(hipMemcpyAsync(dm, moves.data(), 4 * moves.size(), hipMemcpyHostToDevice, ctx->stream));
        launch(ctx, "balance_apply", 0.0, k_apply_moves, dim3(nblocks(moved, 256)), dim3(256), 0, (int64_t)moved,
               (const int32_t *)dm, where);
        sync(ctx);
        ctx->pool.release(dm);
      }
      ctx->pool.release(segoff);
      ctx->pool.release(idx);
    }
    ctx->pool.release(d_over);
    ctx->pool.release(d_room);
    ctx->pool.release(key);
    ctx->pool.release(to);
    moved_total += moved;
    if (moved == 0) break;
  }
  return moved_total;
}

int balance_level(cm_ctx *ctx, const Level &L, int nparts, int32_t *d_where, std::vector<int64_t> &pw,
                  int64_t tvwgt)
{
  if (L.n == 0 || nparts <= 1) return 0;
  CM_REQUIRE(nparts <= REFINE_MAX_PARTS, CM_ERR_UNSUPPORTED, "balancing supports up to 2048 parts");
  return L.off64 ? balance_impl<int64_t>(ctx, L, nparts, d_where, pw, tvwgt)
                 : balance_impl<int32_t>(ctx, L, nparts, d_where, pw, tvwgt);
}

template <typename OffT>
static int refine_impl(cm_ctx *ctx, const Level &L, int nparts, int32_t *where, std::vector<int64_t> &pw,
                       int64_t tvwgt)
{
  const int64_t n = L.n;
  if (n == 0 || nparts <= 1) return 0;
  CM_REQUIRE(nparts <= REFINE_MAX_PARTS, CM_ERR_UNSUPPORTED, "k-way refinement supports up to 2048 parts");
  const OffT *xadj = (const OffT *)L.xadj;
  int32_t *list = (int32_t *)ctx->pool.alloc(sizeof(int32_t) * n);
  unsigned *count = ctx->d_ctr + 4;
  fill_u32(ctx, count, 0u, 1);
  const int32_t *adj = L.adjncy - L.ebase, *wgt = L.adjwgt - L.ebase;
  const int32_t *where_nb = where - L.vbase;
  launch(ctx, "refine_candidates", 0.0, k_refine_candidates<OffT>, dim3(nblocks(n, 256)), dim3(256), 0, n,
         xadj, adj, wgt, (const int32_t *)where, where_nb, list, count);
  unsigned hc = 0;
  CM_HIP(hipMemcpyAsync(&hc, count, sizeof(unsigned), hipMemcpyDeviceToHost, ctx->stream));
  sync(ctx);
  int accepted = 0;
  if (hc > 0) {
    uint64_t *key = (uint64_t *)ctx->pool.alloc(sizeof(uint64_t) * hc);
    uint32_t *idx = (uint32_t *)ctx->pool.alloc(sizeof(uint32_t) * hc);
    int32_t *from = (int32_t *)ctx->pool.alloc(sizeof(int32_t) * hc);
    int32_t *vw = (int32_t *)ctx->pool.alloc(sizeof(int32_t) * hc);
    const unsigned grid = (unsigned)std::min<int64_t>((hc + 3) / 4, 16384);
    launch(ctx, "refine_dest", 0.0, k_refine_dest<OffT>, dim3(grid), dim3(256), sizeof(int32_t) * 4 * nparts,
           (const int32_t *)list, (const unsigned *)count, xadj, adj, wgt, (const int32_t *)where, where_nb,
           (const int32_t *)L.vwgt, nparts, key, idx, from, vw);
    int64_t *segoff = (int64_t *)ctx->pool.alloc(sizeof(int64_t) * 2);
    int64_t hseg[2] = {0, (int64_t)hc};
    CM_HIP(hipMemcpyAsync(segoff, hseg, sizeof(hseg), hipMemcpyHostToDevice, ctx->stream));
    segsort_u64(ctx, 1, segoff, true, hc, key, idx);
    // cu:3401-3412: limits from float targets 1/k and ub 1.03
    const float tp = (float)(1.0 / nparts);
    const int64_t maxw = (int)(tp * (float)tvwgt * 1.03f);
    const int64_t minw = (int)(tp * (float)tvwgt / 1.03f);
    int64_t *d_pw = (int64_t *)ctx->pool.alloc(sizeof(int64_t) * nparts);
    CM_HIP(hipMemcpyAsync(d_pw, pw.data(), sizeof(int64_t) * nparts, hipMemcpyHostToDevice, ctx->stream));
    unsigned *d_out = ctx->d_ctr + 6;
    int32_t *from_s = (int32_t *)ctx->pool.alloc(sizeof(int32_t) * hc);
    int32_t *vw_s = (int32_t *)ctx->pool.alloc(sizeof(int32_t) * hc);
    launch(ctx, "refine_gather", 0.0, k_sorted_gather, dim3(nblocks(hc, 256)), dim3(256), 0, hc,
           (const uint32_t *)idx, (const int32_t *)from, (const int32_t *)vw, from_s, vw_s);
    launch(ctx, "refine_accept", 0.0, k_accept_moves, dim3(1), dim3(CM_WAVE), sizeof(int64_t) * nparts,
           (const uint64_t *)key, (const int32_t *)from_s, (const int32_t *)vw_s, hc, nparts, maxw, minw, d_pw,
           where, d_out);
    unsigned hacc = 0;
    CM_HIP(hipMemcpyAsync(pw.data(), d_pw, sizeof(int64_t) * nparts, hipMemcpyDeviceToHost, ctx->stream));
    CM_HIP(hipMemcpyAsync(&hacc, d_out, sizeof(unsigned), hipMemcpyDeviceToHost, ctx->stream));
    sync(ctx);
    accepted = (int)hacc;
    ctx->pool.release(d_pw);
    ctx->pool.release(key);
    ctx->pool.release(idx);
    ctx->pool.release(from);
    ctx->pool.release(vw);
    ctx->pool.release(from_s);
    ctx->pool.release(vw_s);
    ctx->pool.release(segoff);
  }
  ctx->pool.release(list);
  return accepted;
}

int refine_level(cm_ctx *ctx, const Level &L, int nparts, int32_t *d_where, std::vector<int64_t> &pw,
                 int64_t tvwgt)
{
  return L.off64 ? refine_impl<int64_t>(ctx, L, nparts, d_where, pw, tvwgt)
                 : refine_impl<int32_t>(ctx, L, nparts, d_where, pw, tvwgt);
}

// A block owns 256 rows and strides over their entries, finding each entry's
// row by binary search in the block's offsets: a long row is spread over the
// block instead of one thread.
template <typename OffT>
__global__ __launch_bounds__(256) void k_level_cut(int64_t n, const OffT *__restrict__ xadj,
                                                   const int32_t *__restrict__ adjncy,
                                                   const int32_t *__restrict__ adjwgt,
                                                   const int32_t *__restrict__ where,
                                                   const int32_t *__restrict__ where_nb,
                                                   unsigned long long *sum)
{
  __shared__ int64_t s_off[257];
  __shared__ int32_t s_part[256];
  const int64_t v0 = (int64_t)blockIdx.x * 256;
  const int rows = (int)min<int64_t>(256, n - v0);
  const int t = threadIdx.x;
  if (t < rows) {
    s_off[t] = (int64_t)xadj[v0 + t];
    s_part[t] = where[v0 + t];
  }
  if (t == 0) s_off[rows] = (int64_t)xadj[v0 + rows];
  __syncthreads();
  int64_t cut = 0;
  for (int64_t e = s_off[0] + t; e < s_off[rows]; e += 256) {
    int lo = 0, hi = rows;  // s_off[lo] <= e < s_off[hi]
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (s_off[mid] <= e) lo = mid;
      else hi = mid;
    }
    if (where_nb[adjncy[e]] != s_part[lo]) cut += adjwgt[e];
  }
  cut = wave_sum<int64_t>(cut);
  if (lane() == 0 && cut) atomicAdd(sum, (unsigned long long)cut);
}

void level_cut(cm_ctx *ctx, const Level &L, const int32_t *d_where, unsigned long long *d_sum)
{
  fill_u64(ctx, d_sum, 0ull, 1);
  if (L.n == 0) return;
  const int32_t *adj = L.adjncy - L.ebase, *wgt = L.adjwgt - L.ebase;
  const int32_t *where_nb = d_where - L.vbase;
  const double bytes = (double)L.m * 12.0 + (double)L.n * (L.off64 ? 12.0 : 8.0);
  if (L.off64)
    launch(ctx, "level_cut", bytes, k_level_cut<int64_t>, dim3(nblocks(L.n, 256)), dim3(256), 0, L.n,
           (const int64_t *)L.xadj, adj, wgt, d_where, where_nb, d_sum);
  else
    launch(ctx, "level_cut", bytes, k_level_cut<int32_t>, dim3(nblocks(L.n, 256)), dim3(256), 0, L.n,
           (const int32_t *)L.xadj, adj, wgt, d_where, where_nb, d_sum);
}

}  // namespace cm
