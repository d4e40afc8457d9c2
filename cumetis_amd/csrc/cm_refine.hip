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
//   4. the host walks the sorted list once and accepts moves that keep both
//      parts inside [int(w/1.03), int(w*1.03)] of their target
//      (cu:3401-3488), destinations ascending; accepted moves are applied on
//      the device.
#include <vector>

#include "cm_common.hpp"
#include "cm_internal.hpp"
#include "cm_refine.hpp"
#include "host/cm_host.hpp"

namespace cm {

constexpr int REFINE_MAX_PARTS = 2048;

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
  if (v < n) {
    const int me = where[v];
    int64_t in = 0, ex = 0;
    for (OffT j = xadj[v]; j < xadj[v + 1]; ++j) {
      if (where_nb[adjncy[j]] == me) in += adjwgt[j];
      else ex += adjwgt[j];
    }
    cand = ex > in;
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
      key[q] = ((uint64_t)(uint32_t)to << 32) | (uint32_t)v;
      idx[q] = (uint32_t)q;
      from[q] = me;
      vw[q] = vwgt[v];
    }
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
    std::vector<uint64_t> hkey(hc);
    std::vector<uint32_t> hidx(hc);
    std::vector<int> hfrom(hc), hvw(hc), sfrom(hc), svw(hc);
    CM_HIP(hipMemcpyAsync(hkey.data(), key, 8 * hc, hipMemcpyDeviceToHost, ctx->stream));
    CM_HIP(hipMemcpyAsync(hidx.data(), idx, 4 * hc, hipMemcpyDeviceToHost, ctx->stream));
    CM_HIP(hipMemcpyAsync(hfrom.data(), from, 4 * hc, hipMemcpyDeviceToHost, ctx->stream));
    CM_HIP(hipMemcpyAsync(hvw.data(), vw, 4 * hc, hipMemcpyDeviceToHost, ctx->stream));
    sync(ctx);
    for (unsigned q = 0; q < hc; ++q) {
      sfrom[q] = hfrom[hidx[q]];
      svw[q] = hvw[hidx[q]];
    }
    std::vector<int> moves;
    accepted = cmh::schedule_moves(nparts, tvwgt, hkey, sfrom, svw, pw, moves);
    if (accepted > 0) {
      int32_t *dm = (int32_t *)ctx->pool.alloc(sizeof(int32_t) * moves.size());
      CM_HIP(hipMemcpyAsync(dm, moves.data(), 4 * moves.size(), hipMemcpyHostToDevice, ctx->stream));
      launch(ctx, "refine_apply", 0.0, k_apply_moves, dim3(nblocks(accepted, 256)), dim3(256), 0,
             (int64_t)accepted, (const int32_t *)dm, where);
      sync(ctx);
      ctx->pool.release(dm);
    }
    ctx->pool.release(key);
    ctx->pool.release(idx);
    ctx->pool.release(from);
    ctx->pool.release(vw);
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

}  // namespace cm
