// cm_batch.hip -- batched coarsening of many graphs of uneven sizes.
//
// The graphs still coarsening form one disjoint union per level, stored as
// one CSR with global vertex ids, so each level is a single pass of every
// matching, numbering and contraction kernel over the whole batch.  The
// canonical numbering of a union restricted to one graph is that graph's own
// numbering shifted by a constant (representatives are ordered by vertex id
// and every graph owns a contiguous id range), and the contraction of a union
// is the union of the contractions, so with per-graph maxvwgt, seed and
// graph-local ids in the matching's edge order, every graph's levels are
// bit-identical to coarsening it alone.  After each level the graphs whose
// own stop rule (cu:1843-1845) ends them leave the union; the rest are
// compacted into the next level's union.
#include <algorithm>
#include <memory>
#include <vector>

#include "cm_common.hpp"
#include "cm_internal.hpp"

namespace cm {

int ref_threshold(int64_t n, int nparts);

namespace {

// slot of position x given ascending bases (nslots + 1 entries)
__device__ __forceinline__ int find_slot(const int64_t *base, int nslots, int64_t x)
{
  int lo = 0, hi = nslots;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (base[mid] <= x) lo = mid;
    else hi = mid;
  }
  return lo;
}

// Validation of a concatenated input before any id is made global: for
// every vertex v of graph s (bad |= 1 offsets, 2 id outside [0, n_s), 4 self
// loop, 8 weight <= 0): the row lies inside the graph's entries
// [eoff[s], eoff[s+1]) and ascends, the graph's first row starts at eoff[s],
// and every local id is a vertex of the graph other than v itself.  Thread N
// checks xadj[N] == M.  Empty graphs own no vertex: find_slot gives the
// graph that does.
template <typename OffT>
__global__ __launch_bounds__(256) void k_check_batch(int64_t N, int64_t M, const OffT *xadj, const int32_t *adj_local,
                                                     const int32_t *wgt, const int64_t *voff, const int64_t *eoff,
                                                     int nb, unsigned *bad)
{
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v > N) return;
  unsigned f = 0;
  if (v == N) {
    if ((int64_t)xadj[N] != M) f |= 1;
  } else {
    const int s = find_slot(voff, nb, v);
    const int64_t b = (int64_t)xadj[v], e = (int64_t)xadj[v + 1];
    const int64_t lo = eoff[s], hi = eoff[s + 1];
    if (v == voff[s] && b != lo) f |= 1;
    if (b > e || b < lo || e > hi) {
      f |= 1;
    } else {
      const int64_t nv = voff[s + 1] - voff[s], self = v - voff[s];
      for (int64_t j = b; j < e; ++j) {
        const int32_t u = adj_local[j];
        if (u < 0 || u >= nv) f |= 2;
        else if (u == self) f |= 4;
        if (wgt && wgt[j] <= 0) f |= 8;
      }
    }
  }
  if (f) atomicOr(bad, f);
}

// level 0: local -> global neighbour ids (adj_global may be adj_local).
// Thread t of a block takes entries base + t + k * 256 (k < GLOB_PER): one
// slot search for the first, then the slot advances with the entries.
constexpr int GLOB_PER = 8;
__global__ __launch_bounds__(256) void k_globalize(int64_t m, const int32_t *adj_local,
                                                   const int64_t *eoff, const int64_t *voff, int nb,
                                                   int32_t *adj_global)
{
  const int64_t e0 = (int64_t)blockIdx.x * 256 * GLOB_PER + threadIdx.x;
  if (e0 >= m) return;
  int s = find_slot(eoff, nb, e0);
#pragma unroll
  for (int k = 0; k < GLOB_PER; ++k) {
    const int64_t e = e0 + (int64_t)k * 256;
    if (e >= m) break;
    while (eoff[s + 1] <= e) ++s;
    adj_global[e] = (int32_t)(adj_local[e] + voff[s]);
  }
}

// Total vertex weight of every graph (one block per graph, grids stride).
__global__ __launch_bounds__(1024) void k_graph_vw_sums(int nb, const int64_t *voff, const int32_t *vw,
                                                        int64_t *sums)
{
  __shared__ int64_t sh[16];
  for (int b = blockIdx.x; b < nb; b += gridDim.x) {
    int64_t acc = 0;
    for (int64_t i = voff[b] + threadIdx.x; i < voff[b + 1]; i += 1024) acc += vw[i];
    acc = wave_sum<int64_t>(acc);
    if (lane() == 0) sh[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
      int64_t t = 0;
      for (int w = 0; w < 16; ++w) t += sh[w];
      sums[b] = t;
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) void k_slots(int64_t n, const int64_t *voff, int nb, int32_t *gid)
{
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v < n) gid[v] = find_slot(voff, nb, v);
}

// per slot: first coarse vertex and first coarse entry of the graph
template <typename OffT>
__global__ void k_slot_bases(int ns, const int64_t *vbase, const int32_t *cmap, const OffT *cxadj,
                             int64_t nc, int64_t *cvb, int64_t *ceb)
{
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s > ns) return;
  int64_t c;
  if (s == ns) c = nc;
  else c = cmap[vbase[s]];  // the first vertex of a graph is a representative
  cvb[s] = c;
  ceb[s] = (int64_t)cxadj[c];
}

// next union from the continuing slots: ids and offsets shifted per slot
template <typename OffT>
__global__ __launch_bounds__(256) void k_compact_vertices(int64_t n_new, const int64_t *nvb, int ns,
                                                          const int64_t *ovb, const int64_t *oeb,
                                                          const int64_t *neb, const OffT *oxadj,
                                                          const int32_t *ovwgt, OffT *nxadj,
                                                          int32_t *nvwgt, int32_t *ngid)
{
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v > n_new) return;
  if (v == n_new) {
    nxadj[v] = (OffT)neb[ns];
    return;
  }
  const int s = find_slot(nvb, ns, v);
  const int64_t src = ovb[s] + (v - nvb[s]);
  nxadj[v] = (OffT)((int64_t)oxadj[src] - oeb[s] + neb[s]);
  nvwgt[v] = ovwgt[src];
  ngid[v] = s;
}

__global__ __launch_bounds__(256) void k_compact_edges(int64_t m_new, const int64_t *neb, int ns,
                                                       const int64_t *oeb, const int64_t *ovb,
                                                       const int64_t *nvb, const int32_t *oadj,
                                                       const int32_t *owgt, int32_t *nadj,
                                                       int32_t *nwgt)
{
  // GLOB_PER entries per thread, the slot advancing with them (as k_globalize)
  const int64_t e0 = (int64_t)blockIdx.x * 256 * GLOB_PER + threadIdx.x;
  if (e0 >= m_new) return;
  int s = find_slot(neb, ns, e0);
#pragma unroll
  for (int k = 0; k < GLOB_PER; ++k) {
    const int64_t e = e0 + (int64_t)k * 256;
    if (e >= m_new) break;
    while (neb[s + 1] <= e) ++s;
    const int64_t src = oeb[s] + (e - neb[s]);
    nadj[e] = (int32_t)(oadj[src] - ovb[s] + nvb[s]);
    nwgt[e] = owgt[src];
  }
}

struct Slot {
  int graph;       // index into the caller's arrays
  int64_t n, m;    // sizes at the current level
  int T;
  int32_t maxv;
  uint64_t seed;
};

}  // namespace

template <typename OffT>
static void batch_impl(cm_ctx *ctx, int nb, const cm_graph_view *g, const int64_t *voff, const int64_t *eoff,
                       const cm_coarsen_params *params, cm_hier **out)
{
  const bool off64 = sizeof(OffT) == 8;
  auto shared = std::make_shared<SharedBlocks>();
  shared->ctx = ctx;
  auto own = [&](size_t bytes) {
    void *p = ctx->pool.alloc(bytes ? bytes : 1);
    shared->blocks.push_back(p);
    return p;
  };
  auto tmp_blocks = std::vector<void *>();
  auto tmp = [&](size_t bytes) {
    void *p = ctx->pool.alloc(bytes ? bytes : 1);
    tmp_blocks.push_back(p);
    return p;
  };
  const int64_t N = voff[nb], M = eoff[nb];
  // ---- level-0 union on the device.  A device input is referenced where
  // it can be (offsets and weights are only read; the caller keeps them alive
  // with the hierarchies, as for cm_coarsen); neighbour ids are made global
  // in an array of their own.
  const bool dev = g->on_device != 0;
  OffT *xadj = dev ? (OffT *)const_cast<void *>(g->xadj) : (OffT *)own(sizeof(OffT) * (N + 1));
  int32_t *adj = (int32_t *)own(4 * M);
  int32_t *wgt = dev && g->adjwgt ? const_cast<int32_t *>(g->adjwgt) : (int32_t *)own(4 * M);
  int32_t *vw = dev && g->vwgt ? const_cast<int32_t *>(g->vwgt) : (int32_t *)own(4 * N);
  int64_t *d_voff = (int64_t *)tmp(8 * (nb + 1)), *d_eoff = (int64_t *)tmp(8 * (nb + 1));
  CM_HIP(hipMemcpyAsync(d_voff, voff, 8 * (nb + 1), hipMemcpyHostToDevice, ctx->stream));
  CM_HIP(hipMemcpyAsync(d_eoff, eoff, 8 * (nb + 1), hipMemcpyHostToDevice, ctx->stream));
  if (!dev) CM_HIP(hipMemcpyAsync(xadj, g->xadj, sizeof(OffT) * (N + 1), hipMemcpyHostToDevice, ctx->stream));
  // local ids: read in place from a device input, copied into the union
  // array (and made global there) from a host input
  const int32_t *adj_local = g->adjncy;
  if (!dev) {
    if (M) CM_HIP(hipMemcpyAsync(adj, g->adjncy, 4 * M, hipMemcpyHostToDevice, ctx->stream));
    adj_local = adj;
  }
  if (!g->adjwgt) fill_u32(ctx, wgt, 1u, M);
  else if (!dev && M) CM_HIP(hipMemcpyAsync(wgt, g->adjwgt, 4 * M, hipMemcpyHostToDevice, ctx->stream));
  if (!g->vwgt) fill_u32(ctx, vw, 1u, N);
  else if (!dev && N) CM_HIP(hipMemcpyAsync(vw, g->vwgt, 4 * N, hipMemcpyHostToDevice, ctx->stream));
  bool trusted = true;
  for (int b = 0; b < nb; ++b) trusted = trusted && (params[b].flags & CM_FLAG_TRUSTED_INPUT);
  if (!trusted) {
    unsigned *bad = ctx->d_ctr + 3;
    fill_u32(ctx, bad, 0u, 1);
    launch(ctx, "validate_batch", 0.0, k_check_batch<OffT>, dim3(nblocks(N + 1, 256)), dim3(256), 0, N, M,
           (const OffT *)xadj, (const int32_t *)adj_local, (const int32_t *)(g->adjwgt ? wgt : nullptr),
           (const int64_t *)d_voff, (const int64_t *)d_eoff, nb, bad);
    unsigned h = 0;
    CM_HIP(hipMemcpyAsync(&h, bad, sizeof(unsigned), hipMemcpyDeviceToHost, ctx->stream));
    sync(ctx);
    CM_REQUIRE(!(h & 1), CM_ERR_INPUT, "batch offsets: a graph's rows must lie in [eoff[b], eoff[b+1]) and ascend");
    CM_REQUIRE(!(h & 2), CM_ERR_INPUT, "batch adjncy holds a local vertex id outside [0, n_b)");
    CM_REQUIRE(!(h & 4), CM_ERR_INPUT, "a graph of the batch has a self loop");
    CM_REQUIRE(!(h & 8), CM_ERR_INPUT, "edge weights must be positive");
  }
  launch(ctx, "batch_globalize", 0.0, k_globalize, dim3(nblocks(M, 256 * GLOB_PER)), dim3(256), 0, M,
         adj_local, (const int64_t *)d_eoff, (const int64_t *)d_voff, nb, adj);
  // per-graph totals of vertex weights, summed on the device
  std::vector<int64_t> tvw(nb, 0);
  if (g->vwgt && N) {
    int64_t *d_sums = (int64_t *)tmp(8 * nb);
    launch(ctx, "batch_vw_sums", 4.0 * N, k_graph_vw_sums, dim3((unsigned)std::min(nb, 4096)), dim3(1024), 0, nb,
           (const int64_t *)d_voff, (const int32_t *)vw, d_sums);
    CM_HIP(hipMemcpyAsync(tvw.data(), d_sums, 8 * nb, hipMemcpyDeviceToHost, ctx->stream));
  }
  queue_max_degree(ctx, N, off64, xadj, vw);
  sync(ctx);
  const int64_t maxdeg0 = ctx->h_small[4], maxvw0 = ctx->h_small[3];
  // ---- slots (graphs with vertices); empty graphs get two empty levels
  std::vector<Slot> slots;
  for (int b = 0; b < nb; ++b) {
    out[b] = new cm_hier();
    out[b]->shared = shared;
    const int64_t n = voff[b + 1] - voff[b], m = eoff[b + 1] - eoff[b];
    const int64_t tv = g->vwgt ? tvw[b] : n;
    const cm_coarsen_params &p = params[b];
    const int T = p.threshold > 0 ? p.threshold : ref_threshold(n, p.nparts);
    const int32_t maxv = p.maxvwgt > 0 ? p.maxvwgt : (int32_t)(1.5 * (double)tv / T);
    Level L0;
    L0.n = n;
    L0.m = m;
    L0.off64 = off64;
    L0.xadj = (char *)xadj + sizeof(OffT) * voff[b];
    L0.adjncy = adj + eoff[b];
    L0.adjwgt = wgt + eoff[b];
    L0.vwgt = vw + voff[b];
    L0.tvwgt = tv;
    L0.vbase = voff[b];
    L0.ebase = eoff[b];
    L0.maxdeg = maxdeg0;
    L0.maxvw = maxvw0;  // of the union: a bound for each graph
    out[b]->levels.push_back(L0);
    if (n == 0) {
      Level E = L0;
      E.m = 0;
      out[b]->levels.push_back(E);
      continue;
    }
    slots.push_back(Slot{b, n, m, T, maxv, p.seed});
  }
  // ---- level loop over the union of active slots
  int64_t n = N, m = M;
  int32_t *gid = (int32_t *)own(4 * (N > 0 ? N : 1));
  {
    // union of active slots at level 0 == the input with empty graphs absent;
    // slot ids of vertices index `slots`
    std::vector<int64_t> vb, eb;
    for (auto &s : slots) {
      vb.push_back(out[s.graph]->levels[0].vbase);
      eb.push_back(out[s.graph]->levels[0].ebase);
    }
    vb.push_back(N);
    eb.push_back(M);
    int64_t *d_vb = (int64_t *)tmp(8 * vb.size());
    CM_HIP(hipMemcpyAsync(d_vb, vb.data(), 8 * vb.size(), hipMemcpyHostToDevice, ctx->stream));
    launch(ctx, "batch_slots", 0.0, k_slots, dim3(nblocks(N, 256)), dim3(256), 0, N, (const int64_t *)d_vb,
           (int)slots.size(), gid);
  }
  const int max_levels = params[0].max_levels;
  int level = 0;
  while (!slots.empty()) {
    const int ns = (int)slots.size();
    std::vector<int64_t> vb(ns + 1), eb(ns + 1);
    std::vector<int32_t> hmaxv(ns);
    std::vector<uint64_t> hseed(ns);
    for (int s = 0; s < ns; ++s) {
      const Level &L = out[slots[s].graph]->levels.back();
      vb[s] = L.vbase;
      eb[s] = L.ebase;
      hmaxv[s] = slots[s].maxv;
      hseed[s] = slots[s].seed;
    }
    vb[ns] = n;
    eb[ns] = m;
    int64_t *d_vb = (int64_t *)tmp(8 * (ns + 1)), *d_eb = (int64_t *)tmp(8 * (ns + 1));
    int32_t *d_maxv = (int32_t *)tmp(4 * ns);
    uint64_t *d_seed = (uint64_t *)tmp(8 * ns);
    CM_HIP(hipMemcpyAsync(d_vb, vb.data(), 8 * (ns + 1), hipMemcpyHostToDevice, ctx->stream));
    CM_HIP(hipMemcpyAsync(d_eb, eb.data(), 8 * (ns + 1), hipMemcpyHostToDevice, ctx->stream));
    CM_HIP(hipMemcpyAsync(d_maxv, hmaxv.data(), 4 * ns, hipMemcpyHostToDevice, ctx->stream));
    CM_HIP(hipMemcpyAsync(d_seed, hseed.data(), 8 * ns, hipMemcpyHostToDevice, ctx->stream));
    Level U;  // the union
    U.n = n;
    U.m = m;
    U.off64 = off64;
    U.xadj = xadj;
    U.adjncy = adj;
    U.adjwgt = wgt;
    U.vwgt = vw;
    U.maxdeg = 0;
    U.maxvw = 0;
    for (int s = 0; s < ns; ++s) {
      const Level &S = out[slots[s].graph]->levels.back();
      U.maxdeg = (U.maxdeg < 0 || S.maxdeg < 0) ? -1 : std::max(U.maxdeg, S.maxdeg);
      U.maxvw = (U.maxvw < 0 || S.maxvw < 0) ? -1 : std::max(U.maxvw, S.maxvw);
    }
    int32_t *match = (int32_t *)own(4 * n), *cmap = (int32_t *)own(4 * n);
    MatchOwners mo;
    mo.gid = gid;
    mo.maxv = d_maxv;
    mo.seed = d_seed;
    mo.vbase = d_vb;
    const int rounds = match_handshake(ctx, U, mo, params[0].max_rounds, match, false);  // number() completes it
    int32_t *rep = (int32_t *)tmp(4 * n);
    OffT *segoff = (OffT *)tmp(sizeof(OffT) * (n + 1));
    int32_t *cvw = (int32_t *)own(4 * n);
    const int64_t nc = number(ctx, U, match, cmap, rep, segoff, cvw);
    OffT *cxadj = (OffT *)own(sizeof(OffT) * (nc + 1));
    int32_t *cadj = (int32_t *)own(4 * (m > 0 ? m : 1)), *cwgt = (int32_t *)own(4 * (m > 0 ? m : 1));
    int64_t cmaxdeg = -1;
    const int64_t mc = contract(ctx, U, match, cmap, nc, rep, segoff, params[0].row_order, cxadj, cadj, cwgt,
                                &cmaxdeg);
    const int64_t cmaxvw = ctx->h_small[3];
    // per-slot coarse bases
    int64_t *d_cvb = (int64_t *)tmp(8 * (ns + 1)), *d_ceb = (int64_t *)tmp(8 * (ns + 1));
    launch(ctx, "batch_bases", 0.0, k_slot_bases<OffT>, dim3(nblocks(ns + 1, 256)), dim3(256), 0, ns,
           (const int64_t *)d_vb, (const int32_t *)cmap, (const OffT *)cxadj, nc, d_cvb, d_ceb);
    std::vector<int64_t> cvb(ns + 1), ceb(ns + 1);
    CM_HIP(hipMemcpyAsync(cvb.data(), d_cvb, 8 * (ns + 1), hipMemcpyDeviceToHost, ctx->stream));
    CM_HIP(hipMemcpyAsync(ceb.data(), d_ceb, 8 * (ns + 1), hipMemcpyDeviceToHost, ctx->stream));
    sync(ctx);
    ++level;
    // record the level in every active graph, decide who continues
    std::vector<int> keep;
    for (int s = 0; s < ns; ++s) {
      cm_hier *h = out[slots[s].graph];
      Level &F = h->levels.back();
      F.match = match + vb[s];
      F.cmap = cmap + vb[s];
      F.cbase = cvb[s];
      F.rounds = rounds;
      Level C;
      C.n = cvb[s + 1] - cvb[s];
      C.m = ceb[s + 1] - ceb[s];
      C.off64 = off64;
      C.xadj = (char *)cxadj + sizeof(OffT) * cvb[s];
      C.adjncy = cadj + ceb[s];
      C.adjwgt = cwgt + ceb[s];
      C.vwgt = cvw + cvb[s];
      C.tvwgt = F.tvwgt;
      C.vbase = cvb[s];
      C.ebase = ceb[s];
      C.maxdeg = cmaxdeg;  // of the union: a bound for each slot
      C.maxvw = cmaxvw;
      const int64_t nf = slots[s].n;
      h->levels.push_back(C);
      const bool more = C.n > slots[s].T && C.n < 0.75 * nf && C.m > C.n / 2;
      const bool cap = max_levels > 0 && (int)h->levels.size() > max_levels;
      slots[s].n = C.n;
      slots[s].m = C.m;
      if (more && !cap) keep.push_back(s);
    }
    for (void *q : tmp_blocks) ctx->pool.release(q);
    tmp_blocks.clear();
    if (keep.empty()) break;
    if ((int)keep.size() == ns) {  // everyone continues: the coarse union is the next union
      int32_t *ngid = (int32_t *)own(4 * (nc > 0 ? nc : 1));
      // slot of each coarse vertex: same slots, coarse bases
      int64_t *d_cb = (int64_t *)tmp(8 * (ns + 1));
      CM_HIP(hipMemcpyAsync(d_cb, cvb.data(), 8 * (ns + 1), hipMemcpyHostToDevice, ctx->stream));
      launch(ctx, "batch_slots", 0.0, k_slots, dim3(nblocks(nc, 256)), dim3(256), 0, nc,
             (const int64_t *)d_cb, ns, ngid);
      xadj = cxadj;
      adj = cadj;
      wgt = cwgt;
      vw = cvw;
      gid = ngid;
      n = nc;
      m = mc;
      continue;
    }
    // compaction of the continuing slots
    const int nk = (int)keep.size();
    std::vector<int64_t> nvb(nk + 1), neb(nk + 1), ovb(nk), oeb(nk);
    std::vector<Slot> next;
    nvb[0] = neb[0] = 0;
    for (int i = 0; i < nk; ++i) {
      const int s = keep[i];
      ovb[i] = cvb[s];
      oeb[i] = ceb[s];
      nvb[i + 1] = nvb[i] + (cvb[s + 1] - cvb[s]);
      neb[i + 1] = neb[i] + (ceb[s + 1] - ceb[s]);
      next.push_back(slots[s]);
    }
    const int64_t n2 = nvb[nk], m2 = neb[nk];
    int64_t *d_nvb = (int64_t *)tmp(8 * (nk + 1)), *d_neb = (int64_t *)tmp(8 * (nk + 1));
    int64_t *d_ovb = (int64_t *)tmp(8 * nk), *d_oeb = (int64_t *)tmp(8 * nk);
    CM_HIP(hipMemcpyAsync(d_nvb, nvb.data(), 8 * (nk + 1), hipMemcpyHostToDevice, ctx->stream));
    CM_HIP(hipMemcpyAsync(d_neb, neb.data(), 8 * (nk + 1), hipMemcpyHostToDevice, ctx->stream));
    CM_HIP(hipMemcpyAsync(d_ovb, ovb.data(), 8 * nk, hipMemcpyHostToDevice, ctx->stream));
    CM_HIP(hipMemcpyAsync(d_oeb, oeb.data(), 8 * nk, hipMemcpyHostToDevice, ctx->stream));
    OffT *x2 = (OffT *)own(sizeof(OffT) * (n2 + 1));
    int32_t *a2 = (int32_t *)own(4 * (m2 > 0 ? m2 : 1)), *w2 = (int32_t *)own(4 * (m2 > 0 ? m2 : 1));
    int32_t *v2 = (int32_t *)own(4 * (n2 > 0 ? n2 : 1)), *g2 = (int32_t *)own(4 * (n2 > 0 ? n2 : 1));
    launch(ctx, "batch_compact_v", 0.0, k_compact_vertices<OffT>, dim3(nblocks(n2 + 1, 256)), dim3(256), 0,
           n2, (const int64_t *)d_nvb, nk, (const int64_t *)d_ovb, (const int64_t *)d_oeb,
           (const int64_t *)d_neb, (const OffT *)cxadj, (const int32_t *)cvw, x2, v2, g2);
    launch(ctx, "batch_compact_e", 0.0, k_compact_edges, dim3(nblocks(m2, 256 * GLOB_PER)), dim3(256), 0, m2,
           (const int64_t *)d_neb, nk, (const int64_t *)d_oeb, (const int64_t *)d_ovb,
           (const int64_t *)d_nvb, (const int32_t *)cadj, (const int32_t *)cwgt, a2, w2);
    // the continuing graphs' current level now lives in the compacted copy
    for (int i = 0; i < nk; ++i) {
      Level &C = out[next[i].graph]->levels.back();
      C.xadj = (char *)x2 + sizeof(OffT) * nvb[i];
      C.adjncy = a2 + neb[i];
      C.adjwgt = w2 + neb[i];
      C.vwgt = v2 + nvb[i];
      C.vbase = nvb[i];
      C.ebase = neb[i];
    }
    sync(ctx);
    for (void *q : tmp_blocks) ctx->pool.release(q);
    tmp_blocks.clear();
    slots.swap(next);
    xadj = x2;
    adj = a2;
    wgt = w2;
    vw = v2;
    gid = g2;
    n = n2;
    m = m2;
  }
  sync(ctx);
  for (void *q : tmp_blocks) ctx->pool.release(q);
}

}  // namespace cm

extern "C" int cm_coarsen_batch(cm_ctx *ctx, int32_t nb, const cm_graph_view *concat, const int64_t *voff,
                                const int64_t *eoff, const cm_coarsen_params *params, cm_hier **out)
{
  if (!ctx || !concat || !voff || !eoff || !params || !out || nb < 0) return CM_ERR_INVALID;
  for (int b = 0; b < nb; ++b) out[b] = nullptr;
  try {
    CM_HIP(hipSetDevice(ctx->device));
    CM_REQUIRE(concat->offset_type == CM_OFF32 || concat->offset_type == CM_OFF64, CM_ERR_INVALID,
               "offset_type must be CM_OFF32 or CM_OFF64");
    cm::validate_view(concat);
    CM_REQUIRE(voff[0] == 0 && eoff[0] == 0 && voff[nb] == concat->n && eoff[nb] == concat->m,
               CM_ERR_INVALID, "voff/eoff must start at 0 and end at n/m");
    for (int b = 0; b < nb; ++b) {
      CM_REQUIRE(voff[b] <= voff[b + 1] && eoff[b] <= eoff[b + 1], CM_ERR_INVALID, "voff/eoff must ascend");
      CM_REQUIRE(voff[b] < voff[b + 1] || eoff[b] == eoff[b + 1], CM_ERR_INVALID,
                 "a graph without vertices must have no entries");
      CM_REQUIRE(params[b].nparts >= 1, CM_ERR_INVALID, "nparts must be >= 1");
      // the union's kernels run once per level for all graphs: one row
      // order, one round cap and one level cap
      CM_REQUIRE(params[b].row_order == params[0].row_order, CM_ERR_INVALID,
                 "all graphs of a batch use one row order");
      CM_REQUIRE(params[b].max_rounds == params[0].max_rounds && params[b].max_levels == params[0].max_levels,
                 CM_ERR_INVALID, "all graphs of a batch use one max_rounds and one max_levels");
    }
    CM_REQUIRE(concat->n < (int64_t)INT32_MAX, CM_ERR_UNSUPPORTED, "batch vertex count must fit int32");
    if (concat->offset_type == CM_OFF64)
      cm::batch_impl<int64_t>(ctx, nb, concat, voff, eoff, params, out);
    else
      cm::batch_impl<int32_t>(ctx, nb, concat, voff, eoff, params, out);
  } catch (const cm::Error &e) {
    ctx->err = e.what();
    for (int b = 0; b < nb; ++b) {
      delete out[b];
      out[b] = nullptr;
    }
    return e.code;
  }
  return CM_OK;
}
