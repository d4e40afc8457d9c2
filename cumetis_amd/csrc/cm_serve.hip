// cm_serve.hip -- continuous-batching partition server.
//
// Requests (host graphs) are admitted as slots free up and the HBM footprint
// budget allows.  One host thread steps coarsening: every step is one level
// for ALL admitted jobs at once, as one disjoint union (the batched kernels
// of cm_batch.hip), so a job submitted while others are mid-hierarchy joins
// the very next step at its level 0 and its work overlaps theirs on the
// device.  Jobs whose own stop rule ends them move to a second thread (own
// context and stream) that runs the initial partition and the uncoarsening,
// again overlapping the coarsening steps.  Every job's hierarchy stays
// device-resident until its result is produced.  Because the union kernels
// give each graph exactly its stand-alone result, a job's partition is the
// same alone and in any mix.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "cm_common.hpp"
#include "cm_internal.hpp"

namespace cm {

int ref_threshold(int64_t n, int nparts);

namespace {

double now_ms()
{
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

// one slot of a union step: copy a job's current level into the union
struct SlotSrc {
  const int64_t *xadj;
  const int32_t *adj, *wgt, *vw;
  int64_t sv, se;  // bases of the source values
  int64_t dv, de;  // first vertex / entry in the union
};

__device__ __forceinline__ int slot_of(const int64_t *base, int ns, int64_t x)
{
  int lo = 0, hi = ns;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (base[mid] <= x) lo = mid;
    else hi = mid;
  }
  return lo;
}

__global__ __launch_bounds__(256) void k_union_vertices(int64_t n, int ns, const SlotSrc *src,
                                                        const int64_t *dv, int64_t m_total, int64_t *xadj,
                                                        int32_t *vw, int32_t *gid)
{
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v > n) return;
  if (v == n) {
    xadj[n] = m_total;
    return;
  }
  const int s = slot_of(dv, ns, v);
  const SlotSrc q = src[s];
  const int64_t i = v - q.dv;
  xadj[v] = q.xadj[i] - q.se + q.de;
  vw[v] = q.vw[i];
  gid[v] = s;
}

__global__ __launch_bounds__(256) void k_union_edges(int64_t m, int ns, const SlotSrc *src, const int64_t *de,
                                                     int32_t *adj, int32_t *wgt)
{
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= m) return;
  const int s = slot_of(de, ns, e);
  const SlotSrc q = src[s];
  const int64_t j = e - q.de;
  adj[e] = (int32_t)(q.adj[j] - q.sv + q.dv);
  wgt[e] = q.wgt[j];
}

template <typename OffT>
__global__ void k_union_bases(int ns, const int64_t *vbase, const int32_t *cmap, const OffT *cxadj, int64_t nc,
                              int64_t *cvb, int64_t *ceb)
{
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s > ns) return;
  const int64_t c = s == ns ? nc : (int64_t)cmap[vbase[s]];
  cvb[s] = c;
  ceb[s] = (int64_t)cxadj[c];
}

}  // namespace

// A job's host graph, shared by its trials, with its largest row and vertex
// weight (taken while the submit checks it).
struct JobGraph {
  std::vector<int64_t> xadj;
  std::vector<int32_t> adjncy, adjwgt, vwgt;
  int64_t maxdeg = 0;
  int64_t maxvw = 0;
};

struct Request;

// One trial of a request: one hierarchy in the union steps, coarsened with
// the request's seed + trial.
struct Job {
  Request *req = nullptr;
  int trial = 0;
  std::shared_ptr<const JobGraph> gr;
  int64_t n = 0, m = 0;
  cm_coarsen_params p{};
  int T = 0;
  int32_t maxv = 0;
  int64_t tvwgt = 0;
  size_t footprint = 0;
  cm_hier *h = nullptr;
  int64_t cur_n = 0;  // vertices at the current level
  int status = CM_OK;
};

// A submitted graph: its trials run as separate jobs; the request is done
// when all have finished, with the lowest cut kept (the earlier trial on
// ties), as cm_partition does.
struct Request {
  uint64_t id = 0;
  int trials = 1;
  int finished = 0;
  double t_submit = 0, t_done = 0;
  std::vector<int32_t> part;
  cm_part_report rep{};
  int64_t best_cut = -1;
  int status = CM_OK;
  bool done = false;
  std::vector<std::unique_ptr<Job>> jobs;
  // level 0 on the device, uploaded when the first trial is admitted and
  // released (to the coarsening context's pool) when the last one finishes
  void *d_xadj = nullptr;
  int32_t *d_adjncy = nullptr, *d_adjwgt = nullptr, *d_vwgt = nullptr;
  int dev_left = 0;  // trials that have not finished
};

}  // namespace cm

struct cm_server {
  cm_ctx *cctx = nullptr;
  std::vector<cm_ctx *> fctx;  // one context (stream) per finisher thread
  std::thread loop;
  std::vector<std::thread> finishers;
  std::mutex mu;
  std::condition_variable cv_loop, cv_fin, cv_done;
  std::deque<cm::Job *> pending, to_finish;
  std::vector<cm::Job *> active;
  std::map<uint64_t, std::unique_ptr<cm::Request>> jobs;
  uint64_t next_id = 1;
  size_t budget = 0, used = 0;
  int max_slots = 0;
  int max_active = 0;
  bool stop = false;
  std::vector<double> latencies;
  double t_first = 0, t_last = 0;
  int64_t completed = 0;
};

namespace cm {

static size_t job_footprint(int64_t n, int64_t m)
{
  // level 0 + coarser levels (about as much again) + one union copy + the
  // contraction scratch of its share of a step
  return (size_t)(64 * m + 64 * n + 4096);
}

// One coarsening step over the union of the active jobs' current levels.
static void union_step(cm_server *S, std::vector<Job *> &act)
{
  cm_ctx *ctx = S->cctx;
  const int ns = (int)act.size();
  auto shared = std::make_shared<SharedBlocks>();
  shared->ctx = ctx;
  auto own = [&](size_t bytes) {
    void *p = ctx->pool.alloc(bytes ? bytes : 1);
    shared->blocks.push_back(p);
    return p;
  };
  std::vector<void *> tmp_blocks;
  auto tmp = [&](size_t bytes) {
    void *p = ctx->pool.alloc(bytes ? bytes : 1);
    tmp_blocks.push_back(p);
    return p;
  };
  std::vector<SlotSrc> src(ns);
  std::vector<int64_t> dv(ns + 1), de(ns + 1);
  std::vector<int32_t> hmaxv(ns);
  std::vector<uint64_t> hseed(ns);
  int64_t n = 0, m = 0;
  for (int s = 0; s < ns; ++s) {
    const Level &L = act[s]->h->levels.back();
    src[s] = SlotSrc{(const int64_t *)L.xadj, L.adjncy, L.adjwgt, L.vwgt, L.vbase, L.ebase, n, m};
    dv[s] = n;
    de[s] = m;
    n += L.n;
    m += L.m;
    hmaxv[s] = act[s]->maxv;
    hseed[s] = act[s]->p.seed;
  }
  dv[ns] = n;
  de[ns] = m;
  SlotSrc *d_src = (SlotSrc *)tmp(sizeof(SlotSrc) * ns);
  int64_t *d_dv = (int64_t *)tmp(8 * (ns + 1)), *d_de = (int64_t *)tmp(8 * (ns + 1));
  int32_t *d_maxv = (int32_t *)tmp(4 * ns);
  uint64_t *d_seed = (uint64_t *)tmp(8 * ns);
  CM_HIP(hipMemcpyAsync(d_src, src.data(), sizeof(SlotSrc) * ns, hipMemcpyHostToDevice, ctx->stream));
  CM_HIP(hipMemcpyAsync(d_dv, dv.data(), 8 * (ns + 1), hipMemcpyHostToDevice, ctx->stream));
  CM_HIP(hipMemcpyAsync(d_de, de.data(), 8 * (ns + 1), hipMemcpyHostToDevice, ctx->stream));
  CM_HIP(hipMemcpyAsync(d_maxv, hmaxv.data(), 4 * ns, hipMemcpyHostToDevice, ctx->stream));
  CM_HIP(hipMemcpyAsync(d_seed, hseed.data(), 8 * ns, hipMemcpyHostToDevice, ctx->stream));
  int64_t *xadj = (int64_t *)own(8 * (n + 1));
  int32_t *adj = (int32_t *)own(4 * (m > 0 ? m : 1)), *wgt = (int32_t *)own(4 * (m > 0 ? m : 1));
  int32_t *vw = (int32_t *)own(4 * (n > 0 ? n : 1)), *gid = (int32_t *)tmp(4 * (n > 0 ? n : 1));
  launch(ctx, "serve_union_v", 0.0, k_union_vertices, dim3(nblocks(n + 1, 256)), dim3(256), 0, n, ns,
         (const SlotSrc *)d_src, (const int64_t *)d_dv, m, xadj, vw, gid);
  launch(ctx, "serve_union_e", 0.0, k_union_edges, dim3(nblocks(m, 256)), dim3(256), 0, m, ns,
         (const SlotSrc *)d_src, (const int64_t *)d_de, adj, wgt);
  Level U;
  U.n = n;
  U.m = m;
  U.off64 = true;
  U.xadj = xadj;
  U.adjncy = adj;
  U.adjwgt = wgt;
  U.vwgt = vw;
  U.maxdeg = 0;
  U.maxvw = 0;
  for (int s = 0; s < ns; ++s) {
    const Level &S = act[s]->h->levels.back();
    U.maxdeg = (U.maxdeg < 0 || S.maxdeg < 0) ? -1 : std::max(U.maxdeg, S.maxdeg);
    U.maxvw = (U.maxvw < 0 || S.maxvw < 0) ? -1 : std::max(U.maxvw, S.maxvw);
  }
  int32_t *match = (int32_t *)own(4 * (n > 0 ? n : 1)), *cmap = (int32_t *)own(4 * (n > 0 ? n : 1));
  MatchOwners mo;
  mo.gid = gid;
  mo.maxv = d_maxv;
  mo.seed = d_seed;
  mo.vbase = d_dv;
  const int rounds = match_handshake(ctx, U, mo, 0, match, false);  // number() completes it
  int32_t *rep = (int32_t *)tmp(4 * (n > 0 ? n : 1));
  int64_t *segoff = (int64_t *)tmp(8 * (n + 1));
  int32_t *cvw = (int32_t *)own(4 * (n > 0 ? n : 1));
  const int64_t nc = number(ctx, U, match, cmap, rep, segoff, cvw);
  int64_t *cxadj = (int64_t *)own(8 * (nc + 1));
  int32_t *cadj = (int32_t *)own(4 * (m > 0 ? m : 1)), *cwgt = (int32_t *)own(4 * (m > 0 ? m : 1));
  int64_t cmaxdeg = -1;
  contract(ctx, U, match, cmap, nc, rep, segoff, act[0]->p.row_order, cxadj, cadj, cwgt, &cmaxdeg);
  const int64_t cmaxvw = ctx->h_small[3];
  int64_t *d_cvb = (int64_t *)tmp(8 * (ns + 1)), *d_ceb = (int64_t *)tmp(8 * (ns + 1));
  launch(ctx, "serve_bases", 0.0, k_union_bases<int64_t>, dim3(nblocks(ns + 1, 256)), dim3(256), 0, ns,
         (const int64_t *)d_dv, (const int32_t *)cmap, (const int64_t *)cxadj, nc, d_cvb, d_ceb);
  std::vector<int64_t> cvb(ns + 1), ceb(ns + 1);
  CM_HIP(hipMemcpyAsync(cvb.data(), d_cvb, 8 * (ns + 1), hipMemcpyDeviceToHost, ctx->stream));
  CM_HIP(hipMemcpyAsync(ceb.data(), d_ceb, 8 * (ns + 1), hipMemcpyDeviceToHost, ctx->stream));
  sync(ctx);
  for (void *q : tmp_blocks) ctx->pool.release(q);
  for (int s = 0; s < ns; ++s) {
    cm_hier *h = act[s]->h;
    Level &F = h->levels.back();
    // the level now lives in the union copy, next to its match and cmap
    F.xadj = xadj + dv[s];
    F.adjncy = adj + de[s];
    F.adjwgt = wgt + de[s];
    F.vwgt = vw + dv[s];
    F.vbase = dv[s];
    F.ebase = de[s];
    F.match = match + dv[s];
    F.cmap = cmap + dv[s];
    F.cbase = cvb[s];
    F.rounds = rounds;
    Level C;
    C.n = cvb[s + 1] - cvb[s];
    C.m = ceb[s + 1] - ceb[s];
    C.off64 = true;
    C.xadj = cxadj + cvb[s];
    C.adjncy = cadj + ceb[s];
    C.adjwgt = cwgt + ceb[s];
    C.vwgt = cvw + cvb[s];
    C.tvwgt = F.tvwgt;
    C.vbase = cvb[s];
    C.ebase = ceb[s];
    C.maxdeg = cmaxdeg;  // of the union: a bound for each job
    C.maxvw = cmaxvw;
    h->levels.push_back(C);
    h->shared_more.push_back(shared);
  }
}

static void loop_main(cm_server *S)
{
  hipSetDevice(S->cctx->device);
  for (;;) {
    std::vector<Job *> act, fresh;
    {
      std::unique_lock<std::mutex> lk(S->mu);
      S->cv_loop.wait(lk, [&] { return S->stop || !S->pending.empty() || !S->active.empty(); });
      if (S->stop) return;
      // admission: free slots and HBM budget (one job always fits when idle)
      while (!S->pending.empty() && (int)S->active.size() < S->max_slots) {
        Job *j = S->pending.front();
        if (!S->active.empty() && S->used + j->footprint > S->budget) break;
        S->pending.pop_front();
        S->used += j->footprint;
        S->active.push_back(j);
        fresh.push_back(j);
      }
      S->max_active = std::max(S->max_active, (int)S->active.size());
      act = S->active;
    }
    std::vector<Job *> retire;
    try {
      for (Job *j : fresh) {  // level 0 on the device, once per request
        cm_ctx *ctx = S->cctx;
        Request *r = j->req;
        const JobGraph &G = *j->gr;
        if (!r->d_xadj) {
          void *b[4] = {};
          try {
            b[0] = ctx->pool.alloc(8 * (j->n + 1));
            b[1] = ctx->pool.alloc(4 * (j->m > 0 ? j->m : 1));
            b[2] = ctx->pool.alloc(4 * (j->m > 0 ? j->m : 1));
            b[3] = ctx->pool.alloc(4 * (j->n > 0 ? j->n : 1));
          } catch (...) {  // all four or none: later trials test d_xadj alone
            for (void *q : b)
              if (q) ctx->pool.release(q);
            throw;
          }
          r->d_xadj = b[0];
          r->d_adjncy = (int32_t *)b[1];
          r->d_adjwgt = (int32_t *)b[2];
          r->d_vwgt = (int32_t *)b[3];
          CM_HIP(hipMemcpyAsync(r->d_xadj, G.xadj.data(), 8 * (j->n + 1), hipMemcpyHostToDevice, ctx->stream));
          if (j->m) {
            CM_HIP(hipMemcpyAsync(r->d_adjncy, G.adjncy.data(), 4 * j->m, hipMemcpyHostToDevice, ctx->stream));
            CM_HIP(hipMemcpyAsync(r->d_adjwgt, G.adjwgt.data(), 4 * j->m, hipMemcpyHostToDevice, ctx->stream));
          }
          if (j->n) CM_HIP(hipMemcpyAsync(r->d_vwgt, G.vwgt.data(), 4 * j->n, hipMemcpyHostToDevice, ctx->stream));
        }
        j->h = new cm_hier();
        Level L0;
        L0.n = j->n;
        L0.m = j->m;
        L0.off64 = true;
        L0.xadj = r->d_xadj;
        L0.adjncy = r->d_adjncy;
        L0.adjwgt = r->d_adjwgt;
        L0.vwgt = r->d_vwgt;
        L0.tvwgt = j->tvwgt;
        L0.maxdeg = G.maxdeg;
        L0.maxvw = G.maxvw;
        j->h->levels.push_back(L0);
        j->cur_n = j->n;
      }
      std::vector<Job *> work;
      for (Job *j : act)
        if (j->n > 0) work.push_back(j);
        else retire.push_back(j);
      if (!work.empty()) union_step(S, work);
      for (Job *j : work) {
        const Level &C = j->h->levels.back();
        const int64_t nf = j->cur_n;
        j->cur_n = C.n;
        const bool more = C.n > j->T && C.n < 0.75 * nf && C.m > C.n / 2;  // cu:1843-1845
        const bool cap = j->p.max_levels > 0 && (int)j->h->levels.size() > j->p.max_levels;
        if (!more || cap) retire.push_back(j);
      }
    } catch (const Error &e) {
      for (Job *j : act) j->status = e.code;
      retire = act;
    } catch (const std::exception &) {  // host allocation: the thread must survive
      for (Job *j : act) j->status = CM_ERR_NOMEM;
      retire = act;
    }
    std::lock_guard<std::mutex> lk(S->mu);
    for (Job *j : retire) {
      S->active.erase(std::find(S->active.begin(), S->active.end(), j));
      S->to_finish.push_back(j);
    }
    if (!retire.empty()) S->cv_fin.notify_all();
  }
}

static void release_level0(cm_server *S, Request *r)
{
  for (void *q : {r->d_xadj, (void *)r->d_adjncy, (void *)r->d_adjwgt, (void *)r->d_vwgt})
    if (q) S->cctx->pool.release(q);
  r->d_xadj = nullptr;
  r->d_adjncy = r->d_adjwgt = r->d_vwgt = nullptr;
}

static void finish_main(cm_server *S, cm_ctx *fctx)
{
  hipSetDevice(fctx->device);
  for (;;) {
    Job *j = nullptr;
    {
      std::unique_lock<std::mutex> lk(S->mu);
      S->cv_fin.wait(lk, [&] { return S->stop || !S->to_finish.empty(); });
      if (S->stop && S->to_finish.empty()) return;
      j = S->to_finish.front();
      S->to_finish.pop_front();
    }
    std::vector<int32_t> part;
    cm_part_report rep{};
    int64_t cut = 0;
    if (j->status == CM_OK) {
      try {
        part.assign(j->n, 0);  // one part: a cut of 0
        if (j->p.nparts > 1 && j->n > 0)
          partition_hierarchy(fctx, j->h, j->p.nparts, j->p.flags, part.data(), &rep, &cut);
        rep.levels = (int32_t)j->h->levels.size();
        rep.coarsest_n = (int32_t)j->h->levels.back().n;
      } catch (const Error &e) {
        j->status = e.code;
      } catch (const std::exception &) {  // host allocation in the initial partition
        j->status = CM_ERR_NOMEM;
      }
    }
    if (j->h) {
      for (void *q : j->h->owned) S->cctx->pool.release(q);
      delete j->h;  // drops its references to the union steps
      j->h = nullptr;
    }
    std::lock_guard<std::mutex> lk(S->mu);
    Request *r = j->req;
    S->used -= j->footprint;
    if (--r->dev_left == 0) release_level0(S, r);
    if (j->status != CM_OK && r->status == CM_OK) r->status = j->status;
    if (j->status == CM_OK) {
      r->rep.init_ms += rep.init_ms;
      r->rep.uncoarsen_ms += rep.uncoarsen_ms;
      // lowest cut, then the earlier trial (trials finish in any order)
      if (r->best_cut < 0 || cut < r->best_cut || (cut == r->best_cut && j->trial < r->rep.best_trial)) {
        r->best_cut = cut;
        r->part.swap(part);
        r->rep.best_trial = j->trial;
        r->rep.levels = rep.levels;
        r->rep.coarsest_n = rep.coarsest_n;
        r->rep.edgecut = cut;
      }
    }
    if (++r->finished == r->trials) {
      r->t_done = now_ms();
      r->rep.total_ms = r->t_done - r->t_submit;
      r->rep.trials = r->trials;
      r->done = true;
      S->latencies.push_back(r->rep.total_ms);
      S->t_last = r->t_done;
      ++S->completed;
      S->cv_done.notify_all();
    }
    S->cv_loop.notify_one();  // budget freed
  }
}

}  // namespace cm

extern "C" {

int cm_server_create(int device, int64_t hbm_budget, int32_t max_slots, cm_server **out)
{
  if (!out || max_slots < 1) return CM_ERR_INVALID;
  *out = nullptr;
  cm_server *S = new cm_server();
  int rc = cm_ctx_create(device, nullptr, &S->cctx);
  // host initial partitions run in parallel: each finisher has its own stream
  const unsigned hw = std::thread::hardware_concurrency();
  const int nfin = (int)std::max(1u, std::min(hw ? hw / 2 : 1u, 4u));
  for (int i = 0; i < nfin && rc == CM_OK; ++i) {
    cm_ctx *f = nullptr;
    rc = cm_ctx_create(device, nullptr, &f);
    if (rc == CM_OK) S->fctx.push_back(f);
  }
  if (rc != CM_OK) {
    for (cm_ctx *f : S->fctx) cm_ctx_destroy(f);
    if (S->cctx) cm_ctx_destroy(S->cctx);
    delete S;
    return rc;
  }
  S->budget = hbm_budget > 0 ? (size_t)hbm_budget : (size_t)64 << 30;
  S->max_slots = max_slots;
  S->loop = std::thread(cm::loop_main, S);
  for (cm_ctx *f : S->fctx) S->finishers.emplace_back(cm::finish_main, S, f);
  *out = S;
  return CM_OK;
}

int cm_server_submit(cm_server *S, const cm_graph_view *g, const cm_coarsen_params *p, uint64_t *ticket)
{
  if (!S || !g || !p || !ticket || g->on_device || p->nparts < 1 || g->n < 0 || p->trials < 0)
    return CM_ERR_INVALID;
  if (p->row_order != CM_ROW_FIRST_SEEN) return CM_ERR_UNSUPPORTED;  // one row order per union step
  if (g->offset_type != CM_OFF32 && g->offset_type != CM_OFF64) return CM_ERR_INVALID;
  if (g->n >= (int64_t)INT32_MAX || (g->m > 0 && !g->adjncy) || !g->xadj) return CM_ERR_INVALID;
  auto gr = std::make_shared<cm::JobGraph>();
  gr->xadj.resize(g->n + 1);
  for (int64_t i = 0; i <= g->n; ++i)
    gr->xadj[i] = g->offset_type == CM_OFF64 ? ((const int64_t *)g->xadj)[i] : ((const int32_t *)g->xadj)[i];
  if (gr->xadj[0] != 0 || gr->xadj[g->n] != g->m) return CM_ERR_INPUT;
  gr->adjncy.assign(g->adjncy, g->adjncy + g->m);
  gr->adjwgt = g->adjwgt ? std::vector<int32_t>(g->adjwgt, g->adjwgt + g->m) : std::vector<int32_t>(g->m, 1);
  gr->vwgt = g->vwgt ? std::vector<int32_t>(g->vwgt, g->vwgt + g->n) : std::vector<int32_t>(g->n, 1);
  for (int64_t i = 0; i < g->n; ++i) {
    if (gr->xadj[i] > gr->xadj[i + 1]) return CM_ERR_INPUT;
    gr->maxdeg = std::max(gr->maxdeg, gr->xadj[i + 1] - gr->xadj[i]);
    gr->maxvw = std::max<int64_t>(gr->maxvw, gr->vwgt[i]);
    for (int64_t e = gr->xadj[i]; e < gr->xadj[i + 1]; ++e)
      if (gr->adjncy[e] < 0 || gr->adjncy[e] >= g->n || gr->adjncy[e] == i || gr->adjwgt[e] <= 0)
        return CM_ERR_INPUT;
  }
  int64_t tvwgt = 0;
  for (int32_t w : gr->vwgt) tvwgt += w;
  const int T = p->threshold > 0 ? p->threshold : cm::ref_threshold(g->n, p->nparts);
  const int32_t maxv = p->maxvwgt > 0 ? p->maxvwgt : (int32_t)(1.5 * (double)tvwgt / T);
  auto r = std::make_unique<cm::Request>();
  r->trials = p->nparts > 1 && g->n > 0 ? (p->trials > 0 ? p->trials : CM_DEFAULT_TRIALS) : 1;
  r->dev_left = r->trials;
  r->t_submit = cm::now_ms();
  for (int t = 0; t < r->trials; ++t) {
    auto j = std::make_unique<cm::Job>();
    j->req = r.get();
    j->trial = t;
    j->gr = gr;
    j->n = g->n;
    j->m = g->m;
    j->p = *p;
    j->p.seed = p->seed + (uint64_t)t;
    j->tvwgt = tvwgt;
    j->T = T;
    j->maxv = maxv;
    j->footprint = cm::job_footprint(g->n, g->m);
    r->jobs.push_back(std::move(j));
  }
  std::lock_guard<std::mutex> lk(S->mu);
  r->id = S->next_id++;
  *ticket = r->id;
  if (S->t_first == 0) S->t_first = r->t_submit;
  for (auto &j : r->jobs) S->pending.push_back(j.get());
  S->jobs[r->id] = std::move(r);
  S->cv_loop.notify_one();
  return CM_OK;
}

int cm_server_wait(cm_server *S, uint64_t ticket, int32_t *part, cm_part_report *rep)
{
  if (!S) return CM_ERR_INVALID;
  std::unique_lock<std::mutex> lk(S->mu);
  auto it = S->jobs.find(ticket);
  if (it == S->jobs.end()) return CM_ERR_INVALID;
  cm::Request *r = it->second.get();
  S->cv_done.wait(lk, [&] { return r->done; });
  const int st = r->status;
  if (st == CM_OK) {
    if (part) std::copy(r->part.begin(), r->part.end(), part);
    if (rep) *rep = r->rep;
  }
  S->jobs.erase(it);
  return st;
}

int cm_server_stats(cm_server *S, int64_t *completed, double *p50_ms, double *p99_ms, double *graphs_per_s,
                    int32_t *max_active)
{
  if (!S) return CM_ERR_INVALID;
  std::lock_guard<std::mutex> lk(S->mu);
  std::vector<double> l = S->latencies;
  std::sort(l.begin(), l.end());
  auto pct = [&](double q) { return l.empty() ? 0.0 : l[std::min(l.size() - 1, (size_t)(q * (l.size() - 1) + 0.5))]; };
  if (completed) *completed = S->completed;
  if (p50_ms) *p50_ms = pct(0.50);
  if (p99_ms) *p99_ms = pct(0.99);
  if (graphs_per_s)
    *graphs_per_s = S->t_last > S->t_first ? S->completed / ((S->t_last - S->t_first) / 1e3) : 0.0;
  if (max_active) *max_active = S->max_active;
  return CM_OK;
}

int cm_server_destroy(cm_server *S)
{
  if (!S) return CM_ERR_INVALID;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->stop = true;
  }
  S->cv_loop.notify_all();
  S->cv_fin.notify_all();
  if (S->loop.joinable()) S->loop.join();
  for (auto &t : S->finishers)
    if (t.joinable()) t.join();
  for (auto &kv : S->jobs) {
    for (auto &j : kv.second->jobs)
      if (j->h) {
        for (void *q : j->h->owned) S->cctx->pool.release(q);
        delete j->h;
      }
    cm::release_level0(S, kv.second.get());
  }
  S->jobs.clear();
  for (cm_ctx *f : S->fctx) cm_ctx_destroy(f);
  cm_ctx_destroy(S->cctx);
  delete S;
  return CM_OK;
}

}  // extern "C"
