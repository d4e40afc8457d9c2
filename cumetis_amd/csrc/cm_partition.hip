// cm_partition.hip -- the multilevel k-way pipeline around the GPU coarsening
// (cuMetis_kway_partition, cu:3544-3616, and cuMetis_PartGraph,
// cu:3622-3670):
//   GPU coarsening -> coarsest level to the host -> recursive-bisection
//   initial partition (host) -> refinement of the coarsest level -> for each
//   finer level: GPU projection, then one refinement pass.
// Unless switched off, every level's refinement is preceded by a balancing
// pass (no counterpart in the reference) that empties parts above the
// refinement's upper bound int(w/k * 1.03): the refinement never moves a
// vertex out of such a part, so an overweight initial part stays overweight.
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <exception>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "cm_internal.hpp"
#include "cm_refine.hpp"
#include "host/cm_host.hpp"

namespace {

double now_ms()
{
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

cmh::HostGraph level_to_host(cm_ctx *ctx, const cm::Level &L)
{
  cmh::HostGraph g;
  g.n = (int)L.n;
  g.xadj.resize(L.n + 1);
  g.adjncy.resize(L.m);
  g.adjwgt.resize(L.m);
  g.vwgt.resize(L.n);
  if (L.off64) {
    cm::download(ctx, g.xadj.data(), L.xadj, 8 * (L.n + 1));
  } else {
    std::vector<int32_t> x(L.n + 1);
    cm::download(ctx, x.data(), L.xadj, 4 * (L.n + 1));
    for (int64_t i = 0; i <= L.n; ++i) g.xadj[i] = x[i];
  }
  cm::download(ctx, g.adjncy.data(), L.adjncy, 4 * L.m);
  cm::download(ctx, g.adjwgt.data(), L.adjwgt, 4 * L.m);
  cm::download(ctx, g.vwgt.data(), L.vwgt, 4 * L.n);
  if (L.ebase)
    for (auto &x : g.xadj) x -= L.ebase;
  if (L.vbase)
    for (auto &a : g.adjncy) a -= (int)L.vbase;
  return g;
}

// The parallel initial partition's large recursion steps coarsen here, each
// on a context lent from the owner's step contexts (at most STEP_CTX_MAX,
// created on first use): the step's host graph is uploaded, coarsened by
// cm_coarsen with the step's threshold, maxvwgt and seed, and its coarser
// levels and maps read back, both through the context's pinned chunks.  The
// contexts' cached device memory is freed when the initial partition ends.
class DeviceSteps final : public cmh::StepCoarsener {
 public:
  explicit DeviceSteps(cm_ctx *owner) : owner_(owner) {}
  ~DeviceSteps() override
  {
    if (!used_.load()) return;
    std::lock_guard<std::mutex> lk(owner_->steps.mu);
    for (cm_ctx *c : owner_->steps.idle) c->pool.trim();
  }
  int steps() const { return steps_.load(); }

  void coarsen(int n, const int64_t *xadj, const int *adjncy, const int *adjwgt, const int *vwgt, int threshold,
               int maxvwgt, uint64_t seed, std::vector<cmh::StepLevel> &levels) override
  {
    used_.store(true);
    Lease lease(owner_);
    cm_ctx *c = lease.c;
    // the step's graph on the device (staged), held until the hierarchy
    // built over it is released
    struct Inputs {
      cm_ctx *c;
      std::vector<void *> d;
      ~Inputs()
      {
        for (void *q : d) c->pool.release(q);
      }
    } in{c, {}};
    auto up = [&](const void *src, size_t bytes) -> void * {
      void *d = c->pool.alloc(bytes ? bytes : 1);
      in.d.push_back(d);
      cm::upload(c, d, src, bytes);
      return d;
    };
    cm_graph_view v{};
    v.n = n;
    v.m = xadj[n];
    v.offset_type = CM_OFF64;
    v.on_device = 1;
    v.xadj = up(xadj, 8 * ((size_t)n + 1));
    v.adjncy = (const int32_t *)up(adjncy, 4 * (size_t)v.m);
    v.adjwgt = (const int32_t *)up(adjwgt, 4 * (size_t)v.m);
    v.vwgt = (const int32_t *)up(vwgt, 4 * (size_t)n);
    cm_coarsen_params p{};
    p.nparts = 2;
    p.threshold = threshold;
    p.maxvwgt = maxvwgt;
    p.row_order = CM_ROW_FIRST_SEEN;
    p.seed = seed;
    p.flags = CM_FLAG_TRUSTED_INPUT;  // a subgraph the pipeline built from a validated graph
    cm_hier *h = nullptr;
    const int rc = cm_coarsen(c, &v, &p, &h);
    if (rc != CM_OK) throw cm::Error(rc, "initial-partition step coarsening: " + c->err);
    struct Release {
      cm_ctx *c;
      cm_hier *h;
      ~Release() { cm_hier_release(c, h); }
    } rel{c, h};
    const int nl = (int)h->levels.size();
    levels.clear();
    levels.resize(nl - 1);
    for (int l = 0; l + 1 < nl; ++l) {
      const cm::Level &F = h->levels[l];
      cmh::StepLevel &S = levels[l];
      S.cmap.resize(F.n);
      cm::download(c, S.cmap.data(), F.cmap, 4 * F.n);
      cmh::HostGraph g = level_to_host(c, h->levels[l + 1]);
      S.n = g.n;
      S.xadj = std::move(g.xadj);
      S.adjncy = std::move(g.adjncy);
      S.adjwgt = std::move(g.adjwgt);
      S.vwgt = std::move(g.vwgt);
    }
    steps_.fetch_add(1);
  }

 private:
  struct Lease {
    cm_ctx *owner, *c = nullptr;
    explicit Lease(cm_ctx *o) : owner(o)
    {
      auto &st = owner->steps;
      std::unique_lock<std::mutex> lk(st.mu);
      st.cv.wait(lk, [&] { return !st.idle.empty() || (int)st.all.size() < cm::STEP_CTX_MAX; });
      if (!st.idle.empty()) {
        c = st.idle.back();
        st.idle.pop_back();
      } else {
        const int rc = cm_ctx_create(owner->device, nullptr, &c);
        CM_REQUIRE(rc == CM_OK, rc, "context for an initial-partition step");
        st.all.push_back(c);
      }
      c->tune = owner->tune;
    }
    ~Lease()
    {
      {
        std::lock_guard<std::mutex> lk(owner->steps.mu);
        owner->steps.idle.push_back(c);
      }
      owner->steps.cv.notify_one();
    }
  };
  cm_ctx *owner_;
  std::atomic<int> steps_{0};
  std::atomic<bool> used_{false};
};

// A trial's initial-partition error, with host allocation failures as
// CM_ERR_NOMEM.
[[noreturn]] void rethrow_init(std::exception_ptr e)
{
  try {
    std::rethrow_exception(e);
  } catch (const std::bad_alloc &) {
    throw cm::Error(CM_ERR_NOMEM, "initial partition failed (host allocation)");
  }
}

}  // namespace

namespace cm {

cmh::HostGraph coarsest_to_host(cm_ctx *ctx, const cm_hier *h) { return level_to_host(ctx, h->levels.back()); }

void uncoarsen_hierarchy(cm_ctx *ctx, const cm_hier *h, int nparts, const cmh::HostGraph &hc,
                         const std::vector<int> &where, int32_t *part, bool balance, int64_t *cut)
{
  const int L = (int)h->levels.size();
  const Level &C = h->levels[L - 1];
  const int64_t tvwgt = h->levels[0].tvwgt;
  std::vector<int64_t> pw(nparts, 0);
  for (int i = 0; i < hc.n; ++i) pw[where[i]] += hc.vwgt[i];
  int32_t *w0 = (int32_t *)ctx->pool.alloc(4 * (C.n > 0 ? C.n : 1)), *w1 = nullptr;
  unsigned long long *d_cut = nullptr, h_cut = 0;
  try {
    if (C.n) CM_HIP(hipMemcpyAsync(w0, where.data(), 4 * C.n, hipMemcpyHostToDevice, ctx->stream));
    if (balance) balance_level(ctx, C, nparts, w0, pw, tvwgt);
    refine_level(ctx, C, nparts, w0, pw, tvwgt);
    for (int lev = L - 2; lev >= 0; --lev) {
      const Level &F = h->levels[lev];
      w1 = (int32_t *)ctx->pool.alloc(4 * (F.n > 0 ? F.n : 1));
      project(ctx, F.n, F.cmap, w0 - F.cbase, w1);  // stored cmap = local + cbase
      ctx->pool.release(w0);
      w0 = w1;
      w1 = nullptr;
      if (balance) balance_level(ctx, F, nparts, w0, pw, tvwgt);
      refine_level(ctx, F, nparts, w0, pw, tvwgt);
    }
    const int64_t n0 = h->levels[0].n;
    if (cut) {
      d_cut = (unsigned long long *)ctx->pool.alloc(sizeof(unsigned long long));
      level_cut(ctx, h->levels[0], w0, d_cut);
      CM_HIP(hipMemcpyAsync(&h_cut, d_cut, sizeof(h_cut), hipMemcpyDeviceToHost, ctx->stream));
    }
    if (n0) CM_HIP(hipMemcpyAsync(part, w0, 4 * n0, hipMemcpyDeviceToHost, ctx->stream));
    sync(ctx);
  } catch (...) {
    ctx->pool.release(w0);
    if (w1) ctx->pool.release(w1);
    if (d_cut) ctx->pool.release(d_cut);
    throw;
  }
  ctx->pool.release(w0);
  if (d_cut) ctx->pool.release(d_cut);
  if (cut) *cut = (int64_t)(h_cut / 2);
}

void partition_hierarchy(cm_ctx *ctx, const cm_hier *h, int nparts, int32_t flags, int32_t *part,
                         cm_part_report *rep, int64_t *cut)
{
  const double t1 = now_ms();
  cmh::HostGraph hc = coarsest_to_host(ctx, h);
  std::vector<int> where;
  {
    DeviceSteps dev(ctx);
    cmh::initial_partition(hc, nparts, where,
                           (flags & CM_FLAG_EXACT_INITPART) ? cmh::InitForm::EXACT : cmh::InitForm::AUTO,
                           (flags & CM_FLAG_HOST_INIT_COARSEN) ? nullptr : &dev);
    rep->init_device_steps = dev.steps();
  }
  const double t2 = now_ms();
  uncoarsen_hierarchy(ctx, h, nparts, hc, where, part, !(flags & CM_FLAG_NO_BALANCE), cut);
  rep->init_ms = t2 - t1;
  rep->uncoarsen_ms = now_ms() - t2;
}

}  // namespace cm

extern "C" {

// Trials: the graph goes to the device once; trial t coarsens it with seed
// + t and runs the same initial partition (glibc seed-1 stream) and
// uncoarsening; the partition with the lowest cut is kept, the earlier trial
// on ties.  Single multilevel runs spread widely in cut (one refinement pass
// per level, a coarsest graph that depends on the matching), so the best of
// a few seeds is the standard remedy (METIS's ncuts).
int cm_partition(cm_ctx *ctx, const cm_graph_view *g, const cm_coarsen_params *p, int32_t *part,
                 cm_part_report *rep)
{
  if (!ctx || !g || !p || !part) return CM_ERR_INVALID;
  std::vector<void *> dev;
  auto release_dev = [&]() {
    for (void *q : dev) ctx->pool.release(q);
    dev.clear();
  };
  try {
    CM_REQUIRE(!g->on_device, CM_ERR_INVALID, "cm_partition takes a host graph");
    CM_REQUIRE(p->nparts >= 1, CM_ERR_INVALID, "nparts must be >= 1");
    CM_REQUIRE(p->trials >= 0, CM_ERR_INVALID, "trials must be >= 0");
    const int trials = p->trials > 0 ? p->trials : CM_DEFAULT_TRIALS;
    const double t0 = now_ms();
    if (p->nparts == 1 || g->n == 0) {
      for (int64_t i = 0; i < g->n; ++i) part[i] = 0;
      if (rep) *rep = cm_part_report{0, 0, 0, 0, 1, (int32_t)g->n, 0, 0, 0};
      return CM_OK;
    }
    CM_HIP(hipSetDevice(ctx->device));
    cm::validate_view(g);
    // the graph on the device once, shared by the trials (weights stay
    // implicit when not given)
    const size_t ob = g->offset_type == CM_OFF64 ? 8 : 4;
    auto up = [&](const void *src, size_t bytes) -> void * {
      if (!src) return nullptr;
      void *d = ctx->pool.alloc(bytes ? bytes : 1);
      dev.push_back(d);
      cm::upload(ctx, d, src, bytes);
      return d;
    };
    cm_graph_view dv = *g;
    dv.on_device = 1;
    dv.xadj = up(g->xadj, ob * (g->n + 1));
    dv.adjncy = (const int32_t *)up(g->adjncy, 4 * g->m);
    dv.adjwgt = (const int32_t *)up(g->adjwgt, 4 * g->m);
    dv.vwgt = (const int32_t *)up(g->vwgt, 4 * g->n);
    // Trials overlap: the host initial partition of trial t runs in a thread
    // of its own while the GPU coarsens the later trials; the uncoarsenings
    // then run one after another on the stream.  Each trial's result is the
    // one it has alone (the initial partition draws from a private stream).
    // the trials' large initial-partition steps coarsen on the device
    // (declared before the trials, so that it ends after their threads)
    DeviceSteps steps(ctx);
    struct Trial {
      cm_hier *h = nullptr;
      cmh::HostGraph hc;
      std::vector<int> where;
      std::thread th;
      std::exception_ptr err;
    };
    struct Trials {
      cm_ctx *ctx;
      std::vector<Trial> t;
      ~Trials()
      {
        for (auto &x : t) {
          if (x.th.joinable()) x.th.join();
          if (x.h) cm_hier_release(ctx, x.h);
        }
      }
    } tr{ctx, std::vector<Trial>(trials)};
    int64_t best = -1;
    cm_part_report r{};
    r.trials = trials;
    const int nparts = p->nparts;
    const double p0 = now_ms();
    cmh::StepCoarsener *const sdev = (p->flags & CM_FLAG_HOST_INIT_COARSEN) ? nullptr : &steps;
    for (int t = 0; t < trials; ++t) {
      cm_coarsen_params q = *p;
      q.seed = p->seed + (uint64_t)t;
      if (t > 0) q.flags |= CM_FLAG_TRUSTED_INPUT;  // validated by the first trial
      const double c0 = now_ms();
      int rc = cm_coarsen(ctx, &dv, &q, &tr.t[t].h);
      if (rc != CM_OK) {
        release_dev();
        return rc;
      }
      r.coarsen_ms += now_ms() - c0;
      Trial &x = tr.t[t];
      x.hc = cm::coarsest_to_host(ctx, x.h);
      const bool exact = (p->flags & CM_FLAG_EXACT_INITPART) != 0;
      x.th = std::thread([&x, nparts, exact, sdev]() {
        try {
          cmh::initial_partition(x.hc, nparts, x.where, exact ? cmh::InitForm::EXACT : cmh::InitForm::AUTO, sdev);
        } catch (...) {
          x.err = std::current_exception();
        }
      });
    }
    for (auto &x : tr.t) x.th.join();
    const double p1 = now_ms();
    r.init_ms = p1 - p0 - r.coarsen_ms;  // host time the trials' initial partitions add
    // the uncoarsenings at once: trial t > 0 on helper context t - 1 (a
    // stream, pool and counters of its own, driven by a host thread of its
    // own), each ending with its level-0 edge cut counted on the device.  The
    // levels they read were complete before the initial partitions' read-backs.
    r.init_device_steps = steps.steps();
    for (auto &x : tr.t)
      if (x.err) rethrow_init(x.err);
    while ((int)ctx->helpers.size() < trials - 1) {
      cm_ctx *hc = nullptr;
      const int rc = cm_ctx_create(ctx->device, nullptr, &hc);
      CM_REQUIRE(rc == CM_OK, rc, "helper context for the trials");
      ctx->helpers.push_back(hc);
    }
    for (cm_ctx *hc : ctx->helpers) hc->tune = ctx->tune;
    std::vector<std::vector<int32_t>> parts(trials);
    std::vector<int64_t> cut(trials, -1);
    std::vector<int32_t> levels(trials), coarsest(trials);
    std::vector<std::exception_ptr> errs(trials);
    auto uncoarsen = [&](int t) {
      cm_ctx *c = t == 0 ? ctx : ctx->helpers[t - 1];
      try {
        if (t > 0) CM_HIP(hipSetDevice(c->device));
        Trial &x = tr.t[t];
        parts[t].resize(g->n);
        cm::uncoarsen_hierarchy(c, x.h, nparts, x.hc, x.where, parts[t].data(), !(p->flags & CM_FLAG_NO_BALANCE),
                                &cut[t]);
        levels[t] = (int32_t)x.h->levels.size();
        coarsest[t] = (int32_t)x.h->levels.back().n;
      } catch (...) {
        errs[t] = std::current_exception();
        (void)hipStreamSynchronize(c->stream);  // nothing of it still reads the graph
      }
    };
    {
      std::vector<std::thread> th;
      for (int t = 1; t < trials; ++t) th.emplace_back(uncoarsen, t);
      uncoarsen(0);
      for (auto &x : th) x.join();
    }
    for (auto &e : errs)
      if (e) std::rethrow_exception(e);
    for (auto &x : tr.t) {
      cm_hier_release(ctx, x.h);
      x.h = nullptr;
    }
    r.uncoarsen_ms = now_ms() - p1;
    for (int t = 0; t < trials; ++t)
      if (best < 0 || cut[t] < best) {  // the earlier trial on equal cuts
        best = cut[t];
        r.best_trial = t;
        r.levels = levels[t];
        r.coarsest_n = coarsest[t];
        r.edgecut = cut[t];
      }
    std::copy(parts[r.best_trial].begin(), parts[r.best_trial].end(), part);
    cm::sync(ctx);
    release_dev();
    r.total_ms = now_ms() - t0;
    if (rep) *rep = r;
  } catch (const cm::Error &e) {
    ctx->err = e.what();
    (void)hipStreamSynchronize(ctx->stream);
    release_dev();
    return e.code;
  } catch (const std::exception &e) {
    ctx->err = e.what();
    (void)hipStreamSynchronize(ctx->stream);
    release_dev();
    return CM_ERR_INVALID;
  }
  return CM_OK;
}

// cu:3622-3670.  One process-wide context on device 0, created by the first
// call.  Calls are serialised: the context (its stream, pool and read-back
// buffers) serves one host thread at a time, and the reference itself is not
// reentrant (its globals, cu:4-144).
int cuMetis_PartGraph(int *nvtxs, int *xadj, int *adjncy, int *vwgt, int *adjwgt, int *nparts,
                      float *tpwgts, float *ubvec, int *objval, int *part)
{
  (void)tpwgts;
  (void)ubvec;
  static std::mutex mu;
  static cm_ctx *ctx = nullptr;
  if (!nvtxs || !xadj || !nparts || !part) return CM_ERR_INVALID;
  std::lock_guard<std::mutex> lk(mu);
  if (!ctx) {
    int rc = cm_ctx_create(0, nullptr, &ctx);
    if (rc != CM_OK) return rc;
  }
  cm_graph_view v{*nvtxs, (int64_t)xadj[*nvtxs], CM_OFF32, 0, xadj, adjncy, adjwgt, vwgt};
  cm_coarsen_params p{};
  p.nparts = *nparts;
  p.row_order = CM_ROW_FIRST_SEEN;
  p.seed = 1;
  cm_part_report rep{};
  int rc = cm_partition(ctx, &v, &p, part, &rep);
  if (rc != CM_OK) return rc;
  if (objval) *objval = (int)rep.edgecut;
  return 1;
}

}  // extern "C"
