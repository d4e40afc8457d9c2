// cm_initpart.cpp -- the reference's host initial partition, restated in
// C++: recursive multilevel bisection with its own CPU coarsening (HEM +
// hash contraction, T = 200), BFS region growing, 2-way balancing and FM
// refinement (cuMetis.cu cu:1879-3062, cu:3091-3108).  The random choices
// draw glibc rand() in the reference's order (cu:287-300), so results follow
// the reference stream exactly.
//
// Coarsest graphs of PAR_MIN_N vertices or more (the GPU coarsening stalls
// early on graphs with many isolated vertices and stars, R-MAT: 3.7 M
// vertices at scale 22, where the serial recursion takes minutes; a 160^3
// grid in 128 parts stops at 27 K vertices, 77 ms serial) take a
// parallel form of the same algorithm: isolated vertices are set aside (they
// add nothing to the cut) and placed afterwards on the lightest parts, and
// every recursion step draws from its own stream, seeded from its place in
// the recursion, so the two halves of a split run on separate threads and
// the result does not depend on how many threads there are.
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <deque>
#include <exception>
#include <memory>
#include <queue>
#include <stdexcept>
#include <thread>
#include <vector>

#include "cm_host.hpp"

namespace cmh {
namespace {

constexpr int PAR_MIN_N = 1 << 14;

// Order-preserving parallel loops for the passes over every entry of a large
// graph (gains, projection, split, contraction): f(lo, hi) runs on contiguous
// ranges of [0, n), each writing only its own range's results, so the output
// is the serial loop's in both forms.  Ranges below LOOP_MIN_WORK entries of
// work run inline.
constexpr int64_t LOOP_MIN_WORK = 1 << 20;

int loop_threads()
{
  static const int t = (int)std::max(1u, std::min(std::thread::hardware_concurrency(), 16u));
  return t;
}

// Ranges: ceil(work / LOOP_MIN_WORK) of them, at most loop_threads(); b[c] is
// the first index of range c, balanced by the entries of xadj.
std::vector<int> split_ranges(int n, int64_t work, const int64_t *xadj)
{
  int nr = (int)std::min<int64_t>(loop_threads(), std::max<int64_t>(1, work / LOOP_MIN_WORK));
  if (n < 2 * nr) nr = 1;
  std::vector<int> b(nr + 1, n);
  b[0] = 0;
  for (int c = 1; c < nr; ++c)  // the first vertex whose entries start at or after c/nr of them
    b[c] = (int)(std::lower_bound(xadj, xadj + n, xadj[n] / nr * c) - xadj);
  return b;
}

// Threads whose exceptions reach the thread that started them: join()
// waits for all and rethrows the first; the destructor (unwinding from an
// exception of the starting thread) joins without rethrowing.
class Threads {
 public:
  template <typename F>
  void run(F f)
  {
    errs_.emplace_back();
    std::exception_ptr *e = &errs_.back();  // deque: stays valid as it grows
    th_.emplace_back([f, e]() mutable {
      try {
        f();
      } catch (...) {
        *e = std::current_exception();
      }
    });
  }
  void join()
  {
    for (auto &t : th_)
      if (t.joinable()) t.join();
    for (auto &e : errs_)
      if (e) std::rethrow_exception(e);
  }
  ~Threads()
  {
    for (auto &t : th_)
      if (t.joinable()) t.join();
  }

 private:
  std::vector<std::thread> th_;
  std::deque<std::exception_ptr> errs_;
};

// f(c, lo, hi) for every range c, range 0 on the calling thread
template <typename F>
void run_ranges(const std::vector<int> &b, F &&f)
{
  Threads th;
  for (int c = 1; c + 1 < (int)b.size(); ++c) th.run([&f, &b, c]() { f(c, b[c], b[c + 1]); });
  f(0, b[0], b[1]);
  th.join();
}

template <typename F>
void par_for(int n, int64_t work, const int64_t *xadj, F &&f)
{
  run_ranges(split_ranges(n, work, xadj), [&](int, int lo, int hi) { f(lo, hi); });
}

// cu:287-300 draws from glibc rand() with its default seed 1.  rand() is
// random() over a TYPE_3 (128-byte) state; random_r over a private state
// seeded with 1 yields the same sequence without sharing the process-wide
// state, so concurrent partitions stay deterministic.
struct RandStream {
  random_data rd;
  char state[128];
  explicit RandStream(unsigned seed = 1)
  {
    memset(&rd, 0, sizeof(rd));
    memset(state, 0, sizeof(state));
    initstate_r(seed, state, sizeof(state), &rd);
    srandom_r(seed, &rd);
  }
  int32_t next()
  {
    int32_t r;
    random_r(&rd, &r);
    return r;
  }
};
thread_local RandStream *g_stream = nullptr;

int rnd(int max) { return (int)((uint32_t)g_stream->next() % (uint32_t)max); }

// cu:639-671: identity (flag) then m 4-element block swaps (n >= 10) or n
// pair swaps (n < 10)
void rand_perm(int n, int *p, int m, bool init)
{
  if (init)
    for (int i = 0; i < n; ++i) p[i] = i;
  if (n <= 0) return;
  if (n < 10) {
    for (int i = 0; i < n; ++i) {
      const int v = rnd(n), u = rnd(n);
      std::swap(p[v], p[u]);
    }
  } else {
    for (int i = 0; i < m; ++i) {
      const int v = rnd(n - 3), u = rnd(n - 3);
      std::swap(p[v + 0], p[u + 2]);
      std::swap(p[v + 1], p[u + 3]);
      std::swap(p[v + 2], p[u + 0]);
      std::swap(p[v + 3], p[u + 1]);
    }
  }
}

// Max-priority queue with float keys and a position index (cu:1976-2242).
class MaxQueue {
 public:
  MaxQueue() = default;
  explicit MaxQueue(int cap) : heap_(cap > 0 ? cap : 1), pos_(cap > 0 ? cap : 1, -1) {}
  // room for vertex ids below cap; an empty queue stays empty
  void reserve(int cap)
  {
    if ((size_t)cap > heap_.size()) {
      heap_.resize(cap);
      pos_.resize(cap, -1);
    }
  }
  void clear()
  {
    for (long i = size_ - 1; i >= 0; --i) pos_[heap_[i].val] = -1;
    size_ = 0;
  }
  void push(int node, int key)
  {
    long i = size_++;
    const float k = (float)key;
    while (i > 0) {
      const long j = (i - 1) >> 1;
      if (!(k > heap_[j].key)) break;
      move(i, j);
      i = j;
    }
    put(i, k, node);
  }
  int pop()
  {
    if (size_ == 0) return -1;
    --size_;
    const int top = heap_[0].val;
    pos_[top] = -1;
    if (size_ > 0) sink(0, heap_[size_].key, heap_[size_].val);
    return top;
  }
  void erase(int node)
  {
    const long i = pos_[node];
    pos_[node] = -1;
    if (--size_ > 0 && heap_[size_].val != node) {
      const int last = heap_[size_].val;
      const float nk = heap_[size_].key, ok = heap_[i].key;
      if (nk > ok) swim(i, nk, last);
      else sink(i, nk, last);
    }
  }
  void update(int node, int key)
  {
    const long i = pos_[node];
    const float nk = (float)key, ok = heap_[i].key;
    if (nk > ok) swim(i, nk, node);
    else sink(i, nk, node);
  }

 private:
  struct Item {
    float key;
    int val;
  };
  std::vector<Item> heap_;
  std::vector<long> pos_;
  long size_ = 0;
  void move(long to, long from)
  {
    heap_[to] = heap_[from];
    pos_[heap_[to].val] = to;
  }
  void put(long i, float k, int node)
  {
    heap_[i].key = k;
    heap_[i].val = node;
    pos_[node] = i;
  }
  void swim(long i, float k, int node)
  {
    while (i > 0) {
      const long j = (i - 1) >> 1;
      if (!(k > heap_[j].key)) break;
      move(i, j);
      i = j;
    }
    put(i, k, node);
  }
  void sink(long i, float k, int node)
  {
    long j;
    while ((j = 2 * i + 1) < size_) {
      if (heap_[j].key > k) {
        if (j + 1 < size_ && heap_[j + 1].key > heap_[j].key) ++j;
      } else if (j + 1 < size_ && heap_[j + 1].key > k) {
        ++j;
      } else {
        break;
      }
      move(i, j);
      i = j;
    }
    put(i, k, node);
  }
};

// Per-thread FM scratch kept between calls.  The reference creates two
// queues with an n-entry position index per FM call (cu:1976-1990, 2393-
// 2394), an O(n) cost that dominated the initial partition of large graphs;
// here the queues and the `moved` marks live on, and every FM call leaves
// them as it found them (queues empty, marks -1).
struct FmWork {
  MaxQueue q[2];
  std::vector<int> moved, swaps, perm;  // swaps and perm are written before read
  void reserve(int n)
  {
    q[0].reserve(n);
    q[1].reserve(n);
    if ((size_t)n > moved.size()) moved.resize(n, -1);
    if ((size_t)n + 1 > swaps.size()) {
      swaps.resize(n + 1);
      perm.resize(n + 1);
    }
  }
};
thread_local FmWork tl_fm;

// Graph of one recursion step plus its 2-way state.
struct Part2 {
  int n = 0;
  int64_t m = 0;
  hvec<int64_t> xadj;  // graph arrays: sized, then written whole
  hvec<int> adjncy, adjwgt, vwgt, cmap;
  std::vector<int> label;
  std::vector<int> where, bndptr, bndlist, id, ed;
  int pwgts[2] = {0, 0};
  int tvwgt = 0, mincut = 0, nbnd = 0;
  std::unique_ptr<Part2> coarser;
  Part2 *finer = nullptr;

  void set_tvwgt()
  {
    int s = 0;
    for (int i = 0; i < n; ++i) s += vwgt[i];
    tvwgt = s;
  }
  void alloc2way()
  {
    where.assign(n, 0);
    bndptr.assign(n, -1);
    bndlist.assign(n, 0);
    id.assign(n, 0);
    ed.assign(n, 0);
  }
  void bnd_insert(int i)
  {
    bndlist[nbnd] = i;
    bndptr[i] = nbnd++;
  }
  void bnd_delete(int i)
  {
    bndlist[bndptr[i]] = bndlist[--nbnd];
    bndptr[bndlist[nbnd]] = bndptr[i];
    bndptr[i] = -1;
  }
};

struct Settings {
  float ub = 1.03f;      // cu:390
  int niter = 10;        // cu:377
  int threshold = 200;   // cu:378
  float balance[2] = {0, 0};
  unsigned try_seed = 0;  // parallel form: the region-growing tries' streams (0: the step's stream)
  StepCoarsener *dev = nullptr;  // parallel form: the device coarsener of the large steps
};

// cu:830-903 (HEM with degree-bucketed random visit order) + cu:737-825
int coarsen_once(Part2 &g, int maxvwgt, Part2 &c)
{
  const int n = g.n;
  std::vector<int> match(n, -1), perm(n), tperm(n), d(n);
  g.cmap.assign(n, 0);
  rand_perm(n, tperm.data(), n / 8, true);
  const int aved = (int)(0.7 * (double)(int)(g.xadj[n] / n));
  for (int i = 0; i < n; ++i) d[i] = (int)std::min<int64_t>(g.xadj[i + 1] - g.xadj[i], aved);
  {
    std::vector<int> cnt(aved + 2, 0);
    for (int i = 0; i < n; ++i) cnt[d[i]]++;
    int run = 0;
    for (int k = 0; k <= aved + 1; ++k) {
      const int t = cnt[k];
      cnt[k] = run;
      run += t;
    }
    for (int k = 0; k < n; ++k) perm[cnt[d[tperm[k]]]++] = tperm[k];
  }
  for (int k = 0; k < n; ++k) {
    const int i = perm[k];
    if (match[i] != -1) continue;
    int best = i, bw = -1;
    for (int64_t j = g.xadj[i]; j < g.xadj[i + 1]; ++j) {
      const int u = g.adjncy[j];
      if (match[u] == -1 && bw < g.adjwgt[j] && g.vwgt[i] + g.vwgt[u] <= maxvwgt) {
        best = u;
        bw = g.adjwgt[j];
      }
      if (best == i && 3 * g.vwgt[i] < maxvwgt) best = -1;
    }
    if (best != -1) {
      match[i] = best;
      match[best] = i;
    }
  }
  int nc = 0;
  for (int i = 0; i < n; ++i) {
    if (match[i] == -1) {
      match[i] = i;
      g.cmap[i] = nc++;
    } else if (i <= match[i]) {
      g.cmap[i] = g.cmap[match[i]] = nc++;
    }
  }
  // first-seen hash contraction
  c.n = nc;
  c.xadj.assign(nc + 1, 0);
  c.adjncy.resize(std::max<int64_t>(g.m, 1));
  c.adjwgt.resize(std::max<int64_t>(g.m, 1));
  c.vwgt.assign(nc, 0);
  std::vector<int64_t> slot(nc, -1);
  int64_t out = 0;
  int cv = 0;
  for (int v = 0; v < n; ++v) {
    const int u = match[v];
    if (u < v) continue;
    const int64_t base = out;
    auto absorb = [&](int x) {
      for (int64_t j = g.xadj[x]; j < g.xadj[x + 1]; ++j) {
        const int k = g.cmap[g.adjncy[j]];
        if (slot[k] == -1) {
          slot[k] = out;
          c.adjncy[out] = k;
          c.adjwgt[out++] = g.adjwgt[j];
        } else {
          c.adjwgt[slot[k]] += g.adjwgt[j];
        }
      }
    };
    c.vwgt[cv] = g.vwgt[v];
    absorb(v);
    if (u != v) {
      c.vwgt[cv] += g.vwgt[u];
      absorb(u);
      const int64_t s = slot[cv];
      if (s != -1) {  // self entry takes the last entry's place (cu:804-808)
        --out;
        c.adjncy[s] = c.adjncy[out];
        c.adjwgt[s] = c.adjwgt[out];
        slot[cv] = -1;
      }
    }
    for (int64_t j = base; j < out; ++j) slot[c.adjncy[j]] = -1;
    c.xadj[++cv] = out;
  }
  c.m = out;
  c.adjncy.resize(out);  // (sized for g.m entries above)
  c.adjwgt.resize(out);
  c.set_tvwgt();
  return nc;
}

// cu:1854-1875 with T = 200
Part2 *cpu_coarsen(Part2 &g, Settings &s)
{
  const int maxvwgt = (int)(1.5 * g.tvwgt / s.threshold);
  Part2 *cur = &g;
  do {
    cur->coarser.reset(new Part2());
    Part2 *c = cur->coarser.get();
    coarsen_once(*cur, maxvwgt, *c);
    c->finer = cur;
    cur = c;
  } while (cur->n > s.threshold && cur->n < 0.75 * cur->finer->n && cur->m > cur->n / 2);
  return cur;
}

// The same chain built by the device coarsener (the parallel form's large
// steps; seeded by the step's try stream).
Part2 *dev_coarsen(Part2 &g, Settings &s)
{
  const int maxvwgt = (int)(1.5 * g.tvwgt / s.threshold);
  std::vector<StepLevel> lv;
  s.dev->coarsen(g.n, g.xadj.data(), g.adjncy.data(), g.adjwgt.data(), g.vwgt.data(), s.threshold, maxvwgt,
                 s.try_seed, lv);
  if (lv.empty()) throw std::runtime_error("device coarsening returned no level");
  Part2 *cur = &g;
  for (StepLevel &L : lv) {
    cur->cmap = std::move(L.cmap);
    cur->coarser.reset(new Part2());
    Part2 *c = cur->coarser.get();
    c->n = L.n;
    c->xadj = std::move(L.xadj);
    c->adjncy = std::move(L.adjncy);
    c->adjwgt = std::move(L.adjwgt);
    c->vwgt = std::move(L.vwgt);
    c->m = c->xadj.back();
    c->set_tvwgt();
    c->finer = cur;
    cur = c;
  }
  return cur;
}

void compute_2way(Part2 &g)  // cu:1895-1944
{
  g.pwgts[0] = g.pwgts[1] = 0;
  std::fill(g.bndptr.begin(), g.bndptr.end(), -1);
  g.nbnd = 0;
  for (int i = 0; i < g.n; ++i) g.pwgts[g.where[i]] += g.vwgt[i];
  par_for(g.n, g.m, g.xadj.data(), [&](int lo, int hi) {
    for (int i = lo; i < hi; ++i) {
      int in = 0, ex = 0;
      const int me = g.where[i];
      for (int64_t j = g.xadj[i]; j < g.xadj[i + 1]; ++j)
        (g.where[g.adjncy[j]] == me ? in : ex) += g.adjwgt[j];
      g.id[i] = in;
      g.ed[i] = ex;
    }
  });
  int cut = 0;
  for (int i = 0; i < g.n; ++i)
    if (g.ed[i] > 0 || g.xadj[i] == g.xadj[i + 1]) {
      g.bnd_insert(i);
      cut += g.ed[i];
    }
  g.mincut = cut / 2;
}

// Move higain to `to` and update neighbour gains (shared by balance and FM).
template <typename OnNeighbour>
void flip(Part2 &g, int h, int to, OnNeighbour &&on_nb)
{
  g.where[h] = to;
  std::swap(g.id[h], g.ed[h]);
  if (g.ed[h] == 0 && g.xadj[h] < g.xadj[h + 1]) g.bnd_delete(h);
  for (int64_t j = g.xadj[h]; j < g.xadj[h + 1]; ++j) {
    const int k = g.adjncy[j];
    const int kw = to == g.where[k] ? g.adjwgt[j] : -g.adjwgt[j];
    g.id[k] += kw;
    g.ed[k] -= kw;
    on_nb(k);
  }
}

void balance_moves(Part2 &g, const float *ntp)  // cu:2247-2362
{
  const int n = g.n;
  const int t0 = (int)(g.tvwgt * ntp[0]);
  const int tp[2] = {t0, g.tvwgt - t0};
  const int mindiff = abs(tp[0] - g.pwgts[0]);
  const int from = g.pwgts[0] < tp[0] ? 1 : 0, to = 1 - from;
  std::vector<int> moved(n, -1), perm(std::max(n, 1));
  MaxQueue q(n);
  rand_perm(g.nbnd, perm.data(), g.nbnd / 5, true);
  auto eligible = [&](int k) { return moved[k] == -1 && g.where[k] == from && g.vwgt[k] <= mindiff; };
  for (int ii = 0; ii < g.nbnd; ++ii) {
    const int v = g.bndlist[perm[ii]];
    if (g.where[v] == from && g.vwgt[v] <= mindiff) q.push(v, g.ed[v] - g.id[v]);
  }
  int cut = g.mincut;
  for (int s = 0; s < n; ++s) {
    const int h = q.pop();
    if (h == -1 || g.pwgts[to] + g.vwgt[h] > tp[to]) break;
    cut -= g.ed[h] - g.id[h];
    g.pwgts[to] += g.vwgt[h];
    g.pwgts[from] -= g.vwgt[h];
    moved[h] = s;
    flip(g, h, to, [&](int k) {
      if (g.bndptr[k] != -1) {
        if (g.ed[k] == 0) {
          g.bnd_delete(k);
          if (eligible(k)) q.erase(k);
        } else if (eligible(k)) {
          q.update(k, g.ed[k] - g.id[k]);
        }
      } else if (g.ed[k] > 0) {
        g.bnd_insert(k);
        if (eligible(k)) q.push(k, g.ed[k] - g.id[k]);
      }
    });
  }
  g.mincut = cut;
}

void balance_2way(Part2 &g, Settings &s, const float *ntp)  // cu:2367-2378 (2way_bal)
{
  float worst = -1.0f;
  for (int j = 0; j < 2; ++j) worst = std::max(worst, g.pwgts[j] * s.balance[j] - s.ub);
  if (worst <= 0) return;
  if (fabsf(ntp[0] * g.tvwgt - g.pwgts[0]) < 3 * g.tvwgt / g.n) return;
  balance_moves(g, ntp);
}

void fm_refine(Part2 &g, const float *ntp, int niter)  // cu:2367-2531
{
  const int n = g.n;
  const int t0 = (int)(g.tvwgt * ntp[0]);
  const int tp[2] = {t0, g.tvwgt - t0};
  const double lim = std::min(std::max(0.01 * n, 15.0), 100.0);
  const int limit = (int)lim;
  const int tot = g.pwgts[0] + g.pwgts[1];
  const int avgvwgt = std::min(tot / 20, 2 * tot / n);
  tl_fm.reserve(n);
  std::vector<int> &moved = tl_fm.moved;  // all -1 between passes
  std::vector<int> &swaps = tl_fm.swaps, &perm = tl_fm.perm;
  MaxQueue &q0 = tl_fm.q[0], &q1 = tl_fm.q[1];
  MaxQueue *qs[2] = {&q0, &q1};
  struct Empty {  // the queues go back empty, also on an early return
    MaxQueue &a, &b;
    ~Empty()
    {
      a.clear();
      b.clear();
    }
  } empty_on_exit{q0, q1};
  const int origdiff = abs(tp[0] - g.pwgts[0]);
  for (int pass = 0; pass < niter; ++pass) {
    q0.clear();
    q1.clear();
    int best_at = -1, newcut, mincut, initcut;
    newcut = mincut = initcut = g.mincut;
    int mindiff = abs(tp[0] - g.pwgts[0]);
    rand_perm(g.nbnd, perm.data(), g.nbnd, true);
    for (int ii = 0; ii < g.nbnd; ++ii) {
      const int v = g.bndlist[perm[ii]];
      qs[g.where[v]]->push(v, g.ed[v] - g.id[v]);
    }
    int ns = 0;
    for (; ns < n; ++ns) {
      const int from = (tp[0] - g.pwgts[0] < tp[1] - g.pwgts[1]) ? 0 : 1, to = 1 - from;
      const int h = qs[from]->pop();
      if (h == -1) break;
      newcut -= g.ed[h] - g.id[h];
      g.pwgts[to] += g.vwgt[h];
      g.pwgts[from] -= g.vwgt[h];
      const int diff = abs(tp[0] - g.pwgts[0]);
      if ((newcut < mincut && diff <= origdiff + avgvwgt) || (newcut == mincut && diff < mindiff)) {
        mincut = newcut;
        mindiff = diff;
        best_at = ns;
      } else if (ns - best_at > limit) {
        newcut += g.ed[h] - g.id[h];
        g.pwgts[from] += g.vwgt[h];
        g.pwgts[to] -= g.vwgt[h];
        break;
      }
      moved[h] = ns;
      swaps[ns] = h;
      flip(g, h, to, [&](int k) {
        if (g.bndptr[k] != -1) {
          if (g.ed[k] == 0) {
            g.bnd_delete(k);
            if (moved[k] == -1) qs[g.where[k]]->erase(k);
          } else if (moved[k] == -1) {
            qs[g.where[k]]->update(k, g.ed[k] - g.id[k]);
          }
        } else if (g.ed[k] > 0) {
          g.bnd_insert(k);
          if (moved[k] == -1) qs[g.where[k]]->push(k, g.ed[k] - g.id[k]);
        }
      });
    }
    for (int i = 0; i < ns; ++i) moved[swaps[i]] = -1;
    for (--ns; ns > best_at; --ns) {  // roll back past the best prefix
      const int h = swaps[ns];
      const int to = g.where[h] = (g.where[h] + 1) % 2;
      std::swap(g.id[h], g.ed[h]);
      if (g.ed[h] == 0 && g.bndptr[h] != -1 && g.xadj[h] < g.xadj[h + 1]) g.bnd_delete(h);
      else if (g.ed[h] > 0 && g.bndptr[h] == -1) g.bnd_insert(h);
      g.pwgts[to] += g.vwgt[h];
      g.pwgts[(to + 1) % 2] -= g.vwgt[h];
      for (int64_t j = g.xadj[h]; j < g.xadj[h + 1]; ++j) {
        const int k = g.adjncy[j];
        const int kw = to == g.where[k] ? g.adjwgt[j] : -g.adjwgt[j];
        g.id[k] += kw;
        g.ed[k] -= kw;
        if (g.bndptr[k] != -1 && g.ed[k] == 0) g.bnd_delete(k);
        if (g.bndptr[k] == -1 && g.ed[k] > 0) g.bnd_insert(k);
      }
    }
    g.mincut = mincut;
    if (best_at <= 0 || mincut == initcut) break;
  }
}

// Unseen vertices for the BFS restarts of grow_bisection: the reference picks
// the k-th unseen vertex in index order with a linear scan (cu:2585-2595),
// which is quadratic on graphs with many components (R-MAT's isolated
// vertices: millions of restarts).  A Fenwick tree over the unseen flags
// gives the same vertex in O(log n).
struct Unseen {
  std::vector<int> t;  // 1-based partial sums
  int n = 0, top = 1;
  void reset(int nn)
  {
    n = nn;
    t.assign(n + 1, 0);
    for (int i = 1; i <= n; ++i) t[i] = i & -i;  // all vertices unseen
    top = 1;
    while (top * 2 <= n) top *= 2;
  }
  void mark(int v)
  {
    for (int i = v + 1; i <= n; i += i & -i) --t[i];
  }
  // index of the unseen vertex with k unseen vertices before it
  int kth(int k) const
  {
    int pos = 0;
    for (int step = top; step > 0; step >>= 1)
      if (pos + step <= n && t[pos + step] <= k) {
        pos += step;
        k -= t[pos];
      }
    return pos;
  }
};

// One region-growing try (cu:2563-2650): BFS from a random vertex into part
// 0, then balancing and FM; leaves the try's partition in g.where, its cut in
// g.mincut.  queue and seen hold n entries.
void grow_try(Part2 &g, Settings &s, const float *ntp, std::vector<int> &queue, std::vector<int> &seen,
              Unseen &unseen)
{
  const int n = g.n;
  const int onemax = (int)(s.ub * g.tvwgt * ntp[1]);
  const int onemin = (int)((1.0 / s.ub) * g.tvwgt * ntp[1]);
  std::fill(g.where.begin(), g.where.end(), 1);
  std::fill(seen.begin(), seen.end(), 0);
  unseen.reset(n);
  int pw0 = 0, pw1 = g.tvwgt;
  queue[0] = rnd(n);
  seen[queue[0]] = 1;
  unseen.mark(queue[0]);
  int first = 0, last = 1, nleft = n - 1;
  bool stalled = false;
  for (;;) {
    if (first == last) {  // component exhausted: restart from a random unseen vertex
      if (nleft == 0 || stalled) break;
      const int i = unseen.kth(rnd(nleft));
      queue[0] = i;
      seen[i] = 1;
      unseen.mark(i);
      first = 0;
      last = 1;
      --nleft;
    }
    const int i = queue[first++];
    if (pw0 > 0 && pw1 - g.vwgt[i] < onemin) {
      stalled = true;
      continue;
    }
    g.where[i] = 0;
    pw0 += g.vwgt[i];
    pw1 -= g.vwgt[i];
    if (pw1 <= onemax) break;
    stalled = false;
    for (int64_t j = g.xadj[i]; j < g.xadj[i + 1]; ++j) {
      const int k = g.adjncy[j];
      if (!seen[k]) {
        queue[last++] = k;
        seen[k] = 1;
        unseen.mark(k);
        --nleft;
      }
    }
  }
  compute_2way(g);
  balance_2way(g, s, ntp);
  fm_refine(g, ntp, s.niter);
}

// Tries run concurrently (parallel form) on graphs of this many entries.
constexpr int64_t GROW_PAR_MIN_M = 1 << 20;

void grow_bisection(Part2 &g, Settings &s, const float *ntp, int niparts)  // cu:2536-2661
{
  const int n = g.n;
  g.alloc2way();
  if (s.try_seed != 0) {
    // parallel form: try t draws from its own stream, on its own copy of the
    // graph's 2-way state; the first try with the lowest cut wins, as in the
    // serial loop
    std::vector<std::unique_ptr<Part2>> tg(niparts);
    auto one = [&](int t) {
      Part2 &x = *tg[t];
      RandStream stream(s.try_seed + 0x9E3779B9u * (unsigned)(t + 1));
      RandStream *const outer = g_stream;
      g_stream = &stream;
      std::vector<int> queue(std::max(n, 1)), seen(n);
      Unseen unseen;
      Settings st = s;
      grow_try(x, st, ntp, queue, seen, unseen);
      g_stream = outer;
    };
    for (int t = 0; t < niparts; ++t) {
      tg[t].reset(new Part2());
      Part2 &x = *tg[t];
      x.n = g.n;
      x.m = g.m;
      x.tvwgt = g.tvwgt;
      if (g.m < GROW_PAR_MIN_M) {  // the inline tries read g's arrays through copies too
        x.xadj = g.xadj;
        x.adjncy = g.adjncy;
        x.adjwgt = g.adjwgt;
        x.vwgt = g.vwgt;
      }
      x.alloc2way();
    }
    if (g.m >= GROW_PAR_MIN_M) {
      // try 0 reads g's graph arrays in place (lent to it until the tries
      // end; the tries only read them); the others copy them on their threads
      Part2 &x0 = *tg[0];
      struct Lend {
        Part2 &a, &b;
        void swap()
        {
          a.xadj.swap(b.xadj);
          a.adjncy.swap(b.adjncy);
          a.adjwgt.swap(b.adjwgt);
          a.vwgt.swap(b.vwgt);
        }
        Lend(Part2 &p, Part2 &q) : a(p), b(q) { swap(); }
        ~Lend() { swap(); }
      } lend(g, x0);
      Threads th;  // (joined before the arrays go back)
      for (int t = 1; t < niparts; ++t)
        th.run([&, t]() {
          Part2 &x = *tg[t];
          x.xadj = x0.xadj;
          x.adjncy = x0.adjncy;
          x.adjwgt = x0.adjwgt;
          x.vwgt = x0.vwgt;
          one(t);
        });
      one(0);
      th.join();
    } else {
      for (int t = 0; t < niparts; ++t) one(t);
    }
    int bt = 0;
    for (int t = 1; t < niparts; ++t)
      if (tg[t]->mincut < tg[bt]->mincut) bt = t;
    g.mincut = tg[bt]->mincut;
    g.where = tg[bt]->where;
    return;
  }
  std::vector<int> best(n), queue(std::max(n, 1)), seen(n);
  Unseen unseen;
  int bestcut = 0;
  for (int it = 0; it < niparts; ++it) {
    grow_try(g, s, ntp, queue, seen, unseen);
    if (it == 0 || bestcut > g.mincut) {
      bestcut = g.mincut;
      best = g.where;
      if (bestcut == 0) break;
    }
  }
  g.mincut = bestcut;
  g.where = best;
}

void project_2way(Part2 &g)  // cu:2693-2764
{
  Part2 &c = *g.coarser;
  g.alloc2way();
  for (int i = 0; i < g.n; ++i) {
    const int k = g.cmap[i];
    g.where[i] = c.where[k];
    g.cmap[i] = c.bndptr[k];
  }
  g.nbnd = 0;
  par_for(g.n, g.m, g.xadj.data(), [&](int lo, int hi) {
    for (int i = lo; i < hi; ++i) {
      int in = 0, ex = 0;
      if (g.cmap[i] == -1) {
        for (int64_t j = g.xadj[i]; j < g.xadj[i + 1]; ++j) in += g.adjwgt[j];
      } else {
        const int me = g.where[i];
        for (int64_t j = g.xadj[i]; j < g.xadj[i + 1]; ++j)
          (g.where[g.adjncy[j]] == me ? in : ex) += g.adjwgt[j];
      }
      g.id[i] = in;
      g.ed[i] = ex;
    }
  });
  for (int i = 0; i < g.n; ++i)
    if (g.ed[i] > 0 || g.xadj[i] == g.xadj[i + 1]) g.bnd_insert(i);
  g.mincut = c.mincut;
  g.pwgts[0] = c.pwgts[0];
  g.pwgts[1] = c.pwgts[1];
  g.coarser.reset();
}

int multilevel_bisect(Part2 &g, Settings &s, const float *tp)  // cu:2806-2829
{
  const float rev = (float)(1.0 / (g.tvwgt > 0 ? g.tvwgt : 1));
  for (int i = 0; i < 2; ++i) s.balance[i] = rev / tp[i];
  Part2 *c = s.dev && g.m >= STEP_DEVICE_MIN_M ? dev_coarsen(g, s) : cpu_coarsen(g, s);
  grow_bisection(*c, s, tp, 5);
  Part2 *x = c;  // cu:2769-2802
  compute_2way(*x);
  for (;;) {
    balance_2way(*x, s, tp);
    fm_refine(*x, tp, s.niter);
    if (x == &g) break;
    x = x->finer;
    project_2way(*x);
  }
  return g.mincut;
}

void split(Part2 &g, Part2 &l, Part2 &r)  // cu:2834-2959
{
  Part2 *side[2] = {&l, &r};
  std::vector<int> rename(g.n);
  int cnt[2] = {0, 0};
  for (int i = 0; i < g.n; ++i) rename[i] = cnt[g.where[i]]++;
  // entries each vertex keeps: all of an interior vertex's, the same-side
  // ones of a boundary vertex's
  std::vector<int64_t> kept(g.n);
  par_for(g.n, g.m, g.xadj.data(), [&](int lo, int hi) {
    for (int i = lo; i < hi; ++i) {
      if (g.bndptr[i] == -1) {
        kept[i] = g.xadj[i + 1] - g.xadj[i];
        continue;
      }
      const int p = g.where[i];
      int64_t c = 0;
      for (int64_t j = g.xadj[i]; j < g.xadj[i + 1]; ++j) c += g.where[g.adjncy[j]] == p;
      kept[i] = c;
    }
  });
  int64_t off[2] = {0, 0};
  for (int p = 0; p < 2; ++p) {
    side[p]->n = cnt[p];
    side[p]->xadj.resize(cnt[p] + 1);
    side[p]->vwgt.resize(cnt[p]);
    side[p]->label.resize(cnt[p]);
  }
  for (int i = 0; i < g.n; ++i) {
    const int p = g.where[i];
    Part2 &s = *side[p];
    s.xadj[rename[i]] = off[p];
    off[p] += kept[i];
    s.vwgt[rename[i]] = g.vwgt[i];
    s.label[rename[i]] = g.label[i];
  }
  for (int p = 0; p < 2; ++p) {
    side[p]->xadj[cnt[p]] = off[p];
    side[p]->adjncy.resize(off[p]);
    side[p]->adjwgt.resize(off[p]);
  }
  par_for(g.n, g.m, g.xadj.data(), [&](int lo, int hi) {
    for (int i = lo; i < hi; ++i) {
      const int p = g.where[i];
      Part2 &s = *side[p];
      const bool interior = g.bndptr[i] == -1;
      int64_t o = s.xadj[rename[i]];
      for (int64_t j = g.xadj[i]; j < g.xadj[i + 1]; ++j) {
        const int k = g.adjncy[j];
        if (interior || g.where[k] == p) {
          s.adjncy[o] = rename[k];
          s.adjwgt[o++] = g.adjwgt[j];
        }
      }
    }
  });
  for (int p = 0; p < 2; ++p) {
    side[p]->m = side[p]->xadj.back();
    side[p]->set_tvwgt();
  }
}

int recursive_bisect(std::unique_ptr<Part2> g, Settings &s, int nparts, int *part, float *tp,
                     int fpart)  // cu:2964-3025
{
  if (g->n == 0) return 0;
  float sum = 0;
  for (int i = 0; i < (nparts >> 1); ++i) sum += tp[i];
  const float tp2[2] = {sum, 1.0f - sum};
  int obj = multilevel_bisect(*g, s, tp2);
  for (int i = 0; i < g->n; ++i) part[g->label[i]] = g->where[i] + fpart;
  std::unique_ptr<Part2> l, r;
  if (nparts > 2) {
    l.reset(new Part2());
    r.reset(new Part2());
    split(*g, *l, *r);
  }
  g.reset();
  const float f0 = (float)(1.0 / sum), f1 = (float)(1.0 / (1.0 - sum));
  for (int i = 0; i < (nparts >> 1); ++i) tp[i] *= f0;
  for (int i = nparts >> 1; i < nparts; ++i) tp[i] *= f1;
  if (nparts > 3) {
    obj += recursive_bisect(std::move(l), s, nparts >> 1, part, tp, fpart);
    obj += recursive_bisect(std::move(r), s, nparts - (nparts >> 1), part, tp + (nparts >> 1),
                            fpart + (nparts >> 1));
  } else if (nparts == 3) {
    obj += recursive_bisect(std::move(r), s, nparts - (nparts >> 1), part, tp + (nparts >> 1),
                            fpart + (nparts >> 1));
  }
  return obj;
}

// The parallel form: the step for parts [fpart, fpart + nparts) draws from a
// stream seeded by (fpart, nparts); its halves run on a new thread and this
// one when the half is large and a thread is free, inline otherwise.
constexpr int PAR_TASK_MIN_N = 4096;

struct ParCtx {
  std::atomic<int> threads{0};
  int max_threads = 1;
};

unsigned step_seed(int fpart, int nparts)
{
  uint64_t z = ((uint64_t)(uint32_t)fpart << 32 | (uint32_t)nparts) + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return (unsigned)((z ^ (z >> 31)) >> 32) | 1u;
}

int recursive_bisect_par(std::unique_ptr<Part2> g, Settings s, int nparts, int *part, float *tp, int fpart,
                         ParCtx &pc)
{
  if (g->n == 0) return 0;
  RandStream stream(step_seed(fpart, nparts));
  RandStream *const outer = g_stream;
  g_stream = &stream;
  s.try_seed = step_seed(fpart + nparts, ~(unsigned)nparts);  // the region-growing tries' streams
  float sum = 0;
  for (int i = 0; i < (nparts >> 1); ++i) sum += tp[i];
  const float tp2[2] = {sum, 1.0f - sum};
  int obj = multilevel_bisect(*g, s, tp2);
  for (int i = 0; i < g->n; ++i) part[g->label[i]] = g->where[i] + fpart;
  std::unique_ptr<Part2> l, r;
  if (nparts > 2) {
    l.reset(new Part2());
    r.reset(new Part2());
    split(*g, *l, *r);
  }
  g.reset();
  g_stream = outer;
  const float f0 = (float)(1.0 / sum), f1 = (float)(1.0 / (1.0 - sum));
  for (int i = 0; i < (nparts >> 1); ++i) tp[i] *= f0;
  for (int i = nparts >> 1; i < nparts; ++i) tp[i] *= f1;
  const int nl = nparts >> 1, nr = nparts - nl;
  if (nparts > 3) {
    int objl = 0;
    Threads th;
    bool spawned = false;
    if (l->n >= PAR_TASK_MIN_N && pc.threads.fetch_add(1) < pc.max_threads) {
      spawned = true;
      th.run([&, s]() { objl = recursive_bisect_par(std::move(l), s, nl, part, tp, fpart, pc); });
    } else {
      if (l->n >= PAR_TASK_MIN_N) pc.threads.fetch_sub(1);
      objl = recursive_bisect_par(std::move(l), s, nl, part, tp, fpart, pc);
    }
    const int objr = recursive_bisect_par(std::move(r), s, nr, part, tp + nl, fpart + nl, pc);
    th.join();
    if (spawned) pc.threads.fetch_sub(1);
    obj += objl + objr;
  } else if (nparts == 3) {
    obj += recursive_bisect_par(std::move(r), s, nr, part, tp + nl, fpart + nl, pc);
  }
  return obj;
}

int initial_partition_par(const HostGraph &hg, int nparts, std::vector<int> &where, StepCoarsener *dev)
{
  // the graph without its isolated vertices
  std::unique_ptr<Part2> g(new Part2());
  std::vector<int> iso;
  g->xadj.reserve(hg.n + 1);
  g->xadj.push_back(0);
  for (int i = 0; i < hg.n; ++i) {
    const int64_t b = hg.xadj[i], e = hg.xadj[i + 1];
    if (b == e) {
      iso.push_back(i);
      continue;
    }
    g->xadj.push_back(g->xadj.back() + (e - b));
    g->vwgt.push_back(hg.vwgt[i]);
    g->label.push_back(i);
  }
  g->n = (int)g->label.size();
  std::vector<int> newid(hg.n, -1);
  for (int i = 0; i < g->n; ++i) newid[g->label[i]] = i;
  g->adjncy.resize(hg.m());
  g->adjwgt.resize(hg.m());
  par_for(g->n, hg.m(), g->xadj.data(), [&](int lo, int hi) {
    for (int i = lo; i < hi; ++i) {
      const int64_t b = hg.xadj[g->label[i]], o = g->xadj[i];
      for (int64_t j = 0; j < g->xadj[i + 1] - o; ++j) {
        g->adjncy[o + j] = newid[hg.adjncy[b + j]];
        g->adjwgt[o + j] = hg.adjwgt[b + j];
      }
    }
  });
  g->m = g->xadj.back();
  g->set_tvwgt();
  ParCtx pc;
  const unsigned hw = std::thread::hardware_concurrency();
  pc.max_threads = (int)std::max(1u, std::min(hw ? hw : 1u, 16u)) - 1;
  std::vector<float> tp(nparts, (float)(1.0 / nparts));
  int obj = 0;
  Settings s;
  s.dev = dev;
  if (g->n > 0) obj = recursive_bisect_par(std::move(g), s, nparts, where.data(), tp.data(), 0, pc);
  // isolated vertices, in index order, each to the lightest part (the
  // lowest part id among equals)
  std::vector<int64_t> pw(nparts, 0);
  std::vector<char> isolated(hg.n, 0);
  for (int v : iso) isolated[v] = 1;
  for (int i = 0; i < hg.n; ++i)
    if (!isolated[i]) pw[where[i]] += hg.vwgt[i];
  using Item = std::pair<int64_t, int>;
  std::priority_queue<Item, std::vector<Item>, std::greater<Item>> heap;
  for (int p = 0; p < nparts; ++p) heap.push({pw[p], p});
  for (int v : iso) {
    Item t = heap.top();
    heap.pop();
    where[v] = t.second;
    t.first += hg.vwgt[v];
    heap.push(t);
  }
  return obj;
}

}  // namespace

int initial_partition(const HostGraph &hg, int nparts, std::vector<int> &where, InitForm form,
                      StepCoarsener *dev)
{
  where.assign(hg.n, 0);
  if (nparts <= 1 || hg.n == 0) return 0;
  if (form == InitForm::PARALLEL || (form == InitForm::AUTO && hg.n >= PAR_MIN_N))
    return initial_partition_par(hg, nparts, where, dev);
  RandStream stream;  // a fresh stream per call: the reference's first-call sequence
  g_stream = &stream;
  std::unique_ptr<Part2> g(new Part2());
  g->n = hg.n;
  g->m = hg.m();
  g->xadj = hg.xadj;
  g->adjncy = hg.adjncy;
  g->adjwgt = hg.adjwgt;
  g->vwgt = hg.vwgt;
  g->label.resize(hg.n);
  for (int i = 0; i < hg.n; ++i) g->label[i] = i;
  g->set_tvwgt();
  Settings s;
  std::vector<float> tp(nparts, (float)(1.0 / nparts));
  return recursive_bisect(std::move(g), s, nparts, where.data(), tp.data(), 0);
}

}  // namespace cmh
