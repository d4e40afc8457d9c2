// cm_engine.hip -- contexts, device memory pool, the multilevel coarsening
// driver (replaces cuMetis_coarsen, cu:1824-1850) and the C ABI.
#include <string.h>

#include <algorithm>
#include <thread>
#include <atomic>
#include <cmath>
#include <vector>

#include "cm_common.hpp"
#include "cm_internal.hpp"
#include "cm_contract_hash.hpp"
#include "cm_segsort.hpp"
#include "cm_tuned.hpp"

namespace cm {

// ------------------------------------------------------------------- pool
static size_t size_class(size_t bytes)
{
  size_t c = 256;
  while (c < bytes) c <<= 1;
  return c;
}

Pool::~Pool()
{
  for (auto &kv : free_)
    for (void *p : kv.second) (void)hipFree(p);
  for (auto &kv : live_) (void)hipFree(kv.first);
}

void *Pool::alloc(size_t bytes)
{
  std::lock_guard<std::recursive_mutex> lk(mu_);
  const size_t c = size_class(bytes ? bytes : 1);
  auto it = free_.find(c);
  void *p = nullptr;
  if (it != free_.end() && !it->second.empty()) {
    p = *it->second.begin();
    it->second.erase(it->second.begin());
  } else {
    hipError_t e = hipMalloc(&p, c);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      trim();
      e = hipMalloc(&p, c);
      if (e != hipSuccess) {
        (void)hipGetLastError();
        throw Error(CM_ERR_NOMEM, "device allocation of " + std::to_string(c) + " bytes failed");
      }
    }
    reserved_ += c;
  }
  live_[p] = c;
  in_use_ += c;
  return p;
}

void Pool::release(void *p)
{
  if (!p) return;
  std::lock_guard<std::recursive_mutex> lk(mu_);
  auto it = live_.find(p);
  if (it == live_.end()) return;
  in_use_ -= it->second;
  free_[it->second].insert(p);
  live_.erase(it);
}

void Pool::trim()
{
  std::lock_guard<std::recursive_mutex> lk(mu_);
  for (auto &kv : free_) {
    for (void *p : kv.second) {
      (void)hipFree(p);
      reserved_ -= kv.first;
    }
    kv.second.clear();
  }
}

// ----------------------------------------------------------- small kernels
__global__ void k_fill_u32(uint32_t *p, uint32_t v, int64_t count)
{
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void k_fill_u64(uint64_t *p, uint64_t v, int64_t count)
{
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ __launch_bounds__(256) void k_sum_i32(int64_t n, const int32_t *a, unsigned long long *out)
{
  int64_t s = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    s += a[i];
  s = wave_sum<int64_t>(s);
  if (lane() == 0) atomicAdd(out, (unsigned long long)s);
}

void fill_u32(cm_ctx *ctx, void *p, uint32_t v, int64_t count)
{
  if (count <= 0) return;
  launch(ctx, "fill", 4.0 * count, k_fill_u32, dim3((unsigned)std::min<int64_t>(nblocks(count, 256), 4096)),
         dim3(256), 0, (uint32_t *)p, v, count);
}

void fill_u64(cm_ctx *ctx, void *p, uint64_t v, int64_t count)
{
  if (count <= 0) return;
  launch(ctx, "fill", 8.0 * count, k_fill_u64, dim3((unsigned)std::min<int64_t>(nblocks(count, 256), 4096)),
         dim3(256), 0, (uint64_t *)p, v, count);
}

void ensure_status(cm_ctx *ctx, size_t words)
{
  if (words <= ctx->status_cap) return;
  if (ctx->d_status) ctx->pool.release(ctx->d_status);
  size_t cap = 1024;
  while (cap < words) cap <<= 1;
  ctx->d_status = (uint64_t *)ctx->pool.alloc(cap * sizeof(uint64_t));
  ctx->status_cap = cap;
  fill_u64(ctx, ctx->d_status, 0ull, (int64_t)cap);
  ctx->epoch = 0;
}

void ensure_host_flags(cm_ctx *ctx, size_t words)
{
  if (words <= ctx->h_flags_cap) return;
  size_t cap = std::max<size_t>(words, 1024);
  if (ctx->h_flags) {
    sync(ctx);  // a queued copy may still target the old buffer
    CM_HIP(hipHostFree(ctx->h_flags));
    ctx->h_flags = nullptr;
    ctx->h_flags_cap = 0;
  }
  CM_HIP(hipHostMalloc((void **)&ctx->h_flags, sizeof(unsigned) * cap, hipHostMallocDefault));
  ctx->h_flags_cap = cap;
  CM_HIP(hipHostGetDevicePointer((void **)&ctx->h_flags_dev, ctx->h_flags, 0));
}

const int32_t *unit_weights(cm_ctx *ctx, int64_t count)
{
  if (count > ctx->ones_cap) {  // the pool keeps the old block until the context dies
    const int64_t cap = std::max<int64_t>(count, std::max<int64_t>(2 * ctx->ones_cap, 1 << 16));
    int32_t *p = (int32_t *)ctx->pool.alloc(sizeof(int32_t) * cap);
    fill_u32(ctx, p, 1u, cap);
    ctx->ones = p;
    ctx->ones_cap = cap;
  }
  return ctx->ones;
}

Readback readback_begin(cm_ctx *ctx, int first, int count, const unsigned *ovf)
{
  Readback rb;
  rb.dst = ctx->h_small_dev;
  rb.dst_flags = ctx->h_flags_dev;
  rb.flags = ctx->rb_flags;
  rb.nflags = ctx->rb_flags ? (int32_t)ctx->rb_nflags : 0;
  rb.ovf = ovf;
  rb.first = (int16_t)first;
  rb.count = (int16_t)count;
  rb.tag = ++ctx->rb_tag;
  ctx->rb_flags = nullptr;
  ctx->rb_nflags = 0;
  return rb;
}

void readback_end(cm_ctx *ctx, const Readback &rb)
{
  if (ctx->before_wait) {
    std::function<void()> f;
    f.swap(ctx->before_wait);
    f();
  }
  if (ctx->rb_early && !ctx->timing) {
    // The kernel stores the tag after every other word (a system-scope
    // release), so the words are final once the tag is here: the host goes
    // on without waiting for the stream's completion signal.  A stream that
    // completes without sending the tag falls through to the copy below.
    const volatile int64_t *tag = ctx->h_small + RB_TAG_SLOT;
    for (unsigned spin = 1;; ++spin) {
      if (*tag == rb.tag) {
        std::atomic_thread_fence(std::memory_order_acquire);
        return;
      }
      if ((spin & 255) == 0 && hipStreamQuery(ctx->stream) == hipSuccess) break;
    }
  }
  sync(ctx);
  // the stream's completion makes the stores visible; a stale tag would
  // mean they were not, so the words are then copied the ordinary way
  if (ctx->h_small[RB_TAG_SLOT] != rb.tag) {
    CM_HIP(hipMemcpyAsync(ctx->h_small + rb.first, ctx->d_small + rb.first, sizeof(int64_t) * rb.count,
                          hipMemcpyDeviceToHost, ctx->stream));
    if (rb.ovf)
      CM_HIP(hipMemcpyAsync(ctx->h_small + 6, rb.ovf, sizeof(unsigned), hipMemcpyDeviceToHost, ctx->stream));
    if (rb.nflags)
      CM_HIP(hipMemcpyAsync(ctx->h_flags, rb.flags, sizeof(unsigned) * rb.nflags, hipMemcpyDeviceToHost,
                            ctx->stream));
    sync(ctx);
    if (rb.ovf) ctx->h_small[6] = (int64_t)(uint32_t)ctx->h_small[6];
    ctx->stats["readback_fallback"].count += 1;
  }
}

uint32_t lookback_begin(cm_ctx *ctx, size_t words)
{
  ensure_status(ctx, words);
  if (++ctx->epoch > (uint32_t)ST_EPOCH_MASK) {  // wrap: clear and restart
    fill_u64(ctx, ctx->d_status, 0ull, (int64_t)ctx->status_cap);
    ctx->epoch = 1;
  }
  return ctx->epoch;
}

void close_timing_run(cm_ctx *ctx)
{
  if (ctx->pending.empty() || ctx->pending.back().b != nullptr) return;
  hipEvent_t b = ctx->ev_get();
  CM_HIP(hipEventRecord(b, ctx->stream));
  ctx->pending.back().b = b;
}

void sync(cm_ctx *ctx)
{
  if (ctx->timing) close_timing_run(ctx);
  CM_HIP(hipStreamSynchronize(ctx->stream));
  if (ctx->timing) ctx->collect_timing();
}

// Large host<->device copies go through two pinned chunks of the context:
// the DMA of one chunk runs while host threads copy the other in or out, so
// the host side's fresh pages are first touched by several threads at once.
// (A pageable copy into freshly allocated vectors ran at 4.3 GB/s.)
constexpr size_t STAGE_CHUNK = (size_t)32 << 20;
constexpr size_t STAGE_MIN = (size_t)8 << 20;  // smaller copies go direct

static void host_copy(void *dst, const void *src, size_t bytes)
{
  const unsigned hw = std::thread::hardware_concurrency();
  const int nt = (int)std::min<size_t>(std::min(hw ? hw : 1u, 8u), std::max<size_t>(1, bytes >> 22));
  if (nt <= 1) {
    memcpy(dst, src, bytes);
    return;
  }
  std::vector<std::thread> th;
  auto part = [&](int t) {
    const size_t a = bytes * t / nt, b = bytes * (t + 1) / nt;
    memcpy((char *)dst + a, (const char *)src + a, b - a);
  };
  for (int t = 1; t < nt; ++t) th.emplace_back(part, t);
  part(0);
  for (auto &x : th) x.join();
}

static void stage_init(cm_ctx *ctx)
{
  if (ctx->h_stage) return;
  for (hipEvent_t &e : ctx->stage_ev)
    if (!e) CM_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  CM_HIP(hipHostMalloc((void **)&ctx->h_stage, 2 * STAGE_CHUNK, hipHostMallocDefault));
}

// Synchronous device-to-host copy.
void download(cm_ctx *ctx, void *dst, const void *src, size_t bytes)
{
  if (bytes < STAGE_MIN) {
    if (bytes) CM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    sync(ctx);
    return;
  }
  stage_init(ctx);
  const size_t nc = (bytes + STAGE_CHUNK - 1) / STAGE_CHUNK;
  auto len = [&](size_t k) { return std::min(STAGE_CHUNK, bytes - k * STAGE_CHUNK); };
  auto issue = [&](size_t k) {
    char *buf = ctx->h_stage + (k & 1) * STAGE_CHUNK;
    CM_HIP(hipMemcpyAsync(buf, (const char *)src + k * STAGE_CHUNK, len(k), hipMemcpyDeviceToHost, ctx->stream));
    CM_HIP(hipEventRecord(ctx->stage_ev[k & 1], ctx->stream));
  };
  issue(0);
  for (size_t k = 0; k < nc; ++k) {
    if (k + 1 < nc) issue(k + 1);  // its chunk's last reader, chunk k - 1, is copied out
    CM_HIP(hipEventSynchronize(ctx->stage_ev[k & 1]));
    host_copy((char *)dst + k * STAGE_CHUNK, ctx->h_stage + (k & 1) * STAGE_CHUNK, len(k));
  }
}

// Host-to-device copy; queued on the stream when it returns (the staging
// chunks are reused only after their DMA has finished).
void upload(cm_ctx *ctx, void *dst, const void *src, size_t bytes)
{
  if (bytes < STAGE_MIN) {
    if (bytes) CM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    return;
  }
  stage_init(ctx);
  CM_HIP(hipStreamSynchronize(ctx->stream));  // the chunks' earlier users are done
  const size_t nc = (bytes + STAGE_CHUNK - 1) / STAGE_CHUNK;
  for (size_t k = 0; k < nc; ++k) {
    const size_t n = std::min(STAGE_CHUNK, bytes - k * STAGE_CHUNK);
    char *buf = ctx->h_stage + (k & 1) * STAGE_CHUNK;
    if (k >= 2) CM_HIP(hipEventSynchronize(ctx->stage_ev[k & 1]));  // chunk k - 2's DMA has read buf
    host_copy(buf, (const char *)src + k * STAGE_CHUNK, n);
    CM_HIP(hipMemcpyAsync((char *)dst + k * STAGE_CHUNK, buf, n, hipMemcpyHostToDevice, ctx->stream));
    CM_HIP(hipEventRecord(ctx->stage_ev[k & 1], ctx->stream));
  }
  CM_HIP(hipStreamSynchronize(ctx->stream));
}


static int64_t device_sum(cm_ctx *ctx, int64_t n, const int32_t *a)
{
  unsigned long long *d = (unsigned long long *)(ctx->d_small + 8);
  fill_u64(ctx, d, 0ull, 1);
  launch(ctx, "sum", 4.0 * n, k_sum_i32, dim3((unsigned)std::min<int64_t>(nblocks(n, 256), 2048)),
         dim3(256), 0, n, a, d);
  CM_HIP(hipMemcpyAsync(ctx->h_small + 8, d, 8, hipMemcpyDeviceToHost, ctx->stream));
  sync(ctx);
  return ctx->h_small[8];
}

// ------------------------------------------------------ input validation
// bad[0] |= 1 offsets, 2 neighbour id out of range, 4 self loop, 8 weight <= 0
constexpr int64_t CHECK_LONG = 256;
template <typename OffT>
__global__ __launch_bounds__(256) void k_check_graph(int64_t n, int64_t m, const OffT *xadj,
                                                     const int32_t *adjncy, const int32_t *adjwgt,
                                                     unsigned *bad)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned f = 0;
  int64_t b = 0, e = 0;
  bool lng = false;  // rows above CHECK_LONG entries (hubs) are checked by the whole wave
  if (i <= n) {
    if (i == 0 && xadj[0] != 0) f |= 1;
    if (i == n) {
      if ((int64_t)xadj[n] != m) f |= 1;
    } else {
      b = (int64_t)xadj[i];
      e = (int64_t)xadj[i + 1];
      if (b > e || b < 0 || e > m) {
        f |= 1;
      } else if (e - b > CHECK_LONG) {
        lng = true;
      } else {
        for (int64_t j = b; j < e; ++j) {
          const int32_t u = adjncy[j];
          if (u < 0 || u >= n) f |= 2;
          else if (u == i) f |= 4;
          if (adjwgt && adjwgt[j] <= 0) f |= 8;
        }
      }
    }
  }
  for (uint64_t lm = __ballot(lng); lm; lm &= lm - 1) {
    const int src = __ffsll((unsigned long long)lm) - 1;
    const int64_t rb = __shfl(b, src, CM_WAVE), re = __shfl(e, src, CM_WAVE), ri = __shfl(i, src, CM_WAVE);
    unsigned g = 0;
    for (int64_t j = rb + lane(); j < re; j += CM_WAVE) {
      const int32_t u = adjncy[j];
      if (u < 0 || u >= n) g |= 2;
      else if (u == ri) g |= 4;
      if (adjwgt && adjwgt[j] <= 0) g |= 8;
    }
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) g |= __shfl_xor(g, d, CM_WAVE);
    if (lane() == src) f |= g;
  }
  if (f) atomicOr(bad, f);
}

// partner in range and match[match[i]] == i
__global__ __launch_bounds__(256) void k_check_match(int64_t n, const int32_t *match, unsigned *bad)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int32_t u = match[i];
  if (u < 0 || u >= n || match[u] != i) atomicOr(bad, 16u);
}

__global__ __launch_bounds__(256) void k_check_equal(int64_t n, const int32_t *a, const int32_t *b,
                                                     unsigned *bad)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && a[i] != b[i]) atomicOr(bad, 32u);
}

static unsigned read_bad(cm_ctx *ctx)
{
  unsigned h = 0;
  CM_HIP(hipMemcpyAsync(&h, ctx->d_ctr + 3, sizeof(unsigned), hipMemcpyDeviceToHost, ctx->stream));
  sync(ctx);
  return h;
}

static void check_graph(cm_ctx *ctx, int64_t n, int64_t m, bool off64, const void *xadj,
                        const int32_t *adjncy, const int32_t *adjwgt)
{
  fill_u32(ctx, ctx->d_ctr + 3, 0u, 1);
  if (off64)
    launch(ctx, "validate_graph", 0.0, k_check_graph<int64_t>, dim3(nblocks(n + 1, 256)), dim3(256), 0,
           n, m, (const int64_t *)xadj, adjncy, adjwgt, ctx->d_ctr + 3);
  else
    launch(ctx, "validate_graph", 0.0, k_check_graph<int32_t>, dim3(nblocks(n + 1, 256)), dim3(256), 0,
           n, m, (const int32_t *)xadj, adjncy, adjwgt, ctx->d_ctr + 3);
  const unsigned f = read_bad(ctx);
  CM_REQUIRE(!(f & 1), CM_ERR_INPUT, "xadj is not a valid offset array for n, m");
  CM_REQUIRE(!(f & 2), CM_ERR_INPUT, "adjncy holds a vertex id outside [0, n)");
  CM_REQUIRE(!(f & 4), CM_ERR_INPUT, "graph has a self loop");
  CM_REQUIRE(!(f & 8), CM_ERR_INPUT, "edge weights must be positive");
}

static void check_match(cm_ctx *ctx, int64_t n, const int32_t *match)
{
  fill_u32(ctx, ctx->d_ctr + 3, 0u, 1);
  launch(ctx, "validate_match", 0.0, k_check_match, dim3(nblocks(n, 256)), dim3(256), 0, n, match,
         ctx->d_ctr + 3);
  CM_REQUIRE(read_bad(ctx) == 0, CM_ERR_INVALID,
             "match must be a complete involution (partner or self for every vertex)");
}

static void check_equal(cm_ctx *ctx, int64_t n, const int32_t *a, const int32_t *b, const char *msg)
{
  fill_u32(ctx, ctx->d_ctr + 3, 0u, 1);
  launch(ctx, "validate_equal", 0.0, k_check_equal, dim3(nblocks(n, 256)), dim3(256), 0, n, a, b,
         ctx->d_ctr + 3);
  CM_REQUIRE(read_bad(ctx) == 0, CM_ERR_INVALID, msg);
}

// ----------------------------------------------------------- graph inputs
static int offset_bytes(int t) { return t == CM_OFF64 ? 8 : 4; }

void validate_view(const cm_graph_view *g)
{
  CM_REQUIRE(g != nullptr, CM_ERR_INVALID, "graph view is NULL");
  CM_REQUIRE(g->n >= 0 && g->m >= 0, CM_ERR_INVALID, "negative graph size");
  CM_REQUIRE(g->offset_type == CM_OFF32 || g->offset_type == CM_OFF64, CM_ERR_INVALID,
             "offset_type must be CM_OFF32 or CM_OFF64");
  CM_REQUIRE(g->n < (int64_t)INT32_MAX, CM_ERR_UNSUPPORTED, "vertex count must fit int32");
  CM_REQUIRE(g->offset_type == CM_OFF64 || g->m <= (int64_t)INT32_MAX, CM_ERR_UNSUPPORTED,
             "more than 2^31-1 entries needs CM_OFF64");
  CM_REQUIRE(g->m <= (int64_t)UINT32_MAX, CM_ERR_UNSUPPORTED, "more than 2^32-1 entries");
  CM_REQUIRE(g->xadj != nullptr, CM_ERR_INVALID, "xadj is NULL");
  CM_REQUIRE(g->m == 0 || g->adjncy != nullptr, CM_ERR_INVALID, "adjncy is NULL");
}

// Level 0 of a hierarchy from a view: device views are referenced (unit
// weights are materialised), host views are copied into pool memory.
// mx[0] = longest row, mx[-1] = largest vertex weight (vwgt != nullptr).
// Tiles of 1024 * MAXDEG_PER_THREAD rows pass running maxima along two look-back chains; the
// last tile writes them (no word updated by two tiles).
constexpr int MAXDEG_PER_THREAD = 8;  // vertices per thread (faster than 4 or 16 on grid and R-MAT)

template <typename OffT>
__global__ __launch_bounds__(1024) void k_max_degree(int64_t n, const OffT *xadj, const int32_t *vwgt, int64_t *mx,
                                                     uint64_t *st, unsigned *ctr, uint32_t epoch)
{
  __shared__ uint64_t sh[16], shw[16];
  const int tile = tile_ticket(ctr);
  uint64_t d = 0, vw = 0;
#pragma unroll
  for (int t = 0; t < MAXDEG_PER_THREAD; ++t) {
    const int64_t v = ((int64_t)tile * 1024 + threadIdx.x) * MAXDEG_PER_THREAD + t;
    if (v < n) {
      const uint64_t x = (uint64_t)(xadj[v + 1] - xadj[v]);
      d = x > d ? x : d;
      if (vwgt) {
        const uint64_t y = (uint64_t)(uint32_t)vwgt[v];
        vw = y > vw ? y : vw;
      }
    }
  }
  d = wave_max<uint64_t>(d);
  vw = wave_max<uint64_t>(vw);
  if (lane() == 0) {
    sh[threadIdx.x >> 6] = d;
    shw[threadIdx.x >> 6] = vw;
  }
  __syncthreads();
  if (threadIdx.x < CM_WAVE) {
    const uint64_t bm = wave_reduce<LbMax>(threadIdx.x < 16 ? sh[threadIdx.x] : 0);
    const uint64_t bw = wave_reduce<LbMax>(threadIdx.x < 16 ? shw[threadIdx.x] : 0);
    const uint64_t pm = lookback<LbMax>(st, tile, bm, epoch);
    const uint64_t pw = lookback<LbMax>(st + gridDim.x, tile, bw, epoch);
    if (threadIdx.x == 0 && tile == (int)gridDim.x - 1) {
      mx[0] = (int64_t)(pm > bm ? pm : bm);
      if (vwgt) mx[-1] = (int64_t)(pw > bw ? pw : bw);
    }
  }
}

// Queues the longest row of a CSR into d_small[4] and, for vwgt != nullptr,
// the largest vertex weight into d_small[3]; the caller reads them back
// (h_small[3..4]) with its next synchronisation.
void queue_max_degree(cm_ctx *ctx, int64_t n, bool off64, const void *xadj, const int32_t *vwgt)
{
  fill_u32(ctx, ctx->d_small + 3, 0u, 4);
  if (n > 0) {
    const int64_t tiles = (n + 1024 * MAXDEG_PER_THREAD - 1) / (1024 * MAXDEG_PER_THREAD);
    const uint32_t ep = lookback_begin(ctx, 2 * tiles);
    if (off64)
      launch(ctx, "max_degree", 0.0, k_max_degree<int64_t>, dim3((unsigned)tiles), dim3(1024), 0, n,
             (const int64_t *)xadj, vwgt, ctx->d_small + 4, ctx->d_status, ctx->d_ctr, ep);
    else
      launch(ctx, "max_degree", 0.0, k_max_degree<int32_t>, dim3((unsigned)tiles), dim3(1024), 0, n,
             (const int32_t *)xadj, vwgt, ctx->d_small + 4, ctx->d_status, ctx->d_ctr, ep);
  }
  CM_HIP(hipMemcpyAsync(ctx->h_small + 3, ctx->d_small + 3, 2 * sizeof(int64_t), hipMemcpyDeviceToHost,
                        ctx->stream));
}

static Level level0(cm_ctx *ctx, const cm_graph_view *g, cm_hier *h, bool validate = true)
{
  Level L;
  L.n = g->n;
  L.m = g->m;
  L.off64 = g->offset_type == CM_OFF64;
  const size_t ob = offset_bytes(g->offset_type);
  auto own = [&](size_t bytes) {
    void *p = ctx->pool.alloc(bytes);
    h->owned.push_back(p);
    return p;
  };
  if (g->on_device) {
    L.xadj = const_cast<void *>(g->xadj);
    L.adjncy = const_cast<int32_t *>(g->adjncy);
    L.adjwgt = const_cast<int32_t *>(g->adjwgt);
    L.vwgt = const_cast<int32_t *>(g->vwgt);
  } else {
    L.xadj = own(ob * (L.n + 1));
    CM_HIP(hipMemcpyAsync(L.xadj, g->xadj, ob * (L.n + 1), hipMemcpyHostToDevice, ctx->stream));
    L.adjncy = (int32_t *)own(4 * (L.m > 0 ? L.m : 1));
    if (L.m)
      CM_HIP(hipMemcpyAsync(L.adjncy, g->adjncy, 4 * L.m, hipMemcpyHostToDevice, ctx->stream));
    if (g->adjwgt) {
      L.adjwgt = (int32_t *)own(4 * (L.m > 0 ? L.m : 1));
      if (L.m)
        CM_HIP(hipMemcpyAsync(L.adjwgt, g->adjwgt, 4 * L.m, hipMemcpyHostToDevice, ctx->stream));
    }
    if (g->vwgt) {
      L.vwgt = (int32_t *)own(4 * (L.n > 0 ? L.n : 1));
      if (L.n) CM_HIP(hipMemcpyAsync(L.vwgt, g->vwgt, 4 * L.n, hipMemcpyHostToDevice, ctx->stream));
    }
  }
  queue_max_degree(ctx, L.n, L.off64, L.xadj, L.vwgt);  // vwgt: device copy or nullptr
  if (validate) check_graph(ctx, L.n, L.m, L.off64, L.xadj, L.adjncy, L.adjwgt);
  sync(ctx);
  L.maxdeg = ctx->h_small[4];
  L.maxvw = L.vwgt ? ctx->h_small[3] : 1;
  if (!L.adjwgt) L.adjwgt = const_cast<int32_t *>(unit_weights(ctx, L.m));
  if (!L.vwgt) L.vwgt = const_cast<int32_t *>(unit_weights(ctx, L.n));
  if (!g->on_device && g->vwgt) {
    int64_t s = 0;
    for (int64_t i = 0; i < L.n; ++i) s += g->vwgt[i];
    L.tvwgt = s;
  } else if (!g->vwgt) {
    L.tvwgt = L.n;
  } else {
    L.tvwgt = device_sum(ctx, L.n, L.vwgt);
  }
  return L;
}

int log2_floor(int a)
{
  int i;
  for (i = 1; a > 1; i++, a = a >> 1) {
  }
  return i - 1;
}

// cu:3634
int ref_threshold(int64_t n, int nparts)
{
  const int l = log2_floor(nparts);
  const int64_t a = l > 0 ? n / (20 * l) : n;
  const int64_t b = 30 * (int64_t)nparts;
  return (int)std::max(a, b);
}

// One coarsening level: match -> number -> contract (cu:1699-1806, 1291-1406).
// Levels with at most tune.speculate_edges entries run their matching rounds
// without a host check before the contraction (see coarsen_level).  Larger
// levels queue at least tune.large_level_batch launches first: a launch
// after the last live round returns at once (4-7 us on such a level), while
// a second batch costs a host synchronisation (about 40 us) first.
//
// The next level's matching queued ahead: while the host waits for a small
// level's read-back, the GPU would go idle between the contraction and the
// next level's first launch.  When the next level surely exists (every
// coarse vertex has at most two members, so nc >= n / 2 > T) and has no
// long rows (its rows are unions of two rows of at most maxdeg entries), its
// first batch of rounds is queued behind the contraction, sized by the fine
// level's bounds and reading nc from the device.  The batch is dropped when
// more work is queued behind it (the contraction redone) or when the level's
// own matching had to be resumed.
struct Ahead {
  MatchRun *mr = nullptr;
  int32_t *mstate = nullptr;  // the rounds' working match of the next level
  uint64_t seq = 0;           // ctx->launch_seq right after its launches
};

static void ahead_drop(cm_ctx *ctx, Ahead &a)
{
  if (a.mr) match_end(ctx, a.mr);
  if (a.mstate) ctx->pool.release(a.mstate);
  a = Ahead{};
}

static Level coarsen_level(cm_ctx *ctx, cm_hier *h, Level &F, int32_t maxvwgt, uint64_t seed,
                           int max_rounds, int row_order, int batch, Ahead *in, Ahead *out)
{
  const size_t ob = F.off64 ? 8 : 4;
  auto own = [&](size_t bytes) {
    void *p = ctx->pool.alloc(bytes);
    h->owned.push_back(p);
    return p;
  };
  const int64_t n = F.n, m = F.m;
  const bool adopt = in && in->mr;
  F.match = (int32_t *)own(4 * (n > 0 ? n : 1));
  F.cmap = (int32_t *)own(4 * (n > 0 ? n : 1));
  int32_t *rep = (int32_t *)ctx->pool.alloc(4 * (n > 0 ? n : 1));
  // the rounds' working match
  int32_t *mstate = adopt ? in->mstate : (int32_t *)ctx->pool.alloc(4 * (n > 0 ? n : 1));
  void *segoff = ctx->pool.alloc(ob * (n + 1));
  Level C;
  C.off64 = F.off64;
  C.vwgt = (int32_t *)own(4 * (n > 0 ? n : 1));
  C.xadj = own(ob * (n + 1));  // sized by the bound n: nc is read back at the level's end
  C.adjncy = (int32_t *)own(4 * (m > 0 ? m : 1));
  C.adjwgt = (int32_t *)own(4 * (m > 0 ? m : 1));
  MatchOwners mo{};
  mo.maxv0 = maxvwgt;
  mo.seed0 = seed;
  MatchRun *mr = nullptr;
  if (adopt) {
    mr = in->mr;
    *in = Ahead{};
    match_adopt(ctx, mr, F);
  } else {
    try {
      mr = match_begin(ctx, F, mo, max_rounds, mstate);
    } catch (...) {
      ctx->pool.release(rep);
      ctx->pool.release(mstate);
      ctx->pool.release(segoff);
      throw;
    }
  }
  try {
    // Small levels queue one batch of rounds and go straight on to the
    // numbering and contraction of the finished copy of the match: the
    // contraction's synchronisation also brings back the round flags.  In
    // the rare case that the batch did not finish, the rounds continue from
    // their state and the level is redone.  Large levels, where a
    // synchronisation costs little, check the rounds first.  A matching
    // queued ahead (only on small levels) has its first batch queued.
    const bool speculate = m <= ctx->tune.speculate_edges;
    if (n > 0 && !adopt) {
      // speculative: flags come with the contraction's read-back
      match_queue(ctx, mr, speculate ? batch : std::max(batch, (int)ctx->tune.large_level_batch), !speculate);
      if (!speculate) sync(ctx);
    }
    for (int pass = 0;; ++pass) {
      if (!speculate || pass > 0) match_finish(ctx, mr);
      // the numbering writes the completed copy of the rounds' match to
      // F.match; nc stays on the device until the contraction's one
      // synchronisation
      // first-seen levels with hashed segments: the numbering also lists
      // them for the hash contraction (one select pass fewer)
      struct Lists {
        cm_ctx *ctx;
        HashContract hc;
        ~Lists() { hash_release(ctx, hc); }  // after contract() this is a no-op
      } lists{ctx, {}};
      HashContract &hc = lists.hc;
      const bool pre = row_order == CM_ROW_FIRST_SEEN && n > 0 && hash_needed(F);
      if (pre) hash_alloc(ctx, F, n, hc);
      number(ctx, F, F.match, F.cmap, rep, segoff, C.vwgt, false, mstate, pre ? &hc : nullptr);
      if (out && pass == 0) {
        ctx->before_wait = [&]() {
          Level b;  // the next level: its arrays, with this level's sizes as bounds
          b.n = n;
          b.m = m;
          b.off64 = C.off64;
          b.xadj = C.xadj;
          b.adjncy = C.adjncy;
          b.adjwgt = C.adjwgt;
          b.vwgt = C.vwgt;
          b.maxdeg = 2 * F.maxdeg;
          b.maxvw = -1;  // every pair's weight is checked
          out->mstate = (int32_t *)ctx->pool.alloc(4 * n);
          out->mr = match_begin_ahead(ctx, b, mo, max_rounds, out->mstate, ctx->d_small, batch);
          out->seq = ctx->launch_seq;
        };
      }
      C.m = contract(ctx, F, F.match, F.cmap, -1, rep, segoff, row_order, C.xadj, C.adjncy, C.adjwgt,
                     &C.maxdeg, &C.n, pre ? &hc : nullptr);
      ctx->before_wait = nullptr;  // a contraction without read-back did not run it
      if (out && out->mr && ctx->launch_seq != out->seq) ahead_drop(ctx, *out);  // queued before a redo
      C.maxvw = ctx->h_small[3];
      if (ctx->rb_flags) {  // a contraction that read nothing back left the flags
        CM_HIP(hipMemcpyAsync(ctx->h_flags, ctx->rb_flags, sizeof(unsigned) * ctx->rb_nflags,
                              hipMemcpyDeviceToHost, ctx->stream));
        ctx->rb_flags = nullptr;
        sync(ctx);
      }
      if (match_done(ctx, mr)) break;
      if (out) ahead_drop(ctx, *out);  // it read the coarse graph of an unfinished matching
    }
    F.rounds = match_finish(ctx, mr);
  } catch (...) {
    ctx->before_wait = nullptr;
    ctx->rb_flags = nullptr;
    if (out) ahead_drop(ctx, *out);
    match_end(ctx, mr);
    ctx->pool.release(rep);
    ctx->pool.release(mstate);
    ctx->pool.release(segoff);
    throw;
  }
  match_end(ctx, mr);
  C.tvwgt = F.tvwgt;  // cu:1400
  C.owns_graph = true;
  ctx->pool.release(rep);
  ctx->pool.release(mstate);
  ctx->pool.release(segoff);
  return C;
}

}  // namespace cm

namespace cm {
SharedBlocks::~SharedBlocks()
{
  if (!ctx) return;
  (void)hipStreamSynchronize(ctx->stream);
  for (void *q : blocks) ctx->pool.release(q);
}
}  // namespace cm

namespace cm {
cm_tuning tuned_defaults()
{
  cm_tuning t{};
  t.first_round_batch = TUNED_FIRST_ROUND_BATCH;
  t.level_batch_margin = TUNED_LEVEL_BATCH_MARGIN;
  t.large_level_batch = TUNED_LARGE_LEVEL_BATCH;
  t.compact_round = TUNED_COMPACT_ROUND;
  t.compact_min_n = TUNED_COMPACT_MIN_N;
  t.speculate_edges = TUNED_SPECULATE_EDGES;
  t.number_one_tile_max = TUNED_NUMBER_ONE_TILE_MAX;
  t.seg_wave_bins = TUNED_SEG_WAVE_BINS;
  t.graph_rounds = TUNED_GRAPH_ROUNDS;
  return t;
}
}  // namespace cm

// ------------------------------------------------------------- cm_ctx
hipEvent_t cm_ctx::ev_get()
{
  if (!ev_free.empty()) {
    hipEvent_t e = ev_free.back();
    ev_free.pop_back();
    return e;
  }
  hipEvent_t e;
  CM_HIP(hipEventCreate(&e));
  return e;
}

void cm_ctx::collect_timing()
{
  for (auto &t : pending) {
    float ms = 0.f;
    CM_HIP(hipEventElapsedTime(&ms, t.a, t.b));
    auto &s = stats[t.name];
    s.count += t.count;
    s.total_ms += ms;
    s.algo_bytes += t.bytes;
    ev_free.push_back(t.a);
    ev_free.push_back(t.b);
  }
  pending.clear();
}

// ----------------------------------------------------------------- C ABI
#define CM_API_BEGIN try {
#define CM_API_END(ctx)                                     \
  }                                                         \
  catch (const cm::Error &e)                                \
  {                                                         \
    if (ctx) (ctx)->err = e.what();                         \
    return e.code;                                          \
  }                                                         \
  catch (const std::exception &e)                           \
  {                                                         \
    if (ctx) (ctx)->err = e.what();                         \
    return CM_ERR_INVALID;                                  \
  }                                                         \
  return CM_OK;

extern "C" {

int cm_ctx_create(int device, void *stream, cm_ctx **out)
{
  if (!out) return CM_ERR_INVALID;
  *out = nullptr;
  cm_ctx *ctx = new cm_ctx();
  try {
    int ndev = 0;
    CM_HIP(hipGetDeviceCount(&ndev));
    CM_REQUIRE(device >= 0 && device < ndev, CM_ERR_INVALID, "no such device");
    ctx->device = device;
    CM_HIP(hipSetDevice(device));
    if (stream) {
      ctx->stream = (hipStream_t)stream;
    } else {
      CM_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
      ctx->own_stream = true;
    }
    ctx->d_ctr = (unsigned *)ctx->pool.alloc(sizeof(unsigned) * cm::CTR_SLOTS);
    ctx->d_small = (int64_t *)ctx->pool.alloc(sizeof(int64_t) * cm::SMALL_SLOTS);
    CM_HIP(hipHostMalloc((void **)&ctx->h_small, sizeof(int64_t) * cm::SMALL_SLOTS, hipHostMallocDefault));
    CM_HIP(hipHostGetDevicePointer((void **)&ctx->h_small_dev, ctx->h_small, 0));
    cm::fill_u32(ctx, ctx->d_ctr, 0u, cm::CTR_SLOTS);
    ctx->noalloc_scratch = cm::noalloc_init(device);
    ctx->tune = cm::tuned_defaults();
    cm::sync(ctx);
  } catch (const cm::Error &e) {
    int code = e.code;
    delete ctx;
    return code;
  }
  *out = ctx;
  return CM_OK;
}

int cm_ctx_destroy(cm_ctx *ctx)
{
  if (!ctx) return CM_ERR_INVALID;
  for (cm_ctx *h : ctx->helpers) cm_ctx_destroy(h);
  for (cm_ctx *h : ctx->steps.all) cm_ctx_destroy(h);
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  for (auto &t : ctx->pending) {
    (void)hipEventDestroy(t.a);
    if (t.b) (void)hipEventDestroy(t.b);
  }
  for (auto e : ctx->ev_free) (void)hipEventDestroy(e);
  for (auto &kv : ctx->graphs) (void)hipGraphExecDestroy(kv.second);
  if (ctx->h_small) (void)hipHostFree(ctx->h_small);
  if (ctx->h_flags) (void)hipHostFree(ctx->h_flags);
  if (ctx->h_stage) (void)hipHostFree(ctx->h_stage);
  for (hipEvent_t e : ctx->stage_ev)
    if (e) (void)hipEventDestroy(e);
  if (ctx->own_stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return CM_OK;
}

const char *cm_ctx_last_error(const cm_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int cm_ctx_sync(cm_ctx *ctx)
{
  if (!ctx) return CM_ERR_INVALID;
  CM_API_BEGIN
  cm::sync(ctx);
  CM_API_END(ctx)
}

int cm_ctx_set_timing(cm_ctx *ctx, int on)
{
  if (!ctx) return CM_ERR_INVALID;
  CM_API_BEGIN
  CM_REQUIRE(on >= 0 && on <= 2, CM_ERR_INVALID, "timing mode is 0, 1 or 2");
  cm::sync(ctx);
  ctx->timing = on != 0;
  ctx->timing_runs = on == 2;
  CM_API_END(ctx)
}

int cm_ctx_reset_stats(cm_ctx *ctx)
{
  if (!ctx) return CM_ERR_INVALID;
  CM_API_BEGIN
  cm::sync(ctx);
  ctx->stats.clear();
  CM_API_END(ctx)
}

int cm_ctx_kernel_stat(cm_ctx *ctx, int i, char *name, int name_cap, int64_t *count, double *total_ms,
                       double *algo_bytes)
{
  if (!ctx) return CM_ERR_INVALID;
  CM_API_BEGIN
  cm::sync(ctx);
  CM_REQUIRE(i >= 0 && i < (int)ctx->stats.size(), CM_ERR_INVALID, "stat index out of range");
  auto it = ctx->stats.begin();
  std::advance(it, i);
  if (name && name_cap > 0) {
    strncpy(name, it->first.c_str(), name_cap - 1);
    name[name_cap - 1] = 0;
  }
  if (count) *count = it->second.count;
  if (total_ms) *total_ms = it->second.total_ms;
  if (algo_bytes) *algo_bytes = it->second.algo_bytes;
  CM_API_END(ctx)
}

int cm_ctx_get_tuning(const cm_ctx *ctx, cm_tuning *out)
{
  if (!ctx || !out) return CM_ERR_INVALID;
  *out = ctx->tune;
  return CM_OK;
}

int cm_ctx_set_tuning(cm_ctx *ctx, const cm_tuning *t)
{
  if (!ctx || !t) return CM_ERR_INVALID;
  CM_API_BEGIN
  CM_REQUIRE(t->first_round_batch >= 2 && t->first_round_batch <= 64, CM_ERR_INVALID, "first_round_batch in 2..64");
  CM_REQUIRE(t->level_batch_margin >= 1 && t->level_batch_margin <= 8, CM_ERR_INVALID, "level_batch_margin in 1..8");
  CM_REQUIRE(t->large_level_batch >= 2 && t->large_level_batch <= 64, CM_ERR_INVALID, "large_level_batch in 2..64");
  CM_REQUIRE(t->compact_round >= 2 && t->compact_round <= 32, CM_ERR_INVALID, "compact_round in 2..32");
  CM_REQUIRE(t->compact_min_n >= 1, CM_ERR_INVALID, "compact_min_n >= 1");
  CM_REQUIRE(t->speculate_edges >= 0, CM_ERR_INVALID, "speculate_edges >= 0");
  CM_REQUIRE(t->number_one_tile_max >= 0 && t->number_one_tile_max <= 4096, CM_ERR_INVALID,
             "number_one_tile_max in 0..4096");
  CM_REQUIRE(t->seg_wave_bins >= 1 && t->seg_wave_bins <= 6, CM_ERR_INVALID, "seg_wave_bins in 1..6");
  CM_REQUIRE(t->graph_rounds == 0 || t->graph_rounds == 1, CM_ERR_INVALID, "graph_rounds is 0 or 1");
  cm::sync(ctx);
  ctx->tune = *t;
  CM_API_END(ctx)
}

int cm_ctx_pool_bytes(cm_ctx *ctx, int64_t *reserved, int64_t *in_use)
{
  if (!ctx) return CM_ERR_INVALID;
  if (reserved) *reserved = (int64_t)ctx->pool.bytes_reserved();
  if (in_use) *in_use = (int64_t)ctx->pool.bytes_in_use();
  return CM_OK;
}

int cm_dmalloc(cm_ctx *ctx, size_t bytes, void **out)
{
  if (!ctx || !out) return CM_ERR_INVALID;
  CM_API_BEGIN
  CM_HIP(hipSetDevice(ctx->device));
  *out = ctx->pool.alloc(bytes);
  CM_API_END(ctx)
}

int cm_dfree(cm_ctx *ctx, void *p)
{
  if (!ctx) return CM_ERR_INVALID;
  CM_API_BEGIN
  cm::sync(ctx);
  ctx->pool.release(p);
  CM_API_END(ctx)
}

int cm_memcpy_h2d(cm_ctx *ctx, void *dst, const void *src, size_t bytes)
{
  if (!ctx) return CM_ERR_INVALID;
  CM_API_BEGIN
  if (bytes) CM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  cm::sync(ctx);
  CM_API_END(ctx)
}

int cm_memcpy_d2h(cm_ctx *ctx, void *dst, const void *src, size_t bytes)
{
  if (!ctx) return CM_ERR_INVALID;
  CM_API_BEGIN
  if (bytes) CM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  cm::sync(ctx);
  CM_API_END(ctx)
}

int cm_memcpy_d2d(cm_ctx *ctx, void *dst, const void *src, size_t bytes)
{
  if (!ctx) return CM_ERR_INVALID;
  CM_API_BEGIN
  if (bytes) CM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
  CM_API_END(ctx)
}

int cm_coarsen(cm_ctx *ctx, const cm_graph_view *g, const cm_coarsen_params *p, cm_hier **out)
{
  if (!ctx || !out || !p) return CM_ERR_INVALID;
  *out = nullptr;
  cm_hier *h = new cm_hier();
  try {
    CM_HIP(hipSetDevice(ctx->device));
    cm::validate_view(g);
    CM_REQUIRE(p->nparts >= 1, CM_ERR_INVALID, "nparts must be >= 1");
    CM_REQUIRE(p->row_order == CM_ROW_FIRST_SEEN || p->row_order == CM_ROW_ASCENDING, CM_ERR_INVALID,
               "row_order must be CM_ROW_FIRST_SEEN or CM_ROW_ASCENDING");
    h->levels.push_back(cm::level0(ctx, g, h, !(p->flags & CM_FLAG_TRUSTED_INPUT)));
    const int T = p->threshold > 0 ? p->threshold : cm::ref_threshold(g->n, p->nparts);
    const int32_t maxvwgt =
        p->maxvwgt > 0 ? p->maxvwgt : (int32_t)(1.5 * (double)h->levels[0].tvwgt / T);
    int batch = ctx->tune.first_round_batch;
    // levels go on from their read-back; the sync below ends the call
    struct Early {
      cm_ctx *ctx;
      cm::Ahead ahead, next;  // the next level's matching, queued ahead
      Early(cm_ctx *c) : ctx(c) { ctx->rb_early = true; }
      ~Early()
      {
        cm::ahead_drop(ctx, ahead);
        cm::ahead_drop(ctx, next);
        ctx->rb_early = false;
      }
    } early(ctx);
    for (;;) {
      const cm::Level &F = h->levels.back();
      // the level after the one built now is queued ahead when it surely
      // exists: nc >= n / 2 > T, within max_levels, rows of at most
      // 2 * maxdeg <= LONG_ROW entries, a small level (see coarsen_level)
      const bool ahead_ok = !ctx->timing && !ctx->tune.graph_rounds && F.n / 2 > T &&
                            F.m <= ctx->tune.speculate_edges && F.n < ctx->tune.compact_min_n &&
                            F.maxdeg >= 0 && 2 * F.maxdeg <= cm::LONG_ROW &&
                            (p->max_levels <= 0 || (int)h->levels.size() + 1 <= p->max_levels);
      cm::Level C = cm::coarsen_level(ctx, h, h->levels.back(), maxvwgt, p->seed, p->max_rounds,
                                      p->row_order, batch, &early.ahead, ahead_ok ? &early.next : nullptr);
      std::swap(early.ahead, early.next);  // early.next is empty: coarsen_level used up early.ahead
      // the next level's first batch: launches for the rounds this level
      // took plus one (launch r accepts round r-1, so r rounds need r + 1
      // launches to show that they finished); a short batch only costs a
      // resumption
      batch = std::min((int)ctx->tune.first_round_batch,
                       std::max(5, h->levels.back().rounds + (int)ctx->tune.level_batch_margin));
      const int64_t nf = h->levels.back().n;
      h->levels.push_back(C);
      const bool more = C.n > T && C.n < 0.75 * nf && C.m > C.n / 2;  // cu:1843-1845
      if (!more) break;  // (an ahead matching is dropped with `early`)
      if (p->max_levels > 0 && (int)h->levels.size() > p->max_levels) break;
    }
    cm::sync(ctx);
  } catch (const cm::Error &e) {
    ctx->err = e.what();
    for (void *q : h->owned) ctx->pool.release(q);
    delete h;
    return e.code;
  }
  *out = h;
  return CM_OK;
}

int cm_hier_num_levels(const cm_hier *h) { return h ? (int)h->levels.size() : CM_ERR_INVALID; }

int cm_hier_level(const cm_hier *h, int level, cm_level_info *out)
{
  if (!h || !out || level < 0 || level >= (int)h->levels.size()) return CM_ERR_INVALID;
  const cm::Level &L = h->levels[level];
  out->n = L.n;
  out->m = L.m;
  out->offset_type = L.off64 ? CM_OFF64 : CM_OFF32;
  out->rounds = L.rounds;
  out->xadj = L.xadj;
  out->adjncy = L.adjncy;
  out->adjwgt = L.adjwgt;
  out->vwgt = L.vwgt;
  out->match = L.match;
  out->cmap = L.cmap;
  out->tvwgt = L.tvwgt;
  out->vertex_base = L.vbase;
  out->edge_base = L.ebase;
  out->coarse_base = L.cbase;
  return CM_OK;
}

int cm_hier_copy_level(cm_ctx *ctx, const cm_hier *h, int level, void *xadj, int32_t *adjncy,
                       int32_t *adjwgt, int32_t *vwgt, int32_t *match, int32_t *cmap)
{
  if (!ctx || !h || level < 0 || level >= (int)h->levels.size()) return CM_ERR_INVALID;
  CM_API_BEGIN
  const cm::Level &L = h->levels[level];
  const size_t ob = L.off64 ? 8 : 4;
  auto cp = [&](void *dst, const void *src, size_t bytes) {
    if (dst && src && bytes) cm::download(ctx, dst, src, bytes);
  };
  cp(xadj, L.xadj, ob * (L.n + 1));
  cp(adjncy, L.adjncy, 4 * L.m);
  cp(adjwgt, L.adjwgt, 4 * L.m);
  cp(vwgt, L.vwgt, 4 * L.n);
  cp(match, L.match, 4 * L.n);
  cp(cmap, L.cmap, 4 * L.n);
  cm::sync(ctx);
  // local values for batch slices
  if (xadj && L.ebase) {
    if (L.off64)
      for (int64_t i = 0; i <= L.n; ++i) ((int64_t *)xadj)[i] -= L.ebase;
    else
      for (int64_t i = 0; i <= L.n; ++i) ((int32_t *)xadj)[i] -= (int32_t)L.ebase;
  }
  if (L.vbase) {
    if (adjncy)
      for (int64_t j = 0; j < L.m; ++j) adjncy[j] -= (int32_t)L.vbase;
    if (match && L.match)
      for (int64_t i = 0; i < L.n; ++i) match[i] -= (int32_t)L.vbase;
  }
  if (cmap && L.cmap && L.cbase)
    for (int64_t i = 0; i < L.n; ++i) cmap[i] -= (int32_t)L.cbase;
  CM_API_END(ctx)
}

int cm_project(cm_ctx *ctx, const cm_hier *h, int level, const int32_t *d_where_c, int32_t *d_where_f)
{
  if (!ctx || !h || level < 0 || level + 1 >= (int)h->levels.size()) return CM_ERR_INVALID;
  CM_API_BEGIN
  const cm::Level &L = h->levels[level];
  // stored cmap values are local + cbase: shift the coarse array instead
  cm::project(ctx, L.n, L.cmap, d_where_c - L.cbase, d_where_f);
  CM_API_END(ctx)
}

int cm_hier_release(cm_ctx *ctx, cm_hier *h)
{
  if (!ctx || !h) return CM_ERR_INVALID;
  CM_API_BEGIN
  cm::sync(ctx);
  for (void *q : h->owned) ctx->pool.release(q);
  delete h;
  CM_API_END(ctx)
}

int cm_match(cm_ctx *ctx, const cm_graph_view *g, int32_t maxvwgt, uint64_t seed, int32_t max_rounds,
             int32_t *d_match, int32_t *rounds)
{
  if (!ctx) return CM_ERR_INVALID;
  CM_API_BEGIN
  cm::validate_view(g);
  CM_REQUIRE(g->on_device, CM_ERR_INVALID, "cm_match needs a device view");
  cm_hier tmp;
  cm::Level L = cm::level0(ctx, g, &tmp);
  int r = 0;
  try {
    r = cm::match_handshake(ctx, L, maxvwgt, seed, max_rounds, d_match);
  } catch (...) {
    for (void *q : tmp.owned) ctx->pool.release(q);
    throw;
  }
  cm::sync(ctx);
  for (void *q : tmp.owned) ctx->pool.release(q);
  if (rounds) *rounds = r;
  CM_API_END(ctx)
}

int cm_number(cm_ctx *ctx, int64_t n, const void *d_xadj, int32_t offset_type, const int32_t *d_match,
              int32_t *d_cmap, int64_t *nc)
{
  if (!ctx || n < 0) return CM_ERR_INVALID;
  CM_API_BEGIN
  cm::Level L;
  L.n = n;
  L.off64 = offset_type == CM_OFF64;
  L.xadj = const_cast<void *>(d_xadj);
  const size_t ob = L.off64 ? 8 : 4;
  cm::check_match(ctx, n, d_match);
  int32_t *vw = (int32_t *)ctx->pool.alloc(4 * (n > 0 ? n : 1));
  cm::fill_u32(ctx, vw, 1u, n);
  L.vwgt = vw;
  int32_t *rep = (int32_t *)ctx->pool.alloc(4 * (n > 0 ? n : 1));
  int32_t *cv = (int32_t *)ctx->pool.alloc(4 * (n > 0 ? n : 1));
  void *seg = ctx->pool.alloc(ob * (n + 1));
  int64_t r = cm::number(ctx, L, const_cast<int32_t *>(d_match), d_cmap, rep, seg, cv);  // complete: not written
  cm::sync(ctx);
  ctx->pool.release(vw);
  ctx->pool.release(rep);
  ctx->pool.release(cv);
  ctx->pool.release(seg);
  if (nc) *nc = r;
  CM_API_END(ctx)
}

int cm_contract(cm_ctx *ctx, const cm_graph_view *g, const int32_t *d_match, const int32_t *d_cmap,
                int64_t nc, int32_t row_order, void *d_cxadj, int32_t *d_cadjncy, int32_t *d_cadjwgt,
                int32_t *d_cvwgt, int64_t *mc)
{
  if (!ctx) return CM_ERR_INVALID;
  CM_API_BEGIN
  cm::validate_view(g);
  CM_REQUIRE(g->on_device, CM_ERR_INVALID, "cm_contract needs a device view");
  cm_hier tmp;
  cm::Level L = cm::level0(ctx, g, &tmp);
  const size_t ob = L.off64 ? 8 : 4;
  const int64_t n = L.n;
  int32_t *cm2 = (int32_t *)ctx->pool.alloc(4 * (n > 0 ? n : 1));
  int32_t *rep = (int32_t *)ctx->pool.alloc(4 * (n > 0 ? n : 1));
  void *seg = ctx->pool.alloc(ob * (n + 1));
  int64_t out = 0;
  try {
    // numbering recomputed to obtain the segment offsets; it must agree
    // with the caller's cmap, which is what the gather uses
    cm::check_match(ctx, n, d_match);
    const int64_t nc2 = cm::number(ctx, L, const_cast<int32_t *>(d_match), cm2, rep, seg, d_cvwgt);  // complete
    CM_REQUIRE(nc2 == nc, CM_ERR_INVALID, "nc does not match the matching");
    cm::check_equal(ctx, n, cm2, d_cmap, "cmap must be the canonical numbering of match");
    out = cm::contract(ctx, L, d_match, d_cmap, nc, rep, seg, row_order, d_cxadj, d_cadjncy, d_cadjwgt);
  } catch (...) {
    ctx->pool.release(cm2);
    ctx->pool.release(rep);
    ctx->pool.release(seg);
    for (void *q : tmp.owned) ctx->pool.release(q);
    throw;
  }
  cm::sync(ctx);
  ctx->pool.release(cm2);
  ctx->pool.release(rep);
  ctx->pool.release(seg);
  for (void *q : tmp.owned) ctx->pool.release(q);
  if (mc) *mc = out;
  CM_API_END(ctx)
}

int cm_segsort_i32(cm_ctx *ctx, int32_t *d_keys, int32_t *d_vals, int32_t n, const int32_t *d_segs,
                   int32_t length)
{
  if (!ctx || n < 0 || length < 0) return CM_ERR_INVALID;
  CM_API_BEGIN
  cm::segsort_i32(ctx, d_keys, d_vals, n, d_segs, length);
  CM_API_END(ctx)
}

int cm_project_raw(cm_ctx *ctx, int64_t n, const int32_t *d_cmap, const int32_t *d_where_c,
                   int32_t *d_where_f)
{
  if (!ctx || n < 0) return CM_ERR_INVALID;
  CM_API_BEGIN
  cm::project(ctx, n, d_cmap, d_where_c, d_where_f);
  cm::sync(ctx);
  CM_API_END(ctx)
}

}  // extern "C"
