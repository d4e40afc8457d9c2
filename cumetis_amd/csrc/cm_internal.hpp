// cm_internal.hpp -- host-side internals of the coarsening engine: device
// memory pool, per-context state, kernel launch with optional hipEvent timing,
// and the level record of the hierarchy.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <condition_variable>
#include <functional>
#include <map>
#include <memory>
#include <unordered_map>
#include <mutex>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/cumetis_amd.h"

namespace cm {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define CM_HIP(expr)                                                                    \
  do {                                                                                  \
    hipError_t _e = (expr);                                                             \
    if (_e != hipSuccess)                                                               \
      throw ::cm::Error(CM_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
  } while (0)

#define CM_REQUIRE(cond, code, msg)                      \
  do {                                                   \
    if (!(cond)) throw ::cm::Error((code), (msg));       \
  } while (0)

// Caching device allocator: power-of-two size classes, never returns memory
// to the runtime before the context dies, so steady-state coarsening calls
// allocate nothing.  A size class hands out its lowest free address, so a
// call repeated from the same free state gets the same blocks (and its
// kernels the same arguments, which lets replayed launch graphs match).
class Pool {
 public:
  ~Pool();
  void *alloc(size_t bytes);
  void release(void *p);
  size_t bytes_reserved() const { return reserved_; }
  size_t bytes_in_use() const { return in_use_; }
  void trim();

 private:
  std::map<size_t, std::set<void *>> free_;
  std::map<void *, size_t> live_;
  size_t reserved_ = 0, in_use_ = 0;
  std::recursive_mutex mu_;  // hierarchies may be released from another host thread
};

struct KStat {
  int64_t count = 0;
  double total_ms = 0.0;
  double algo_bytes = 0.0;
};

struct Timed {
  const char *name;
  hipEvent_t a, b;
  double bytes;
  int64_t count;  // launches between a and b (run timing: consecutive launches of one name)
  uint64_t seq;   // launch_seq after the last of them
};

// Rows above LONG_ROW entries are scanned by whole waves in the matching.
constexpr int LONG_ROW = 64;
constexpr int STEP_CTX_MAX = 4;  // contexts lent to device-coarsened initial-partition steps

struct Level {
  int64_t n = 0, m = 0;
  bool off64 = false;
  void *xadj = nullptr;
  int32_t *adjncy = nullptr, *adjwgt = nullptr, *vwgt = nullptr;
  int32_t *cmap = nullptr, *match = nullptr;
  int64_t tvwgt = 0;
  bool owns_graph = false;  // xadj/adjncy/adjwgt/vwgt come from the pool
  int rounds = 0;           // handshake rounds used at this level
  // batch slices: stored ids are local id + base (0 for a single graph)
  int64_t vbase = 0;  // added to adjncy and match values
  int64_t ebase = 0;  // added to xadj values
  int64_t cbase = 0;  // added to cmap values
  int64_t maxdeg = -1;  // bound on the longest row (-1 unknown)
  int64_t maxvw = -1;   // largest vertex weight (-1 unknown)
};

}  // namespace cm

namespace cm {
// Device blocks shared by the hierarchies of one batch; released when the
// last of them is released.
struct SharedBlocks {
  cm_ctx *ctx = nullptr;
  std::vector<void *> blocks;
  ~SharedBlocks();
};
}  // namespace cm

struct cm_hier {
  std::vector<cm::Level> levels;
  std::vector<void *> owned;  // pool blocks released with the hierarchy
  std::shared_ptr<cm::SharedBlocks> shared;
  std::vector<std::shared_ptr<cm::SharedBlocks>> shared_more;  // serving: one per union step
};

struct cm_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  std::vector<cm_ctx *> helpers;  // cm_partition's concurrent trials (created on first use)
  // the parallel initial partition's device-coarsened recursion steps: at
  // most cm::STEP_CTX_MAX contexts of their own, created on first use, each lent
  // to one step at a time
  struct {
    std::mutex mu;
    std::condition_variable cv;
    std::vector<cm_ctx *> all, idle;
  } steps;
  cm::Pool pool;
  std::string err;
  bool timing = false;
  // timing mode 2: consecutive launches of one name share one pair of events
  // (no event between them), so a run of short launches is timed without
  // the gap an event adds to every launch
  bool timing_runs = false;
  std::vector<cm::Timed> pending;  // run timing: the last one may be open (b == nullptr)
  std::vector<hipEvent_t> ev_free;
  std::map<std::string, cm::KStat> stats;
  // small persistent device buffers
  unsigned *d_ctr = nullptr;      // tile tickets / counters (CTR_SLOTS)
  int64_t *d_small = nullptr;     // totals written by kernels (SMALL_SLOTS)
  int64_t *h_small = nullptr;     // pinned mirror of d_small
  unsigned *h_flags = nullptr;    // pinned copy of matching round flags
  size_t h_flags_cap = 0;         // in words
  char *h_stage = nullptr;        // two pinned chunks for large staged copies (created on first use)
  hipEvent_t stage_ev[2] = {nullptr, nullptr};
  int64_t *h_small_dev = nullptr;   // device addresses of h_small and h_flags
  unsigned *h_flags_dev = nullptr;  // (written by a level's last kernel)
  // round flags for the next read-back (readback_begin): the matching's
  // flags are then copied with the level's totals instead of on their own
  const unsigned *rb_flags = nullptr;
  int64_t rb_nflags = 0;
  int64_t rb_tag = 0;
  // rb_early: readback_end returns as soon as the read-back has arrived in
  // host memory, while the stream may still run (the caller queues only
  // stream-ordered work after it and synchronises before it returns)
  bool rb_early = false;
  // run once by the next readback_end before it waits (the coarsening
  // queues the next level's first rounds there, behind the contraction)
  std::function<void()> before_wait;
  uint64_t launch_seq = 0;  // kernels launched so far
  // unit weights shared by every graph given without weights (never
  // written; earlier, shorter blocks stay valid until the context dies)
  int32_t *ones = nullptr;
  int64_t ones_cap = 0;
  uint64_t *d_status = nullptr;   // look-back status words
  size_t status_cap = 0;          // in words
  uint32_t epoch = 0;             // look-back launch epoch (1 .. 2^20-1)
  int64_t last_rounds = 0;
  // the hash-contraction item counters (d_ctr + HASH_CTR_SLOT) are zero at
  // this point of the stream: the matching's init kernel clears them
  bool hash_ctr_zero = false;
  // device address of this device's static scratch array (cm_noalloc.hip)
  void *noalloc_scratch = nullptr;
  // launch-shape parameters (cm_tuning; defaults from cm_tuned.hpp)
  cm_tuning tune{};
  // replayable round batches, keyed by every parameter of their launches
  std::unordered_map<std::string, hipGraphExec_t> graphs;

  hipEvent_t ev_get();
  void collect_timing();
};

namespace cm {

constexpr int CTR_SLOTS = 64;
constexpr int HASH_CTR_SLOT = 32;  // 16 counters of the hash contraction
constexpr int SMALL_SLOTS = 4096;

// Device address of the allocation-free entry point's static scratch array
// on `device` (which must be current); looked up once per device.
void *noalloc_init(int device);
// The launch-shape defaults of a new context (cm_tuned.hpp).
cm_tuning tuned_defaults();
// Host-side checks of a graph view (sizes, offset type, NULL arrays).
void validate_view(const cm_graph_view *g);
void ensure_status(cm_ctx *ctx, size_t words);
void ensure_host_flags(cm_ctx *ctx, size_t words);
// Status words for a look-back launch over `words` tiles; returns the epoch.
uint32_t lookback_begin(cm_ctx *ctx, size_t words);

// Run timing: records the end event of the open run, if any.
void close_timing_run(cm_ctx *ctx);

template <typename K, typename... Args>
inline void launch(cm_ctx *ctx, const char *name, double bytes, K kernel, dim3 grid, dim3 block,
                   size_t shmem, Args... args)
{
  if (grid.x == 0 || grid.y == 0 || grid.z == 0) return;
  hipEvent_t a = nullptr, b = nullptr;
  if (ctx->timing && ctx->timing_runs) {
    // run timing: a launch that follows a launch of the same name joins its
    // run (no event between them); the run ends at the next launch of
    // another name or at the next synchronisation (close_timing_run)
    Timed *open = !ctx->pending.empty() && ctx->pending.back().b == nullptr ? &ctx->pending.back() : nullptr;
    const bool extend = open && open->seq == ctx->launch_seq && strcmp(open->name, name) == 0;
    if (!extend) {
      close_timing_run(ctx);
      a = ctx->ev_get();
      CM_HIP(hipEventRecord(a, ctx->stream));
    }
    hipLaunchKernelGGL(kernel, grid, block, shmem, ctx->stream, args...);
    CM_HIP(hipGetLastError());
    ++ctx->launch_seq;
    if (extend) {
      open->bytes += bytes;
      open->count += 1;
      open->seq = ctx->launch_seq;
    } else {
      ctx->pending.push_back(Timed{name, a, nullptr, bytes, 1, ctx->launch_seq});
    }
    return;
  }
  if (ctx->timing) {
    a = ctx->ev_get();
    CM_HIP(hipEventRecord(a, ctx->stream));
  }
  hipLaunchKernelGGL(kernel, grid, block, shmem, ctx->stream, args...);
  CM_HIP(hipGetLastError());
  ++ctx->launch_seq;
  if (ctx->timing) {
    b = ctx->ev_get();
    CM_HIP(hipEventRecord(b, ctx->stream));
    ctx->pending.push_back(Timed{name, a, b, bytes, 1, ctx->launch_seq});
  }
}

inline unsigned nblocks(int64_t work, int per_block)
{
  int64_t b = (work + per_block - 1) / per_block;
  return (unsigned)(b < 1 ? 1 : b);
}

void sync(cm_ctx *ctx);
// Large copies through the context's two pinned chunks (host threads copy
// one chunk in or out while the other's DMA runs).  download returns when
// the data is on the host; upload returns with the copy queued on the stream
// and the host source no longer needed.
void download(cm_ctx *ctx, void *dst, const void *src, size_t bytes);
void upload(cm_ctx *ctx, void *dst, const void *src, size_t bytes);
// The read-back of a contraction (see Readback): d_small[first, first +
// count) to h_small, *ovf (if given) to h_small[6], and the registered round
// flags (ctx->rb_flags, consumed) to h_flags.  begin() describes it for the
// kernel that sends it; end() waits for it (for the stream, unless
// ctx->rb_early) and checks that it arrived.
struct Readback;  // cm_common.hpp
Readback readback_begin(cm_ctx *ctx, int first, int count, const unsigned *ovf = nullptr);
void readback_end(cm_ctx *ctx, const Readback &rb);
void fill_u32(cm_ctx *ctx, void *p, uint32_t v, int64_t count);
// count unit weights (read-only, valid for the context's lifetime): graphs
// without vertex or edge weights share them instead of filling their own.
const int32_t *unit_weights(cm_ctx *ctx, int64_t count);
void fill_u64(cm_ctx *ctx, void *p, uint64_t v, int64_t count);
void scan_excl_i32(cm_ctx *ctx, int64_t n, const int32_t *in, int32_t *out);
// The same scan over caller-provided look-back status words (ceil(n/1024) of
// them, none holding a word of `epoch` from an earlier launch).
void scan_excl_i32_status(cm_ctx *ctx, int64_t n, const int32_t *in, int32_t *out, uint64_t *status,
                          uint32_t epoch);

// Owning graph of every vertex of a batch (disjoint union); gid == nullptr
// means one graph with maxv0 / seed0.
struct MatchOwners {
  const int32_t *gid = nullptr;    // graph slot of each vertex
  const int32_t *maxv = nullptr;   // per slot
  const uint64_t *seed = nullptr;  // per slot
  const int64_t *vbase = nullptr;  // per slot: first vertex in the union
  int32_t maxv0 = 0;
  uint64_t seed0 = 0;
  bool wunit = false;  // every edge weight is 1 (the level uses the shared unit array)
};

// ---- subsystem entry points (device pointers, ctx stream) ----
// (a) matching; returns rounds.  match[] ends complete (partner or self).
// A matching in progress (see cm_match.hip): begin; queue batches of round
// launches (each also queues a copy of the round flags to ctx->h_flags);
// after a synchronisation, done() tells whether the queued launches finished
// it; finish() runs further batches until done and returns the rounds;
// copy_finished() writes the result with unmatched vertices matched to
// themselves; end() releases the state.
struct MatchRun;
MatchRun *match_begin(cm_ctx *ctx, const Level &L, const MatchOwners &own, int max_rounds, int32_t *d_match);
// A matching queued before its level's sizes are on the host: `bound` has
// the level's arrays and bounds on its sizes (no rows above LONG_ROW); the
// first `count` launches read the vertex count from d_n.  adopt() gives it
// the real sizes once they are known (later launches use them) and leaves
// the flags of the queued launches to the next read-back.
MatchRun *match_begin_ahead(cm_ctx *ctx, const Level &bound, const MatchOwners &own, int max_rounds,
                            int32_t *d_match, const int64_t *d_n, int count);
void match_adopt(cm_ctx *ctx, MatchRun *mr, const Level &L);
// copy_flags = false leaves the flags to the next read-back (readback_begin).
void match_queue(cm_ctx *ctx, MatchRun *mr, int count, bool copy_flags = true);
bool match_done(cm_ctx *ctx, MatchRun *mr);
int match_finish(cm_ctx *ctx, MatchRun *mr);
void match_copy_finished(cm_ctx *ctx, MatchRun *mr, int32_t *out);
void match_end(cm_ctx *ctx, MatchRun *mr);
constexpr int MATCH_BATCH = 9;  // launches per batch of a resumed matching
// finish=false leaves unmatched vertices at -1 for number() to complete.
int match_handshake(cm_ctx *ctx, const Level &L, int32_t maxvwgt, uint64_t seed, int max_rounds,
                    int32_t *d_match, bool finish = true);
int match_handshake(cm_ctx *ctx, const Level &L, const MatchOwners &own, int max_rounds, int32_t *d_match,
                    bool finish = true);
// (b) numbering: cmap for every fine vertex; rep/segoff/cvwgt per coarse vertex.
//     Returns nc; segoff[nc] = m.
// With wait=false nc stays in d_small[0] (read it back with the next
// synchronisation) and -1 is returned.
// match entries still -1 (matching run with finish=false) are completed as
// self matches.  With d_match_in, the matching is read from there and its
// completed copy written to d_match.
// With hc (from hash_alloc), the numbering also fills hc's item lists for
// the hash contraction (hand hc to contract() then).
struct HashContract;
int64_t number(cm_ctx *ctx, const Level &L, int32_t *d_match, int32_t *d_cmap,
               int32_t *d_rep, void *d_segoff, int32_t *d_cvwgt, bool wait = true,
               const int32_t *d_match_in = nullptr, HashContract *hc = nullptr);
void queue_max_degree(cm_ctx *ctx, int64_t n, bool off64, const void *xadj, const int32_t *vwgt = nullptr);
// (c) contraction of L into coarse CSR; returns mc.  cxadj/cadjncy/cadjwgt
//     must hold nc+1 / L.m / L.m entries.  nc = -1: the count left on the
//     device by number(wait=false); cxadj then holds L.n+1 entries and *ncp
//     receives nc.  After any number() + contract() pair, h_small[3] holds
//     the largest coarse vertex weight.
int64_t contract(cm_ctx *ctx, const Level &L, const int32_t *d_match, const int32_t *d_cmap,
                 int64_t nc, const int32_t *d_rep, const void *d_segoff, int row_order,
                 void *d_cxadj, int32_t *d_cadjncy, int32_t *d_cadjwgt, int64_t *maxdeg = nullptr,
                 int64_t *ncp = nullptr, HashContract *pre = nullptr);
// segmented sort of 64-bit keys with 32-bit values, segments segoff[0..nseg]
void segsort_u64(cm_ctx *ctx, int64_t nseg, const void *d_segoff, bool off64, int64_t total,
                 uint64_t *d_keys, uint32_t *d_vals);
// multilevel partition of an existing hierarchy: initial partition of the
// coarsest level (host), then per level projection + one refinement pass.
// part receives level-0 assignments (host), *cut (when given) their edge cut;
// flags: CM_FLAG_NO_BALANCE, CM_FLAG_EXACT_INITPART, CM_FLAG_HOST_INIT_COARSEN;
// rep receives init_ms, uncoarsen_ms and init_device_steps.
void partition_hierarchy(cm_ctx *ctx, const cm_hier *h, int nparts, int32_t flags, int32_t *part,
                         cm_part_report *rep, int64_t *cut = nullptr);
// (d) projection
void project(cm_ctx *ctx, int64_t n, const int32_t *d_cmap, const int32_t *d_where_c,
             int32_t *d_where_f);

}  // namespace cm
