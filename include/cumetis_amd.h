/*
 * cumetis_amd.h -- C ABI of the MI355X (gfx950) multilevel coarsening engine.
 *
 * Plain pointers and sizes only.  Citations ("cu:") are to cuMetis.cu of the
 * reference, which exposes one public function (include/cuMetis.h:18-19) and
 * otherwise an internal operator boundary around its GPU coarsening.
 *
 * Conventions
 *   - every function returns CM_OK (0) or a negative cm_status; none exits;
 *   - cm_ctx_last_error() gives the message of the last failure on a context;
 *   - device pointers are stream-ordered on the context's stream;
 *   - the caller owns its inputs; the library owns a hierarchy until
 *     cm_hier_release();
 *   - contexts are independent; one context must not be used by two host
 *     threads at once.
 */
#ifndef CUMETIS_AMD_H
#define CUMETIS_AMD_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum cm_status {
  CM_OK = 0,
  CM_ERR_INVALID = -1,     /* bad argument */
  CM_ERR_HIP = -2,         /* HIP runtime error */
  CM_ERR_NOMEM = -3,       /* device allocation failed */
  CM_ERR_UNSUPPORTED = -4, /* size or combination not supported */
  CM_ERR_INPUT = -5        /* malformed graph input */
};

enum cm_offset_type { CM_OFF32 = 0, CM_OFF64 = 1 };

/* Row order of coarse adjacency lists.
 *   CM_ROW_FIRST_SEEN: byte-identical to the reference CPU contraction
 *     (cu:737-825): distinct coarse neighbours in first-seen order over
 *     row(v) then row(u), v = smaller endpoint, self entry replaced by the
 *     last one.
 *   CM_ROW_ASCENDING: ascending coarse neighbour ids, the order the
 *     reference GPU contraction produces (cu:1332 then cu:1196-1271). */
enum cm_row_order { CM_ROW_FIRST_SEEN = 0, CM_ROW_ASCENDING = 1 };

typedef struct cm_ctx cm_ctx;
typedef struct cm_hier cm_hier;

/* A CSR graph.  xadj has n+1 entries of the offset type; adjwgt/vwgt may be
 * NULL for unit weights (cu:479-503); with adjwgt NULL the first level's
 * matching and contraction read no weight array, so pass NULL rather than an
 * array of ones.  on_device selects device or host
 * pointers.  The graph must be symmetric, without self loops, with positive
 * edge weights. */
typedef struct {
  int64_t n;
  int64_t m;
  int32_t offset_type;
  int32_t on_device;
  const void *xadj;
  const int32_t *adjncy;
  const int32_t *adjwgt;
  const int32_t *vwgt;
} cm_graph_view;

/* Coarsening parameters.  threshold 0 = max(n/(20*floor(log2 k)), 30k)
 * (cu:3634); maxvwgt 0 = (int)(1.5*tvwgt/threshold) (cu:1828). */
typedef struct {
  int32_t nparts;
  int32_t threshold;
  int32_t maxvwgt;
  int32_t row_order;  /* cm_row_order */
  int32_t max_levels; /* 0 = no limit */
  int32_t max_rounds; /* handshake round cap, 0 = default (4096) */
  uint64_t seed;
  int32_t flags;      /* CM_FLAG_* */
  int32_t trials;     /* cm_partition / cuMetis_PartGraph / cm_server: partition
                         trials, 0 = CM_DEFAULT_TRIALS; trial t coarsens with
                         seed + t, the lowest cut is kept (the earlier trial on
                         ties); cm_coarsen ignores it */
} cm_coarsen_params;

#define CM_DEFAULT_TRIALS 4

/* cm_coarsen_params.flags */
#define CM_FLAG_TRUSTED_INPUT 1 /* skip the device-side validation of the input graph */
/* cm_partition / cm_server: skip the balancing pass that moves vertices out
 * of parts heavier than int(w/k * 1.03) before each level's refinement; with
 * it set the pipeline is the reference's (cu:3544-3616), which leaves
 * overweight parts of the initial partition overweight. */
#define CM_FLAG_NO_BALANCE 2
/* cm_partition / cm_server: run the host initial partition in the
 * reference's serial form (one glibc rand() stream, cu:287-300) also when the
 * coarsest graph has 2^14 vertices or more.  Without it such graphs (R-MAT,
 * whose coarsening stalls on isolated vertices and stars; large meshes in many
 * parts) take a parallel form of the
 * same recursive bisection: isolated vertices placed last on the lightest
 * parts, one stream per recursion step seeded from its place, the two halves
 * of each split on separate threads; deterministic, but not the reference's
 * rand order. */
#define CM_FLAG_EXACT_INITPART 4
/* cm_partition / cm_server: the parallel form coarsens every recursion step
 * on the host.  Without it, steps whose graph has 2^20 entries or more
 * coarsen on the device (cm_coarsen with the step's threshold 200, maxvwgt
 * int(1.5*w/200) and stop rule, seeded from the step), the rest of the step
 * (region growing, 2-way FM) staying on the host: a different but equally
 * deterministic chain, and seconds less on graphs whose coarsest level keeps
 * a large dense core (R-MAT). */
#define CM_FLAG_HOST_INIT_COARSEN 8

/* One level of a hierarchy (device pointers owned by the hierarchy; they are
 * read-only for the caller: level 0 of a graph given without weights points
 * at unit weights the context shares between graphs).
 * A level of a batched hierarchy is a slice of the batch's level arrays:
 * its stored values are local values plus a base, i.e. xadj holds
 * local + edge_base, adjncy and match hold local + vertex_base, cmap holds
 * local + coarse_base (all bases are 0 for cm_coarsen).  cm_hier_copy_level
 * returns local values. */
typedef struct {
  int64_t n;
  int64_t m;
  int32_t offset_type;
  int32_t rounds; /* matching rounds spent on this level (0 on the coarsest;
                     for a batched hierarchy, the rounds of the union step,
                     which ends when its slowest graph has finished) */
  void *xadj;
  int32_t *adjncy;
  int32_t *adjwgt;
  int32_t *vwgt;
  int32_t *match; /* NULL on the coarsest level */
  int32_t *cmap;  /* NULL on the coarsest level */
  int64_t tvwgt;
  int64_t vertex_base;
  int64_t edge_base;
  int64_t coarse_base;
} cm_level_info;

/* ---------------------------------------------------------------- context */
/* stream: a hipStream_t, or NULL to create a private non-blocking stream. */
int cm_ctx_create(int device, void *stream, cm_ctx **out);
int cm_ctx_destroy(cm_ctx *ctx);
const char *cm_ctx_last_error(const cm_ctx *ctx);
int cm_ctx_sync(cm_ctx *ctx);
/* Per-kernel hipEvent timing on the context stream (off by default):
 * on = 1 records events around every launch; on = 2 around every run of
 * consecutive launches of one kernel (the events then add no gap between
 * them; a run's time includes the gaps between its launches). */
int cm_ctx_set_timing(cm_ctx *ctx, int on);
int cm_ctx_reset_stats(cm_ctx *ctx);
/* Stats of kernel i (0-based) in name order: returns CM_ERR_INVALID past the
 * end.  name receives up to name_cap bytes. */
int cm_ctx_kernel_stat(cm_ctx *ctx, int i, char *name, int name_cap, int64_t *count,
                       double *total_ms, double *algo_bytes);
int cm_ctx_pool_bytes(cm_ctx *ctx, int64_t *reserved, int64_t *in_use);

/* Launch-shape parameters of a context.  None of them changes a result:
 * they decide how many round launches are queued before a host check, when
 * the matching switches to its frontier lists, which levels are numbered by
 * one tile, which segment lengths the segmented sort keeps in registers,
 * and whether round batches run as replayed hipGraphs.  A new context takes
 * the values tuned for gfx950 (cumetis_amd/tune.py writes them into the
 * build); set them to run the tuner's candidates. */
typedef struct {
  int32_t first_round_batch;   /* round launches queued at level 0 (2..64) */
  int32_t level_batch_margin;  /* later levels: previous rounds + this (1..8) */
  int32_t large_level_batch;   /* at least this many on levels above speculate_edges (2..64) */
  int32_t compact_round;       /* frontier lists after this launch (2..32) */
  int64_t compact_min_n;       /* ... on levels with at least this many vertices */
  int64_t speculate_edges;     /* levels up to this many entries contract before their round check */
  int32_t number_one_tile_max; /* levels up to this many vertices (0..4096) numbered by one tile */
  int32_t seg_wave_bins;       /* segment-length bins 0..this-1 (2..2^this keys) sorted in registers (1..6) */
  int32_t graph_rounds;        /* 1: round batches replay cached hipGraphs */
  int32_t reserved[7];
} cm_tuning;
int cm_ctx_get_tuning(const cm_ctx *ctx, cm_tuning *out);
/* CM_ERR_INVALID if a value is out of range (the context keeps its values). */
int cm_ctx_set_tuning(cm_ctx *ctx, const cm_tuning *t);

/* Device memory helpers for callers without an allocator of their own. */
int cm_dmalloc(cm_ctx *ctx, size_t bytes, void **out);
int cm_dfree(cm_ctx *ctx, void *p);
int cm_memcpy_h2d(cm_ctx *ctx, void *dst, const void *src, size_t bytes);
int cm_memcpy_d2h(cm_ctx *ctx, void *dst, const void *src, size_t bytes);
/* Device-to-device copy queued on the context stream (returns without
 * waiting); with cm_ctx_sync it measures the on-box copy bandwidth. */
int cm_memcpy_d2d(cm_ctx *ctx, void *dst, const void *src, size_t bytes);

/* ------------------------------------------------------------- hierarchy */
/* Multilevel coarsening: replaces cuMetis_coarsen (cu:1824-1850).  Levels are
 * built until the reference stop rule fails (cu:1843-1845): continue while
 * nc > threshold && nc < 0.75 n && mc > nc/2. */
int cm_coarsen(cm_ctx *ctx, const cm_graph_view *g, const cm_coarsen_params *p, cm_hier **out);

/* Batched coarsening of nb graphs concatenated into one CSR: concat->xadj
 * holds the offsets of the concatenation (sum n + 1 entries, concat->m =
 * sum m), adjncy holds vertex ids local to each graph; graph b owns
 * vertices [voff[b], voff[b+1]) and entries [eoff[b], eoff[b+1]); voff and
 * eoff are host arrays of nb+1 entries; params[b] applies to graph b, except
 * that row_order, max_rounds and max_levels must agree across the batch
 * (CM_ERR_INVALID otherwise).  Unless every params[b].flags holds
 * CM_FLAG_TRUSTED_INPUT, each graph is checked on the device before use
 * (rows inside its entries, local ids in [0, n_b), no self loop, positive
 * weights; CM_ERR_INPUT otherwise).  Every level is one pass of each kernel
 * over the union of the graphs still coarsening; graph b stops by its own
 * rule.  out[b] receives a hierarchy whose levels (after subtracting the
 * bases, see cm_level_info) are bit-identical to cm_coarsen of graph b
 * alone. */
int cm_coarsen_batch(cm_ctx *ctx, int32_t nb, const cm_graph_view *concat, const int64_t *voff,
                     const int64_t *eoff, const cm_coarsen_params *params, cm_hier **out);

int cm_hier_num_levels(const cm_hier *h);
int cm_hier_level(const cm_hier *h, int level, cm_level_info *out);
/* Copy one level to host arrays (any of them may be NULL to skip).  xadj
 * uses the hierarchy's offset type. */
int cm_hier_copy_level(cm_ctx *ctx, const cm_hier *h, int level, void *xadj, int32_t *adjncy,
                       int32_t *adjwgt, int32_t *vwgt, int32_t *match, int32_t *cmap);
/* Projection where_f[i] = where_c[cmap[i]] from level+1 to level:
 * replaces cuMetis_kway_project / projectback (cu:3067-3087). */
int cm_project(cm_ctx *ctx, const cm_hier *h, int level, const int32_t *d_where_c,
               int32_t *d_where_f);
int cm_hier_release(cm_ctx *ctx, cm_hier *h);

/* -------------------------------------------------------- unit operators */
/* Seeded handshake heavy-edge matching of one graph (device view):
 * replaces cuMetis_gpu_match's matching steps (cu:1699-1747; kernels
 * cu:1411-1419, 1563-1630).  d_match[i] = partner or i.  rounds may be NULL. */
int cm_match(cm_ctx *ctx, const cm_graph_view *g, int32_t maxvwgt, uint64_t seed,
             int32_t max_rounds, int32_t *d_match, int32_t *rounds);
/* Coarse numbering: replaces findc2..findc4 + scan (cu:1634-1694,
 * cu:1750-1789).  d_match must be complete (partner or self). */
int cm_number(cm_ctx *ctx, int64_t n, const void *d_xadj, int32_t offset_type,
              const int32_t *d_match, int32_t *d_cmap, int64_t *nc);
/* Contraction given match and cmap: replaces cuMetis_gpu_create_cgraph
 * (cu:1291-1406).  d_cxadj: nc+1 offsets; d_cadjncy/d_cadjwgt: room for m
 * entries; d_cvwgt: nc.  *mc receives the coarse entry count. */
int cm_contract(cm_ctx *ctx, const cm_graph_view *g, const int32_t *d_match,
                const int32_t *d_cmap, int64_t nc, int32_t row_order, void *d_cxadj,
                int32_t *d_cadjncy, int32_t *d_cadjwgt, int32_t *d_cvwgt, int64_t *mc);
/* Segmented sort, the bb_segsort<int,int> role (include/bb_segsort.h:35-36):
 * segments start at d_segs[0..length), the last ends at n; keys ascending,
 * values follow their keys.  Unlike bb_segsort the order of equal keys is
 * fixed (by original position). */
int cm_segsort_i32(cm_ctx *ctx, int32_t *d_keys, int32_t *d_vals, int32_t n,
                   const int32_t *d_segs, int32_t length);
/* cm_segsort_i32 without allocating: no counterpart in the reference; for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid it).  Same arguments and the
 * same result as cm_segsort_i32; all scratch comes from one static device
 * array of CM_NOALLOC_SCRATCH_BYTES per device (held until the process
 * ends), so the call performs no hipMalloc/hipFree.  It needs about
 * 24 n + 9 length + 64 KiB bytes of that array (n keys in `length`
 * segments: about one million keys in as many segments, 1.3 million in
 * long segments): larger calls return CM_ERR_UNSUPPORTED.  Every call on one device shares that device's
 * array and it keeps the last call's contents: calls that overlap on one
 * device (other contexts, streams or host threads) race. */
#define CM_NOALLOC_SCRATCH_BYTES (32u << 20)
int cm_segsort_i32_noalloc(cm_ctx *ctx, int32_t *d_keys, int32_t *d_vals, int32_t n,
                           const int32_t *d_segs, int32_t length);
/* Projection on raw arrays (cu:3067-3075). */
int cm_project_raw(cm_ctx *ctx, int64_t n, const int32_t *d_cmap, const int32_t *d_where_c,
                   int32_t *d_where_f);

/* ------------------------------------------------------ METIS-style API */
/* Drop-in for cuMetis_PartGraph (include/cuMetis.h:18-19, cu:3622-3670):
 * host pointers, writes part[n]; vwgt/adjwgt may be NULL; tpwgts/ubvec are
 * ignored as in the reference (uniform targets, ub 1.03).  Unlike the
 * reference, *objval receives the edge cut, and errors return a negative
 * cm_status instead of exiting.  Returns 1 on success like the reference. */
int cuMetis_PartGraph(int *nvtxs, int *xadj, int *adjncy, int *vwgt, int *adjwgt, int *nparts,
                      float *tpwgts, float *ubvec, int *objval, int *part);

/* Same pipeline with explicit options and per-phase timings (ms; levels
 * and coarsest_n are the kept trial's).  The trials' host initial
 * partitions run in threads of their own while the GPU coarsens the later
 * trials: coarsen_ms sums the coarsenings, init_ms is the wall time the
 * initial partitions add beyond them, uncoarsen_ms the wall time of the
 * trials' uncoarsenings (one after another on the stream). */
typedef struct {
  double total_ms, coarsen_ms, init_ms, uncoarsen_ms;
  int32_t levels;
  int32_t coarsest_n;
  int64_t edgecut;
  int32_t trials;     /* trials run */
  int32_t best_trial; /* the kept trial (its seed is params.seed + best_trial) */
  int32_t init_device_steps; /* recursion steps of the initial partitions coarsened on the device */
  int32_t reserved;
} cm_part_report;
int cm_partition(cm_ctx *ctx, const cm_graph_view *g, const cm_coarsen_params *p, int32_t *part,
                 cm_part_report *report);

/* -------------------------------------------------------------- serving */
/* Continuous-batching partition server.  Submitted host graphs (copied at
 * submit) are admitted as slots (max_slots) and the HBM budget (bytes, 0 =
 * 64 GiB; estimated per job from its size) free up.  Each coarsening step
 * runs one level for all admitted jobs as one batched union, so new jobs
 * join the next step while others are mid-hierarchy; finished hierarchies
 * stay device-resident while a second thread runs the initial partition and
 * uncoarsening of each job, overlapping the coarsening steps.  A job's result
 * is identical to cm_partition of the same graph alone.  Only
 * CM_ROW_FIRST_SEEN is accepted. */
typedef struct cm_server cm_server;
int cm_server_create(int device, int64_t hbm_budget, int32_t max_slots, cm_server **out);
int cm_server_submit(cm_server *s, const cm_graph_view *g, const cm_coarsen_params *p, uint64_t *ticket);
/* Blocks until the job is done; part receives n entries; report total_ms is
 * the latency from submit to result.  A ticket can be waited once. */
int cm_server_wait(cm_server *s, uint64_t ticket, int32_t *part, cm_part_report *report);
int cm_server_stats(cm_server *s, int64_t *completed, double *p50_ms, double *p99_ms, double *graphs_per_s,
                    int32_t *max_active);
int cm_server_destroy(cm_server *s);

/* ------------------------------------------------------------ host parts */
/* METIS .graph reader (cu:3712-3851): arrays are malloc'd, release each
 * with cm_free_host.  Malformed input returns CM_ERR_INPUT with the message
 * in err (the reference exits instead).  Self loops are rejected. */
int cm_read_metis(const char *path, int64_t *n, int64_t *m, int64_t **xadj, int32_t **adjncy,
                  int32_t **adjwgt, int32_t **vwgt, char *err, int errcap);
void cm_free_host(void *p);
/* "<graph>.part.<k>" writer: one part id per line (cu:3856-3870). */
int cm_write_partition(const char *path, const int32_t *part, int64_t n);
/* The host initial partition alone (cuMetis_initialpartition, cu:3091-3108;
 * recursive bisection cu:2964-3025).  Draws the glibc rand() sequence of
 * seed 1 (the reference never seeds) from a private, per-call stream;
 * *objval receives the summed bisection cuts. */
int cm_initial_partition(int64_t n, const int64_t *xadj, const int32_t *adjncy, const int32_t *adjwgt,
                         const int32_t *vwgt, int32_t nparts, int32_t *where, int64_t *objval);
/* The parallel form that cm_partition and cm_server use for coarsest graphs
 * of 2^14 vertices or more (see CM_FLAG_EXACT_INITPART), at any size; no
 * counterpart in the reference, whose serial form takes minutes on such
 * graphs.  Deterministic for any thread count; *objval = the edge cut. */
int cm_initial_partition_par(int64_t n, const int64_t *xadj, const int32_t *adjncy, const int32_t *adjwgt,
                             const int32_t *vwgt, int32_t nparts, int32_t *where, int64_t *objval);

/* Seeded benchmark graphs (host, multithreaded; arrays malloc'd, release
 * with cm_free_host): symmetric CSR, rows ascending, no self loops or
 * duplicates, weights per undirected edge mirrored.
 *   R-MAT: 2^scale vertices, edge_factor * 2^scale generated edges with
 *          quadrant probabilities a, b, c (d = 1-a-b-c), unit weights.
 *   RGG:   n uniform points in the unit square, radius sqrt(d/(pi n)) for
 *          mean degree d, weights uniform in [1, wmax] (1 if wmax <= 0). */
int cm_gen_rmat(int32_t scale, int32_t edge_factor, double a, double b, double c, uint64_t seed,
                int64_t *n, int64_t **xadj, int32_t **adjncy, int32_t **adjwgt);
int cm_gen_rgg2d(int64_t n, double mean_degree, int32_t wmax, uint64_t seed, int64_t **xadj,
                 int32_t **adjncy, int32_t **adjwgt);

#ifdef __cplusplus
}
#endif
#endif
