// cm_contract.hip -- (b) coarse numbering and (c) contraction.
//
// (b) replaces findc1..findc4 + thrust scan + host scan (cu:1621-1694,
//     cu:1750-1789, cu:1304): one single-pass decoupled look-back scan over
//     the fine vertices computes, for every representative i (i <= match[i]),
//     its coarse id, the coarse vertex weight and the offset of its gather
//     segment (deg(i) + deg(match[i])); a second pass copies cmap to partners.
// (c) replaces cuMetis_gpu_create_cgraph (cu:1291-1406): gather the two fine
//     rows of each coarse vertex as keys (coarse neighbour, position) ->
//     binned segmented sort -> per-segment duplicate merge -> look-back scan
//     of coarse degrees (cxadj) -> emit rows in ascending or first-seen order.
//     First-seen order reproduces the reference CPU contraction (cu:737-825)
//     byte for byte: distinct neighbours ranked by first position, self entry
//     replaced by the last entry.
#include "cm_common.hpp"
#include "cm_contract_hash.hpp"
#include "cm_internal.hpp"
#include "cm_segsort.hpp"

namespace cm {

constexpr int GATHER_LONG = 64;   // fine rows longer than this gather by wave
constexpr int SMALL_SEG = 64;     // coarse segments up to this merge per thread
constexpr int FIRST_SORTED_BIN = 4;  // segsort bins from REG_SEG + 1 = 17 keys up
static_assert(REG_SEG == 16, "FIRST_SORTED_BIN must be the first bin above REG_SEG");

// ------------------------------------------------------------- numbering
// Tiles of 1024 * NUM_PER vertices; thread x of a tile takes the vertices
// tile_base + t * 1024 + x (t < NUM_PER), so that every load and store of a
// per-vertex array by a wave covers 64 consecutive entries (fewer tiles =
// a shorter look-back chain).  A vertex still unmatched (match = -1, the
// matching left unfinished) becomes its own coarse vertex and its match
// entry is completed here; a representative also gives its partner the
// coarse id.  Coarse count and segment length are scanned together as one
// word (length << 16 | count; a tile counts at most 8192 vertices).  Four
// look-back chains, walked by waves 0..3: coarse count and segment length
// (sums), largest coarse vertex weight and longest segment (maxima); the
// last tile stores the totals: totals[0] = nc, totals[3] =
// largest coarse vertex weight, totals[4] = longest segment (a bound on the
// longest coarse row).
constexpr int NUM_PER = 2;
// Levels of at most tune.number_one_tile_max (<= 1024 * NUM_PER_ONE)
// vertices are numbered by one tile: no look-back at all (each chain step
// goes through the memory-side coherence point, a few us on the small
// levels of a hierarchy).
constexpr int NUM_PER_ONE = 4;  // 8 would spill: 1024-thread tiles have 128 VGPRs

template <typename OffT, int NUM_PER>
__global__ __launch_bounds__(1024) void k_number(int64_t n, const OffT *__restrict__ xadj,
                                                 const int32_t *match_in, int32_t *match,
                                                 const int32_t *__restrict__ vwgt,
                                                 int32_t *__restrict__ cmap, int32_t *__restrict__ rep,
                                                 OffT *__restrict__ segoff, int32_t *__restrict__ cvwgt,
                                                 uint64_t *st_c, uint64_t *st_l, uint64_t *st_m, uint64_t *st_n,
                                                 unsigned *ctr, int64_t *totals, HcLists sel, uint32_t epoch)
{
  static_assert(1024 * NUM_PER < (1 << 16), "a tile's coarse count fits the low 16 bits of the scan word");
  constexpr int NW = 1024 / CM_WAVE;  // waves per tile
  constexpr int NE = NW * NUM_PER;    // (stripe, wave) slots, in vertex order
  __shared__ uint64_t sh_x[NE], sh_p[NE], sh_m[NW], sh_n[NW], sh_t[2];
  const int tile = tile_ticket(ctr);
  const int64_t i0 = (int64_t)tile * 1024 * NUM_PER + threadIdx.x;
  const int w = threadIdx.x >> 6, l = lane();
  uint64_t xin[NUM_PER], mw = 0, ml = 0;
  uint32_t ln[NUM_PER];
  int32_t uu[NUM_PER], cw[NUM_PER], u0[NUM_PER];
  bool fl[NUM_PER];
  // two dependency levels (match, then rows and weights of both ends), each
  // loaded unconditionally through indices clamped to n - 1
#pragma unroll
  for (int t = 0; t < NUM_PER; ++t) {
    const int64_t i = i0 + (int64_t)t * 1024;
    u0[t] = match_in[i < n ? i : n - 1];
  }
#pragma unroll
  for (int t = 0; t < NUM_PER; ++t) {
    const int64_t i = i0 + (int64_t)t * 1024;
    const int64_t ic = i < n ? i : n - 1;
    const int32_t u = u0[t] < 0 ? (int32_t)ic : u0[t];
    const OffT xa = xadj[ic], xb = xadj[ic + 1], ya = xadj[u], yb = xadj[u + 1];
    const int32_t wa = vwgt[ic], wb = vwgt[u];
    uu[t] = u;
    fl[t] = i < n && i <= u;
    ln[t] = fl[t] ? (uint32_t)(xb - xa) + (u != i ? (uint32_t)(yb - ya) : 0u) : 0u;
    cw[t] = fl[t] ? wa + (u != i ? wb : 0) : 0;
    if (i < n && (u0[t] < 0 || match_in != match)) match[i] = u;
    if (fl[t]) {
      mw = (uint64_t)(uint32_t)cw[t] > mw ? (uint64_t)(uint32_t)cw[t] : mw;
      ml = ln[t] > ml ? ln[t] : ml;
    }
  }
  // per stripe: the wave-inclusive scan of (length << 16 | count)
#pragma unroll
  for (int t = 0; t < NUM_PER; ++t) {
    const uint64_t x = ((uint64_t)ln[t] << 16) | (fl[t] ? 1u : 0u);
    const uint64_t inc = wave_incl_scan<uint64_t>(x);
    xin[t] = inc - x;
    if (l == CM_WAVE - 1) sh_x[t * NW + w] = inc;
  }
  mw = wave_max<uint64_t>(mw);
  ml = wave_max<uint64_t>(ml);
  if (l == CM_WAVE - 1) {
    sh_m[w] = mw;
    sh_n[w] = ml;
  }
  __syncthreads();
  // waves 0..3 walk one look-back chain each, at the same time (a walk is
  // a few round trips to the memory-side coherence point); waves 0 and 2
  // both sum the slots, wave 0 also writes their exclusive prefixes (over
  // the NE slots, stripe-major, then wave: the vertex order)
  if (w == 0 || w == 2) {
    constexpr int EPL = (NE + CM_WAVE - 1) / CM_WAVE;  // consecutive slots per lane
    uint64_t s[EPL], sum = 0;
#pragma unroll
    for (int e = 0; e < EPL; ++e) {
      const int j = l * EPL + e;
      s[e] = j < NE ? sh_x[j] : 0;
      sum += s[e];
    }
    const uint64_t inc = wave_incl_scan<uint64_t>(sum);
    const uint64_t tot = __shfl(inc, CM_WAVE - 1, CM_WAVE);
    if (w == 0) {
      uint64_t run = inc - sum;
#pragma unroll
      for (int e = 0; e < EPL; ++e) {
        const int j = l * EPL + e;
        if (j < NE) sh_p[j] = run;
        run += s[e];
      }
      const uint64_t ec = lookback(st_c, tile, tot & 0xFFFFu, epoch);
      if (l == 0) sh_t[0] = ec;
    } else {
      const uint64_t el = lookback(st_l, tile, tot >> 16, epoch);
      if (l == 0) sh_t[1] = el;
    }
  } else if (w == 1 || w == 3) {
    const uint64_t b = wave_reduce<LbMax>(l < NW ? (w == 1 ? sh_m[l] : sh_n[l]) : 0);
    const uint64_t e = lookback<LbMax>(w == 1 ? st_m : st_n, tile, b, epoch);
    if (l == 0 && (int64_t)(tile + 1) * 1024 * NUM_PER >= n)  // the last tile
      totals[w == 1 ? 3 : 4] = (int64_t)(e > b ? e : b);
  }
  __syncthreads();
  const uint64_t ec = sh_t[0], el = sh_t[1];
#pragma unroll
  for (int t = 0; t < NUM_PER; ++t) {
    const int64_t i = i0 + (int64_t)t * 1024;
    const uint64_t ex = sh_p[t * NW + w] + xin[t];
    const uint64_t c = ec + (ex & 0xFFFFu), off = el + (ex >> 16);
    if (i < n) {
      if (fl[t]) {
        const int32_t u = uu[t];
        cmap[i] = (int32_t)c;
        if (u != (int32_t)i) cmap[u] = (int32_t)c;  // the partner (u > i)
        rep[c] = (int32_t)i;
        segoff[c] = (OffT)off;
        cvwgt[c] = cw[t];
      }
      if (i == n - 1) {
        totals[0] = (int64_t)(c + (fl[t] ? 1 : 0));
        segoff[c + (fl[t] ? 1 : 0)] = (OffT)(off + ln[t]);
      }
    }
    uu[t] = (int32_t)c;  // from here on: the coarse id
  }
  // the hash contraction's select step: segments above REG_SEG into the
  // item lists (all threads take part in the block-wide reservation)
  if (sel.ctr) {
    unsigned cnt[HC_N_COUNTERS] = {}, at[HC_N_COUNTERS];
#pragma unroll
    for (int t = 0; t < NUM_PER; ++t)
      if (i0 + (int64_t)t * 1024 < n && fl[t]) hc_count((int64_t)ln[t], cnt);
    hc_reserve<16>(cnt, at, sel.ctr);
#pragma unroll
    for (int t = 0; t < NUM_PER; ++t)
      if (i0 + (int64_t)t * 1024 < n && fl[t]) hc_write(uu[t], (int64_t)ln[t], at, sel);
  }
}

template <typename OffT>
static int64_t number_impl(cm_ctx *ctx, const Level &L, const int32_t *match_in, int32_t *match, int32_t *cmap,
                           int32_t *rep, OffT *segoff, int32_t *cvwgt, bool wait, HashContract *hc)
{
  const int64_t n = L.n;
  if (n == 0) {
    fill_u32(ctx, segoff, 0u, sizeof(OffT) / 4);
    fill_u32(ctx, ctx->d_small, 0u, 10);  // nc and the maxima (slots 0..4)
    return 0;
  }
  const bool one = n <= std::min<int64_t>(ctx->tune.number_one_tile_max, 1024 * NUM_PER_ONE);
  const int per = one ? NUM_PER_ONE : NUM_PER;
  const int64_t tiles = (n + 1024 * per - 1) / (1024 * per);
  const uint32_t ep = lookback_begin(ctx, 4 * tiles);  // 4 chains
  HcLists sel;
  if (hc) {  // the select step into hc's lists, whose counters start at zero
    sel = hc->lists;
    if (!ctx->hash_ctr_zero) fill_u32(ctx, sel.ctr, 0u, HC_CTR_WORDS);
    ctx->hash_ctr_zero = false;
    hc->selected = true;
  }
  launch(ctx, "number_scan", (double)sizeof(OffT) * (n + 1) * 2 + 16.0 * n,
         one ? k_number<OffT, NUM_PER_ONE> : k_number<OffT, NUM_PER>, dim3((unsigned)tiles), dim3(1024), 0, n,
         (const OffT *)L.xadj, match_in, match, (const int32_t *)L.vwgt, cmap, rep, segoff, cvwgt, ctx->d_status,
         ctx->d_status + tiles, ctx->d_status + 2 * tiles, ctx->d_status + 3 * tiles, ctx->d_ctr, ctx->d_small, sel,
         ep);
  if (!wait) return -1;
  CM_HIP(hipMemcpyAsync(ctx->h_small, ctx->d_small, sizeof(int64_t), hipMemcpyDeviceToHost,
                        ctx->stream));
  sync(ctx);
  return ctx->h_small[0];
}

int64_t number(cm_ctx *ctx, const Level &L, int32_t *d_match, int32_t *d_cmap,
               int32_t *d_rep, void *d_segoff, int32_t *d_cvwgt, bool wait, const int32_t *d_match_in,
               HashContract *hc)
{
  const int32_t *in = d_match_in ? d_match_in : d_match;
  return L.off64 ? number_impl(ctx, L, in, d_match, d_cmap, d_rep, (int64_t *)d_segoff, d_cvwgt, wait, hc)
                 : number_impl(ctx, L, in, d_match, d_cmap, d_rep, (int32_t *)d_segoff, d_cvwgt, wait, hc);
}

// ------------------------------------------------- generic look-back scan
// out[i] = sum of in[0..i), out[n] = total; tiles of 1024.
template <typename OffT>
__global__ __launch_bounds__(1024) void k_scan_excl(int64_t n, const int32_t *__restrict__ in,
                                                    OffT *__restrict__ out, uint64_t *st,
                                                    unsigned *ctr, int64_t *total, uint32_t epoch)
{
  __shared__ uint64_t sh[16];
  const int tile = tile_ticket(ctr);
  const int64_t i = (int64_t)tile * 1024 + threadIdx.x;
  const uint64_t v = i < n ? (uint64_t)(uint32_t)in[i] : 0;
  const int w = threadIdx.x >> 6, l = lane();
  const uint64_t iv = wave_incl_scan<uint64_t>(v);
  if (l == CM_WAVE - 1) sh[w] = iv;
  __syncthreads();
  if (w == 0) {
    const uint64_t a = l < 16 ? sh[l] : 0;
    const uint64_t ia = wave_incl_scan<uint64_t>(a);
    const uint64_t ta = __shfl(ia, 15, CM_WAVE);
    const uint64_t ea = lookback(st, tile, ta, epoch);
    if (l < 16) sh[l] = ea + ia - a;
  }
  __syncthreads();
  if (i < n) {
    const uint64_t ex = sh[w] + iv - v;
    out[i] = (OffT)ex;
    if (i == n - 1) {
      out[n] = (OffT)(ex + v);
      total[0] = (int64_t)(ex + v);
    }
  }
}

template <typename OffT>
static int64_t scan_excl(cm_ctx *ctx, int64_t n, const int32_t *in, OffT *out)
{
  if (n == 0) {
    fill_u32(ctx, out, 0u, sizeof(OffT) / 4);
    return 0;
  }
  const int64_t tiles = (n + 1023) / 1024;
  const uint32_t ep = lookback_begin(ctx, tiles);
  launch(ctx, "scan_cxadj", 4.0 * n + sizeof(OffT) * (n + 1), k_scan_excl<OffT>, dim3((unsigned)tiles),
         dim3(1024), 0, n, in, out, ctx->d_status, ctx->d_ctr, ctx->d_small + 1, ep);
  CM_HIP(hipMemcpyAsync(ctx->h_small + 1, ctx->d_small + 1, sizeof(int64_t), hipMemcpyDeviceToHost,
                        ctx->stream));
  sync(ctx);
  return ctx->h_small[1];
}

// exclusive scan of n int32 values into int32 offsets (out[n] = total);
// used by the segment binning
void scan_excl_i32(cm_ctx *ctx, int64_t n, const int32_t *in, int32_t *out)
{
  if (n == 0) {
    fill_u32(ctx, out, 0u, 1);
    return;
  }
  const int64_t tiles = (n + 1023) / 1024;
  const uint32_t ep = lookback_begin(ctx, tiles);
  scan_excl_i32_status(ctx, n, in, out, ctx->d_status, ep);
}

void scan_excl_i32_status(cm_ctx *ctx, int64_t n, const int32_t *in, int32_t *out, uint64_t *status,
                          uint32_t epoch)
{
  if (n == 0) {
    fill_u32(ctx, out, 0u, 1);
    return;
  }
  const int64_t tiles = (n + 1023) / 1024;
  launch(ctx, "scan_bins", 0.0, k_scan_excl<int32_t>, dim3((unsigned)tiles), dim3(1024), 0, n, in, out, status,
         ctx->d_ctr, ctx->d_small + 2, epoch);
}

// ------------------------------------------------------------------ gather
// Fine vertex i writes its row into the segment of its coarse vertex: the
// representative's row first, then the partner's.  Key = (coarse neighbour,
// position in segment), value = edge weight.
template <typename OffT>
__global__ __launch_bounds__(256) void k_gather(int64_t n, const OffT *__restrict__ xadj,
                                                const int32_t *__restrict__ adjncy,
                                                const int32_t *__restrict__ adjwgt,
                                                const int32_t *__restrict__ match,
                                                const int32_t *__restrict__ cmap,
                                                const OffT *__restrict__ segoff,
                                                uint64_t *__restrict__ keys, uint32_t *__restrict__ vals)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const OffT b = xadj[i], e = xadj[i + 1];
  if ((int64_t)(e - b) > GATHER_LONG) return;
  const int32_t u = match[i];
  const int32_t c = cmap[i];
  const int64_t s0 = (int64_t)segoff[c];
  if ((int64_t)segoff[c + 1] - s0 <= REG_SEG) return;  // merged in registers by k_contract_small
  const uint32_t local = u < i ? (uint32_t)(xadj[u + 1] - xadj[u]) : 0u;
  for (OffT j = b; j < e; ++j) {
    const uint32_t pos = local + (uint32_t)(j - b);
    keys[s0 + pos] = ((uint64_t)(uint32_t)cmap[adjncy[j]] << 32) | pos;
    vals[s0 + pos] = (uint32_t)edge_w(adjwgt, j);
  }
}

template <typename OffT>
__global__ __launch_bounds__(256) void k_gather_long(const int32_t *list, const unsigned *list_n,
                                                     const OffT *__restrict__ xadj,
                                                     const int32_t *__restrict__ adjncy,
                                                     const int32_t *__restrict__ adjwgt,
                                                     const int32_t *__restrict__ match,
                                                     const int32_t *__restrict__ cmap,
                                                     const OffT *__restrict__ segoff,
                                                     uint64_t *__restrict__ keys,
                                                     uint32_t *__restrict__ vals)
{
  const unsigned nl = *list_n;
  const int l = lane();
  const int64_t nw = (int64_t)gridDim.x * (blockDim.x / CM_WAVE);
  for (int64_t wi = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / CM_WAVE; wi < nl; wi += nw) {
    const int32_t i = list[wi];
    const OffT b = xadj[i], e = xadj[i + 1];
    const int32_t u = match[i];
    const int64_t s0 = (int64_t)segoff[cmap[i]];
    const uint32_t local = u < i ? (uint32_t)(xadj[u + 1] - xadj[u]) : 0u;
    for (OffT j = b + l; j < e; j += CM_WAVE) {
      const uint32_t pos = local + (uint32_t)(j - b);
      keys[s0 + pos] = ((uint64_t)(uint32_t)cmap[adjncy[j]] << 32) | pos;
      vals[s0 + pos] = (uint32_t)edge_w(adjwgt, j);
    }
  }
}

template <typename OffT>
__global__ __launch_bounds__(256) void k_select_long_rows(int64_t n, const OffT *xadj, int32_t *list,
                                                          unsigned *count)
{
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool is_long = v < n && (int64_t)(xadj[v + 1] - xadj[v]) > GATHER_LONG;
  const uint64_t mk = __ballot(is_long);
  if (!mk) return;
  const int leader = __ffsll((unsigned long long)mk) - 1;
  unsigned base = 0;
  if (lane() == leader) base = atomicAdd(count, (unsigned)__popcll(mk));
  base = __shfl(base, leader, CM_WAVE);
  if (is_long) list[base + mask_rank(mk)] = (int32_t)v;
}

// ---------------------------------------------------------- merge: count
// Sorted segment of coarse vertex c: a head is the first key of a run of
// equal coarse neighbours; its low word is the neighbour's first position.
// cdeg[c] = heads other than c.  First-seen mode also ranks the non-self
// heads by first position (written to rank[s0+pos]) and records the self
// head's position and rank.
__device__ __forceinline__ uint32_t khi(uint64_t k) { return (uint32_t)(k >> 32); }
__device__ __forceinline__ uint32_t klo(uint64_t k) { return (uint32_t)k; }

template <bool FS, typename OffT>
__global__ __launch_bounds__(256) void k_count_small(const int32_t *__restrict__ list, int64_t nlist,
                                                     const OffT *__restrict__ segoff,
                                                     const uint64_t *__restrict__ keys,
                                                     int32_t *__restrict__ rank,
                                                     int32_t *__restrict__ cdeg,
                                                     int32_t *__restrict__ selfpos,
                                                     int32_t *__restrict__ selfrank)
{
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nlist) return;
  const int64_t c = list[q];
  const int64_t s0 = (int64_t)segoff[c], s1 = (int64_t)segoff[c + 1];
  const int len = (int)(s1 - s0);
  if (FS)
    for (int p = 0; p < len; ++p) rank[s0 + p] = -1;
  int cnt = 0, ps = -1;
  uint32_t prev = 0xFFFFFFFFu;
  for (int64_t e = s0; e < s1; ++e) {
    const uint64_t k = keys[e];
    if (khi(k) != prev) {
      prev = khi(k);
      if (prev == (uint32_t)c) ps = (int)klo(k);
      else {
        ++cnt;
        if (FS) rank[s0 + klo(k)] = 1;
      }
    }
  }
  if (FS) {
    int r = 0, sr = -1;
    for (int p = 0; p < len; ++p) {
      if (p == ps) sr = r;
      if (rank[s0 + p] >= 0) rank[s0 + p] = r++;
    }
    selfpos[c] = ps;
    selfrank[c] = sr;
  }
  cdeg[c] = cnt;
}

template <bool FS, typename OffT>
__global__ __launch_bounds__(256) void k_count_large(const int32_t *__restrict__ list, int64_t nlist,
                                                     const OffT *__restrict__ segoff,
                                                     const uint64_t *__restrict__ keys,
                                                     int32_t *__restrict__ rank,
                                                     int32_t *__restrict__ cdeg,
                                                     int32_t *__restrict__ selfpos,
                                                     int32_t *__restrict__ selfrank)
{
  __shared__ int64_t sh[4];
  __shared__ int s_ps, s_sr;
  const int t = threadIdx.x;
  for (int64_t g = blockIdx.x; g < nlist; g += gridDim.x) {
    const int32_t c = list[g];
    const int64_t s0 = (int64_t)segoff[c], s1 = (int64_t)segoff[c + 1];
    const int len = (int)(s1 - s0);
    if (FS)
      for (int p = t; p < len; p += 256) rank[s0 + p] = -1;
    if (t == 0) {
      s_ps = -1;
      s_sr = -1;
    }
    __syncthreads();
    int64_t cnt = 0;
    for (int64_t e = s0 + t; e < s1; e += 256) {
      const uint64_t k = keys[e];
      const bool head = e == s0 || khi(keys[e - 1]) != khi(k);
      if (head) {
        if (khi(k) == (uint32_t)c) s_ps = (int)klo(k);
        else {
          ++cnt;
          if (FS) rank[s0 + klo(k)] = 1;
        }
      }
    }
    int64_t total;
    block256_excl_scan<int64_t>(cnt, total, sh);
    if (FS) {
      const int ps = s_ps;
      int64_t carry = 0;
      for (int base = 0; base < len; base += 256) {
        const int p = base + t;
        const int64_t f = (p < len && rank[s0 + p] >= 0) ? 1 : 0;
        int64_t tot;
        const int64_t ex = block256_excl_scan<int64_t>(f, tot, sh);
        if (f) rank[s0 + p] = (int32_t)(carry + ex);
        if (p == ps) s_sr = (int)(carry + ex);
        carry += tot;
      }
    }
    __syncthreads();
    if (t == 0) {
      cdeg[c] = (int32_t)total;
      if (FS) {
        selfpos[c] = s_ps;
        selfrank[c] = s_sr;
      }
    }
    __syncthreads();
  }
}

// ----------------------------------------------------------- merge: emit
template <bool FS>
__device__ __forceinline__ int64_t final_slot(int r, uint32_t p, int ps, int sr, int64_t cdeg)
{
  if (!FS) return r;
  // position in the reference's list (self included), then swap-with-last
  const int64_t r1 = r + ((ps >= 0 && (uint32_t)ps < p) ? 1 : 0);
  return (ps >= 0 && r1 == cdeg) ? (int64_t)sr : r1;
}

template <bool FS, typename OffT>
__global__ __launch_bounds__(256) void k_emit_small(const int32_t *__restrict__ list, int64_t nlist,
                                                    const OffT *__restrict__ segoff,
                                                    const uint64_t *__restrict__ keys,
                                                    const uint32_t *__restrict__ vals,
                                                    const int32_t *__restrict__ rank,
                                                    const int32_t *__restrict__ selfpos,
                                                    const int32_t *__restrict__ selfrank,
                                                    const OffT *__restrict__ cxadj,
                                                    int32_t *__restrict__ cadjncy,
                                                    int32_t *__restrict__ cadjwgt)
{
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nlist) return;
  const int64_t c = list[q];
  const int64_t s0 = (int64_t)segoff[c], s1 = (int64_t)segoff[c + 1];
  const int64_t out = (int64_t)cxadj[c], deg = (int64_t)cxadj[c + 1] - out;
  const int ps = FS ? selfpos[c] : -1, sr = FS ? selfrank[c] : -1;
  int r = 0;
  for (int64_t e = s0; e < s1;) {
    const uint64_t k = keys[e];
    const uint32_t nb = khi(k);
    uint32_t w = vals[e];
    int64_t e2 = e + 1;
    while (e2 < s1 && khi(keys[e2]) == nb) w += vals[e2++];
    if (nb != (uint32_t)c) {
      const int rr = FS ? rank[s0 + klo(k)] : r;
      const int64_t slot = out + final_slot<FS>(rr, klo(k), ps, sr, deg);
      cadjncy[slot] = (int32_t)nb;
      cadjwgt[slot] = (int32_t)w;
      ++r;
    }
    e = e2;
  }
}

template <bool FS, typename OffT>
__global__ __launch_bounds__(256) void k_emit_large(const int32_t *__restrict__ list, int64_t nlist,
                                                    const OffT *__restrict__ segoff,
                                                    const uint64_t *__restrict__ keys,
                                                    const uint32_t *__restrict__ vals,
                                                    const int32_t *__restrict__ rank,
                                                    const int32_t *__restrict__ selfpos,
                                                    const int32_t *__restrict__ selfrank,
                                                    const OffT *__restrict__ cxadj,
                                                    int32_t *__restrict__ cadjncy,
                                                    int32_t *__restrict__ cadjwgt)
{
  __shared__ int64_t sh[4];
  const int t = threadIdx.x;
  for (int64_t g = blockIdx.x; g < nlist; g += gridDim.x) {
    const int32_t c = list[g];
    const int64_t s0 = (int64_t)segoff[c], s1 = (int64_t)segoff[c + 1];
    const int64_t out = (int64_t)cxadj[c], deg = (int64_t)cxadj[c + 1] - out;
    const int ps = FS ? selfpos[c] : -1, sr = FS ? selfrank[c] : -1;
    int64_t carry = 0;
    for (int64_t base = s0; base < s1; base += 256) {
      const int64_t e = base + t;
      uint64_t k = 0;
      bool head = false;
      if (e < s1) {
        k = keys[e];
        head = (e == s0 || khi(keys[e - 1]) != khi(k)) && khi(k) != (uint32_t)c;
      }
      int64_t ex = 0;
      if (!FS) {
        int64_t tot;
        ex = block256_excl_scan<int64_t>(head ? 1 : 0, tot, sh);
        ex += carry;
        carry += tot;
      }
      if (head) {
        const uint32_t nb = khi(k);
        uint32_t w = vals[e];
        for (int64_t e2 = e + 1; e2 < s1 && khi(keys[e2]) == nb; ++e2) w += vals[e2];
        const int rr = FS ? rank[s0 + klo(k)] : (int)ex;
        const int64_t slot = out + final_slot<FS>(rr, klo(k), ps, sr, deg);
        cadjncy[slot] = (int32_t)nb;
        cadjwgt[slot] = (int32_t)w;
      }
    }
  }
}

// ------------------------------------------ short segments: fused, in registers
// Bitonic network over N register entries (static indices once unrolled).
template <int N>
__device__ __forceinline__ void reg_bitonic(uint64_t (&k)[N], uint64_t (&p)[N])
{
#pragma unroll
  for (int sz = 2; sz <= N; sz <<= 1) {
#pragma unroll
    for (int j = sz >> 1; j > 0; j >>= 1) {
#pragma unroll
      for (int i = 0; i < N; ++i) {
        const int l = i ^ j;
        if (l > i) {
          const bool up = (i & sz) == 0;
          if (up ? (k[i] > k[l]) : (k[i] < k[l])) {
            const uint64_t tk = k[i];
            k[i] = k[l];
            k[l] = tk;
            const uint64_t tp = p[i];
            p[i] = p[l];
            p[l] = tp;
          }
        }
      }
    }
  }
}

// Merge the segment of coarse vertex c (row(v) then row(u), at most N
// entries) entirely in registers.  On return out_key/out_w hold the row in
// its final order (ascending, or first-seen with the self entry replaced by
// the last one) and the count of entries.
template <int N, bool FS, typename OffT>
__device__ __forceinline__ int merge_regs(int32_t c, int32_t v, int32_t u, const OffT *__restrict__ xadj,
                                          const int32_t *__restrict__ adjncy,
                                          const int32_t *__restrict__ adjwgt,
                                          const int32_t *__restrict__ cmap, uint64_t (&k)[N],
                                          uint64_t (&p)[N])
{
  // Loads are unconditional (invalid slots read entry 0 / vertex 0 and are
  // discarded), so all of them issue before the first wait; conditional
  // loads would each get a branch and a wait of their own.
  const int64_t b0 = (int64_t)xadj[v];
  const int64_t e0 = (int64_t)xadj[v + 1];
  const int64_t b1 = (int64_t)xadj[u];
  const int64_t e1 = (int64_t)xadj[u + 1];
  const int d0 = (int)(e0 - b0);
  const int d1 = u != v ? (int)(e1 - b1) : 0;
  int32_t a[N], wt[N];
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const bool valid = i < d0 + d1;
    const int64_t j = valid ? (i < d0 ? b0 + i : b1 + (i - d0)) : 0;
    a[i] = adjncy[j];
    wt[i] = edge_w(adjwgt, j);
  }
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const bool valid = i < d0 + d1;
    const int32_t cu = cmap[valid ? a[i] : 0];
    k[i] = valid ? (((uint64_t)(uint32_t)cu << 32) | (uint32_t)i) : ~0ull;
    p[i] = valid ? (uint64_t)(uint32_t)wt[i] : 0ull;
  }
  reg_bitonic<N>(k, p);
  // run sums (runs are adjacent after the sort), heads = first of a run
  uint64_t sum[N];
#pragma unroll
  for (int i = N - 1; i >= 0; --i) {
    const bool cont = i + 1 < N && (k[i + 1] >> 32) == (k[i] >> 32);
    sum[i] = p[i] + (cont ? sum[(i + 1) < N ? i + 1 : i] : 0ull);
  }
  // order the heads: by first position (first-seen, self included) or by
  // sorted position without self (ascending); a second network compacts
  uint64_t k2[N], p2[N];
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const uint32_t key = (uint32_t)(k[i] >> 32);
    const bool head = k[i] != ~0ull && (i == 0 || (uint32_t)(k[i > 0 ? i - 1 : 0] >> 32) != key);
    if (FS) k2[i] = head ? (uint64_t)(uint32_t)k[i] : ~0ull;
    else k2[i] = (head && key != (uint32_t)c) ? (uint64_t)i : ~0ull;
    p2[i] = ((uint64_t)key << 32) | (uint32_t)sum[i];
  }
  reg_bitonic<N>(k2, p2);
  int len = 0, sidx = -1;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    if (k2[i] != ~0ull) {
      if (FS && (uint32_t)(p2[i] >> 32) == (uint32_t)c) sidx = i;
      ++len;
    }
  }
  if (FS && sidx >= 0) {
    // cu:804-808: the self entry takes the last entry's place
    uint64_t last = 0;
#pragma unroll
    for (int i = 0; i < N; ++i)
      if (i == len - 1) last = p2[i];
#pragma unroll
    for (int i = 0; i < N; ++i)
      if (i == sidx) p2[i] = last;
    --len;
  }
#pragma unroll
  for (int i = 0; i < N; ++i) p[i] = p2[i];
  return len;
}

// First-seen merge of a short segment without sorting (row(v) then row(u),
// cu:766-808): entry i is a head when no earlier entry has its coarse id, a
// head's weight is its own plus that of its later duplicates, heads keep
// their position order, and the head of the coarse vertex itself gives its
// slot to the last head (cu:804-808).  dst[i] = row slot of entry i (-1 for
// non-heads and the self head); returns the row length.  All comparisons use
// static indices, so the arrays stay in registers, and each pair (j < i) is
// compared once and used at once (a later duplicate adds its weight to the
// earlier entry and marks itself), so no comparison result stays live.
template <int N, typename OffT>
__device__ __forceinline__ int merge_fs(int32_t c, int32_t v, int32_t u, const OffT *__restrict__ xadj,
                                        const int32_t *__restrict__ adjncy,
                                        const int32_t *__restrict__ adjwgt,
                                        const int32_t *__restrict__ cmap, int32_t *cu, int32_t *sum,
                                        int32_t *dst)
{
  const int64_t b0 = (int64_t)xadj[v];
  const int64_t e0 = (int64_t)xadj[v + 1];
  const int64_t b1 = (int64_t)xadj[u];
  const int64_t e1 = (int64_t)xadj[u + 1];
  const int d0 = (int)(e0 - b0);
  const int cnt = d0 + (u != v ? (int)(e1 - b1) : 0);
  int32_t a[N];
#pragma unroll
  for (int i = 0; i < N; ++i) {  // unconditional loads (see merge_regs)
    const int64_t j = i < cnt ? (i < d0 ? b0 + i : b1 + (i - d0)) : 0;
    a[i] = adjncy[j];
    sum[i] = edge_w(adjwgt, j);
  }
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const int32_t x = cmap[i < cnt ? a[i] : 0];
    cu[i] = i < cnt ? x : -1 - i;  // distinct negative ids for empty slots never match
  }
  uint32_t dup = 0;  // bit i: entry i repeats an earlier one
#pragma unroll
  for (int i = 1; i < N; ++i) {
#pragma unroll
    for (int j = 0; j < i; ++j) {
      const bool eq = cu[j] == cu[i];
      sum[j] += eq ? sum[i] : 0;
      dup |= eq ? (1u << i) : 0u;
    }
  }
  int r = 0, sidx = -1;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const bool h = i < cnt && !((dup >> i) & 1u);
    dst[i] = h ? r : -1;
    if (h && cu[i] == c) sidx = r;
    r += h ? 1 : 0;
  }
  if (sidx < 0) return r;
#pragma unroll
  for (int i = 0; i < N; ++i)
    if (dst[i] == sidx) dst[i] = -1;  // the self head
#pragma unroll
  for (int i = 0; i < N; ++i)
    if (dst[i] == r - 1) dst[i] = sidx;  // the last head takes its slot
  return r - 1;
}

// One thread per coarse vertex, tiles of SMALL_TILE coarse vertices in
// dispatch order.  Segments of at most REG_SEG entries are merged in
// registers; longer ones have been merged by the sorted or hashed path,
// which left their degree in cdeg.  A look-back scan of the degrees gives
// cxadj, and the short rows are written straight to their final place: the
// fine rows are read once and the coarse rows written once.
// Tiles of 256 keep every CU busy on small levels and balance the large
// ones (the merge is VALU work per coarse vertex, so a tile's time is set by
// its waves per SIMD, not by memory).
constexpr int SMALL_TILE = 256;    // coarse vertices per tile of k_contract_small
constexpr int SMALL_STAGE = 4096;  // output entries staged in LDS per window
constexpr int COPY_BATCH = 8;      // staged entries per step of a tile-copied hashed row

// COPY: hashed rows are copied from the staging area by the tile (tc/tw
// given); without it the copy path is compiled out, which frees registers
// for more waves per SIMD.
template <bool FS, typename OffT, bool COPY = false>
__global__ __launch_bounds__(SMALL_TILE) void k_contract_small(const int64_t *__restrict__ ncp,
                                                        const OffT *__restrict__ segoff,
                                                        const int32_t *__restrict__ rep,
                                                        const int32_t *__restrict__ match,
                                                        const OffT *__restrict__ xadj,
                                                        const int32_t *__restrict__ adjncy,
                                                        const int32_t *__restrict__ adjwgt,
                                                        const int32_t *__restrict__ cmap,
                                                        const int32_t *__restrict__ cdeg_large,
                                                        const int32_t *__restrict__ tc,
                                                        const int32_t *__restrict__ tw,
                                                        OffT *__restrict__ cxadj, int32_t *__restrict__ cadjncy,
                                                        int32_t *__restrict__ cadjwgt, uint64_t *st,
                                                        unsigned *ctr, int64_t *total, uint32_t epoch,
                                                        Readback rb)
{
  __shared__ uint64_t sh[SMALL_TILE / CM_WAVE];
  __shared__ uint64_t s_excl;
  __shared__ int32_t stage_a[SMALL_STAGE], stage_w[SMALL_STAGE];
  const int64_t nc = *ncp;
  const int tile = tile_ticket(ctr);
  // the grid may be sized by a bound on nc: tiles past the end take no part
  // in the look-back chain (no later tile waits for them)
  if ((int64_t)tile * SMALL_TILE >= nc) return;
  const int64_t c = (int64_t)tile * SMALL_TILE + threadIdx.x;
  uint64_t k[REG_SEG], p[REG_SEG];            // ascending order: sorted (id, weight) pairs
  int32_t fcu[REG_SEG], fsum[REG_SEG], fdst[REG_SEG];  // first-seen order: entries and row slots
  int deg = 0;
  // loads by a clamped index, all issued before the first wait
  const int64_t cc = c < nc ? c : nc - 1;
  const int64_t s0 = (int64_t)segoff[cc], s1 = (int64_t)segoff[cc + 1];
  const int32_t v = rep[cc], dl = cdeg_large[cc];
  const int32_t u = match[v];
  const int len = (int)(s1 - s0);
  const bool small = c < nc && len <= REG_SEG;
  if constexpr (FS) {
    // the merge width is chosen per wave (pairs compared: 120, 28 or 6)
    const bool wide = __ballot(small && len > 8) != 0;
    const bool narrow = __ballot(small && len > 4) == 0;
    if (small) {
      if (wide) {
        deg = merge_fs<REG_SEG, OffT>((int32_t)c, v, u, xadj, adjncy, adjwgt, cmap, fcu, fsum, fdst);
      } else if (!narrow) {
        deg = merge_fs<8, OffT>((int32_t)c, v, u, xadj, adjncy, adjwgt, cmap, fcu, fsum, fdst);
#pragma unroll
        for (int i = 8; i < REG_SEG; ++i) fdst[i] = -1;
      } else {
        deg = merge_fs<4, OffT>((int32_t)c, v, u, xadj, adjncy, adjwgt, cmap, fcu, fsum, fdst);
#pragma unroll
        for (int i = 4; i < REG_SEG; ++i) fdst[i] = -1;
      }
    } else if (c < nc) {
      deg = dl;
    }
  } else if (c < nc) {
    if (small) {
      if (len <= 8) {
        uint64_t k8[8], p8[8];
        deg = merge_regs<8, FS, OffT>((int32_t)c, v, u, xadj, adjncy, adjwgt, cmap, k8, p8);
#pragma unroll
        for (int i = 0; i < 8; ++i) p[i] = p8[i];
      } else {
        deg = merge_regs<REG_SEG, FS, OffT>((int32_t)c, v, u, xadj, adjncy, adjwgt, cmap, k, p);
      }
    } else {
      deg = dl;
    }
  }
  uint64_t tot;
  const uint64_t ex = block_excl_scan<uint64_t, SMALL_TILE / CM_WAVE>((uint64_t)deg, tot, sh);
  if (threadIdx.x < CM_WAVE) {
    const uint64_t e = lookback(st, tile, tot, epoch);
    if (threadIdx.x == 0) s_excl = e;
  }
  __syncthreads();
  // the last tile's wave 0 sends the read-back of the small totals (total =
  // d_small + 1, the coarse entry count, is its own slot)
  if ((int64_t)(tile + 1) * SMALL_TILE >= nc && threadIdx.x < CM_WAVE)
    readback_write(rb, total - 1, 1, (int64_t)(s_excl + tot));
  const uint64_t off = s_excl + ex;
  if (c < nc) {
    cxadj[c] = (OffT)off;
    if (c == nc - 1) {
      cxadj[nc] = (OffT)(off + deg);
      total[0] = (int64_t)(off + deg);
    }
  }
  // Rows go out through LDS: each window of STAGE entries of the tile's
  // output range is filled by the rows that fall in it, then stored by
  // consecutive threads.  With tc/tw, the hashed rows (all of them then
  // come from segments of at most TILE_COPY_SEG entries) are copied in from
  // the staging area; otherwise their slots (written later by the emit
  // pass) keep the sentinel and are skipped.  Windows cover the range up to
  // the end of the last row written.  A tile whose range is longer than one
  // window (hashed rows in it, written by the emit pass) moves each window to
  // the first entry written at or after the end of the previous one, so the
  // hashed rows' ranges are not walked.
  const bool copied = COPY && c < nc && !small && len <= TILE_COPY_SEG;
  const bool writes = (small || copied) && deg > 0;
  const uint64_t my_end = writes ? ex + (uint64_t)deg : 0;
  uint64_t span = wave_max<uint64_t>(my_end);
  if (lane() == 0) sh[threadIdx.x >> 6] = span;
  __syncthreads();
  span = 0;
#pragma unroll
  for (int i = 0; i < SMALL_TILE / CM_WAVE; ++i) span = sh[i] > span ? sh[i] : span;
  // the first entry written at or after `from` (block-wide; all threads call)
  auto next_written = [&](uint64_t from) -> uint64_t {
    const uint64_t cand = writes && my_end > from ? (ex > from ? ex : from) : ~0ull;
    const uint64_t wm = wave_max<uint64_t>(~cand);
    __syncthreads();  // sh is free again
    if (lane() == 0) sh[threadIdx.x >> 6] = wm;
    __syncthreads();
    uint64_t r = 0;
#pragma unroll
    for (int i = 0; i < SMALL_TILE / CM_WAVE; ++i) r = sh[i] > r ? sh[i] : r;
    return ~r;
  };
  const bool skip = span > (uint64_t)SMALL_STAGE;  // block-uniform
  for (uint64_t w0 = skip ? next_written(0) : 0; w0 < span;
       w0 = skip ? next_written(w0 + SMALL_STAGE) : w0 + SMALL_STAGE) {
    const int wlen = (int)(span - w0 < (uint64_t)SMALL_STAGE ? span - w0 : (uint64_t)SMALL_STAGE);
    for (int k = threadIdx.x; k < wlen; k += SMALL_TILE) stage_a[k] = -1;
    __syncthreads();
    // the row's offset inside the window, in 32 bits once the row is known
    // to overlap it
    const int64_t rel = (int64_t)ex - (int64_t)w0;
    if (small && rel < wlen && rel + deg > 0) {
      const int r32 = (int)rel;
#pragma unroll
      for (int i = 0; i < REG_SEG; ++i) {
        if constexpr (FS) {
          const int pos = r32 + fdst[i];
          if (fdst[i] >= 0 && (unsigned)pos < (unsigned)wlen) {
            stage_a[pos] = fcu[i];
            stage_w[pos] = fsum[i];
          }
        } else {
          const int pos = r32 + i;
          if (i < deg && (unsigned)pos < (unsigned)wlen) {
            stage_a[pos] = (int32_t)(p[i] >> 32);
            stage_w[pos] = (int32_t)(uint32_t)p[i];
          }
        }
      }
    }
    if (COPY && copied && rel < wlen && rel + deg > 0) {
      const int r32 = (int)rel;
      const int i0 = r32 < 0 ? -r32 : 0, i1 = deg < wlen - r32 ? deg : wlen - r32;
      const int32_t *twp = tw ? tw : tc;  // unconditional loads (packed: the weights are in tc)
      for (int i = i0; i < i1; i += COPY_BATCH) {
        int32_t a[COPY_BATCH], w[COPY_BATCH];
#pragma unroll
        for (int t = 0; t < COPY_BATCH; ++t) {  // clamped loads, issued together
          const int j = i + t < i1 ? i + t : i1 - 1;
          a[t] = tc[s0 + j];
          w[t] = twp[s0 + j];
        }
        if (!tw) {  // packed staging: one word per entry
#pragma unroll
          for (int t = 0; t < COPY_BATCH; ++t) {
            w[t] = hc_unpack_w(a[t]);
            a[t] = hc_unpack_id(a[t]);
          }
        }
#pragma unroll
        for (int t = 0; t < COPY_BATCH; ++t)
          if (i + t < i1) {
            stage_a[r32 + i + t] = a[t];
            stage_w[r32 + i + t] = w[t];
          }
      }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < wlen; k += SMALL_TILE) {
      const int32_t a = stage_a[k];
      if (a != -1) {
        cadjncy[s_excl + w0 + k] = a;
        cadjwgt[s_excl + w0 + k] = stage_w[k];
      }
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------- host side
template <typename OffT>
static int64_t contract_impl(cm_ctx *ctx, const Level &L, const int32_t *match, const int32_t *cmap,
                             int64_t nc, const int32_t *rep, const OffT *segoff, int row_order, OffT *cxadj,
                             int32_t *cadjncy, int32_t *cadjwgt, int64_t *maxdeg, int64_t *ncp,
                             bool force_sorted = false, HashContract *pre = nullptr)
{
  const int64_t n = L.n, m = L.m;
  const OffT *xadj = (const OffT *)L.xadj;
  const bool fs = row_order == CM_ROW_FIRST_SEEN;
  // pre: the hash lists set up before the numbering (and filled by it); they
  // are used by the first-seen hash path and released here on every path
  struct PreGuard {
    cm_ctx *ctx;
    HashContract *hc;
    ~PreGuard()
    {
      if (hc) hash_release(ctx, *hc);
    }
  } pre_guard{ctx, pre};
  if (nc < 0 && (n == 0 || !fs || force_sorted)) {  // the paths below need nc on the host
    CM_HIP(hipMemcpyAsync(ctx->h_small, ctx->d_small, sizeof(int64_t), hipMemcpyDeviceToHost, ctx->stream));
    sync(ctx);
    nc = ctx->h_small[0];
  }
  if (ncp && nc >= 0) *ncp = nc;
  if (nc == 0) {
    fill_u32(ctx, cxadj, 0u, sizeof(OffT) / 4);
    if (maxdeg) *maxdeg = 0;
    return 0;
  }
  if (fs && !force_sorted) {
    // segments above REG_SEG (they exist only if 2 * maxdeg > REG_SEG) by
    // hashing, the rest in registers; one host synchronisation at the end
    // nc < 0: the count is on the device (d_small[0]); grids are sized by
    // the bound n and kernels read the count
    const bool hashed = L.maxdeg < 0 || 2 * L.maxdeg > REG_SEG;
    const int64_t ncb = nc < 0 ? n : nc;
    const int64_t *d_nc = ctx->d_small;
    if (nc >= 0) {
      fill_u64(ctx, ctx->d_small + 7, (uint64_t)nc, 1);
      d_nc = ctx->d_small + 7;
    }
    int32_t *cdeg = (int32_t *)ctx->pool.alloc(sizeof(int32_t) * ncb);
    HashContract hc_local;
    HashContract &hc = pre ? *pre : hc_local;
    try {
      if (hashed)
        hash_prepare_impl<OffT>(ctx, L, match, cmap, ncb, d_nc, rep, segoff, cdeg, L.maxdeg < 0 ? -1 : 2 * L.maxdeg,
                                hc);
      const int64_t tiles = (ncb + SMALL_TILE - 1) / SMALL_TILE;
      const uint32_t ep = lookback_begin(ctx, tiles);
      const double small_bytes = 8.0 * m + 12.0 * n + (double)sizeof(OffT) * (n + 1) + 8.0 * ncb +
                                 (double)sizeof(OffT) * (ncb + 1);
      // hashed rows copied here when the emit pass is not needed
      const bool tile_copy = hashed && hc.seg_max >= 0 && hc.seg_max <= TILE_COPY_SEG;
      // the totals come back with this kernel (the emit pass only copies rows)
      ctx->h_small[6] = 0;
      const Readback rb = readback_begin(ctx, 0, 5, hashed ? hash_overflow_flag(hc) : nullptr);
      launch(ctx, "contract_small", small_bytes,
             tile_copy ? k_contract_small<true, OffT, true> : k_contract_small<true, OffT, false>, dim3((unsigned)tiles),
             dim3(SMALL_TILE),
             0, d_nc, segoff, rep, match, xadj, L.adjncy, L.adjwgt, cmap, (const int32_t *)cdeg,
             tile_copy ? (const int32_t *)hc.tc : nullptr, tile_copy ? (const int32_t *)hc.tw : nullptr, cxadj,
             cadjncy, cadjwgt, ctx->d_status, ctx->d_ctr, ctx->d_small + 1, ep, rb);
      if (hashed) hash_emit_impl<OffT>(ctx, segoff, cdeg, cxadj, cadjncy, cadjwgt, hc);
      readback_end(ctx, rb);
    } catch (...) {
      hash_release(ctx, hc);
      ctx->pool.release(cdeg);
      throw;
    }
    if (nc < 0) nc = ctx->h_small[0];
    if (ncp) *ncp = nc;
    const int64_t mc = ctx->h_small[1];
    const int64_t md = ctx->h_small[4];
    const bool overflow = hashed && (uint32_t)ctx->h_small[6] != 0;
    hash_release(ctx, hc);
    ctx->pool.release(cdeg);
    if (overflow)  // a hub partition held more distinct keys than its table
      return contract_impl<OffT>(ctx, L, match, cmap, nc, rep, segoff, row_order, cxadj, cadjncy, cadjwgt,
                                 maxdeg, ncp, true);
    if (maxdeg) *maxdeg = md;
    return mc;
  }
  // size classes of the gather segments: only segments above REG_SEG
  // entries go through the sorted path
  SegBins bins;
  bin_segments(ctx, nc, segoff, sizeof(OffT) == 8, bins);
  const int32_t *mid = bins.list + bins.off[FIRST_SORTED_BIN];
  const int64_t nmid = bins.off[SEG_LARGE_BIN] - bins.off[FIRST_SORTED_BIN];
  const int32_t *large = bins.list + bins.off[SEG_LARGE_BIN];
  const int64_t nlarge = bins.off[SEG_NBINS] - bins.off[SEG_LARGE_BIN];
  const bool sorted_path = nmid + nlarge > 0;
  int32_t *cdeg = (int32_t *)ctx->pool.alloc(sizeof(int32_t) * nc);
  uint64_t *keys = nullptr;
  uint32_t *vals = nullptr;
  int32_t *rank = nullptr, *selfpos = nullptr, *selfrank = nullptr, *list = nullptr;
  const unsigned glarge = (unsigned)std::min<int64_t>(nlarge > 0 ? nlarge : 1, 8192);
  if (sorted_path) {
    keys = (uint64_t *)ctx->pool.alloc(sizeof(uint64_t) * (m > 0 ? m : 1));
    vals = (uint32_t *)ctx->pool.alloc(sizeof(uint32_t) * (m > 0 ? m : 1));
    list = (int32_t *)ctx->pool.alloc(sizeof(int32_t) * n);
    if (fs) {
      rank = (int32_t *)ctx->pool.alloc(sizeof(int32_t) * (m > 0 ? m : 1));
      selfpos = (int32_t *)ctx->pool.alloc(sizeof(int32_t) * nc);
      selfrank = (int32_t *)ctx->pool.alloc(sizeof(int32_t) * nc);
    }
    unsigned *list_n = ctx->d_ctr + 2;
    fill_u32(ctx, list_n, 0u, 1);
    const unsigned gn = nblocks(n, 256);
    launch(ctx, "contract_select_long", 0.0, k_select_long_rows<OffT>, dim3(gn), dim3(256), 0, n, xadj,
           list, list_n);
    launch(ctx, "contract_gather", 0.0, k_gather<OffT>, dim3(gn), dim3(256), 0, n, xadj, L.adjncy,
           L.adjwgt, match, cmap, segoff, keys, vals);
    launch(ctx, "contract_gather_long", 0.0, k_gather_long<OffT>, dim3(1024), dim3(256), 0,
           (const int32_t *)list, (const unsigned *)list_n, xadj, L.adjncy, L.adjwgt, match, cmap, segoff,
           keys, vals);
    sort_binned(ctx, bins, segoff, sizeof(OffT) == 8, m, keys, vals, FIRST_SORTED_BIN);
#define CM_COUNT(FSV)                                                                                     \
  if (nmid)                                                                                              \
    launch(ctx, "contract_count_mid", 0.0, k_count_small<FSV, OffT>, dim3(nblocks(nmid, 256)), dim3(256), \
           0, mid, nmid, segoff, (const uint64_t *)keys, rank, cdeg, selfpos, selfrank);                  \
  if (nlarge)                                                                                            \
    launch(ctx, "contract_count_large", 0.0, k_count_large<FSV, OffT>, dim3(glarge), dim3(256), 0, large, \
           nlarge, segoff, (const uint64_t *)keys, rank, cdeg, selfpos, selfrank);
    if (fs) {
      CM_COUNT(true)
    } else {
      CM_COUNT(false)
    }
#undef CM_COUNT
  }
  // fused short-segment merge + look-back cxadj + direct row writes
  fill_u64(ctx, ctx->d_small + 7, (uint64_t)nc, 1);
  const int64_t tiles = (nc + SMALL_TILE - 1) / SMALL_TILE;
  const uint32_t ep = lookback_begin(ctx, tiles);
  const double small_bytes = 8.0 * m + 12.0 * n + (double)sizeof(OffT) * (n + 1) + 8.0 * nc +
                             (double)sizeof(OffT) * (nc + 1);
  const Readback rb = readback_begin(ctx, 1, 4);
  if (fs)
    launch(ctx, "contract_small", small_bytes, k_contract_small<true, OffT>, dim3((unsigned)tiles), dim3(SMALL_TILE),
           0, (const int64_t *)(ctx->d_small + 7), segoff, rep, match, xadj, L.adjncy, L.adjwgt, cmap,
           (const int32_t *)cdeg, (const int32_t *)nullptr, (const int32_t *)nullptr, cxadj, cadjncy, cadjwgt,
           ctx->d_status, ctx->d_ctr, ctx->d_small + 1, ep, rb);
  else
    launch(ctx, "contract_small", small_bytes, k_contract_small<false, OffT>, dim3((unsigned)tiles),
           dim3(SMALL_TILE), 0, (const int64_t *)(ctx->d_small + 7), segoff, rep, match, xadj, L.adjncy,
           L.adjwgt, cmap, (const int32_t *)cdeg, (const int32_t *)nullptr, (const int32_t *)nullptr, cxadj,
           cadjncy, cadjwgt, ctx->d_status, ctx->d_ctr, ctx->d_small + 1, ep, rb);
  if (sorted_path) {
#define CM_EMIT(FSV)                                                                                      \
  if (nmid)                                                                                              \
    launch(ctx, "contract_emit_mid", 0.0, k_emit_small<FSV, OffT>, dim3(nblocks(nmid, 256)), dim3(256), 0, \
           mid, nmid, segoff, (const uint64_t *)keys, (const uint32_t *)vals, (const int32_t *)rank,       \
           (const int32_t *)selfpos, (const int32_t *)selfrank, (const OffT *)cxadj, cadjncy, cadjwgt);    \
  if (nlarge)                                                                                            \
    launch(ctx, "contract_emit_large", 0.0, k_emit_large<FSV, OffT>, dim3(glarge), dim3(256), 0, large,   \
           nlarge, segoff, (const uint64_t *)keys, (const uint32_t *)vals, (const int32_t *)rank,          \
           (const int32_t *)selfpos, (const int32_t *)selfrank, (const OffT *)cxadj, cadjncy, cadjwgt);
    if (fs) {
      CM_EMIT(true)
    } else {
      CM_EMIT(false)
    }
#undef CM_EMIT
  }
  readback_end(ctx, rb);
  const int64_t mc = ctx->h_small[1];
  if (maxdeg) *maxdeg = ctx->h_small[4];
  release_bins(ctx, bins);
  ctx->pool.release(cdeg);
  if (sorted_path) {
    ctx->pool.release(keys);
    ctx->pool.release(vals);
    ctx->pool.release(list);
    if (fs) {
      ctx->pool.release(rank);
      ctx->pool.release(selfpos);
      ctx->pool.release(selfrank);
    }
  }
  return mc;
}

int64_t contract(cm_ctx *ctx, const Level &L, const int32_t *d_match, const int32_t *d_cmap,
                 int64_t nc, const int32_t *d_rep, const void *d_segoff, int row_order,
                 void *d_cxadj, int32_t *d_cadjncy, int32_t *d_cadjwgt, int64_t *maxdeg, int64_t *ncp,
                 HashContract *pre)
{
  // a level on the shared array of ones (never written) has unit edge
  // weights: the kernels take no array then (edge_w) and stream one array
  // less
  Level W = L;
  if (W.m > 0 && W.adjwgt == ctx->ones) W.adjwgt = nullptr;
  return W.off64 ? contract_impl(ctx, W, d_match, d_cmap, nc, d_rep, (const int64_t *)d_segoff, row_order,
                                 (int64_t *)d_cxadj, d_cadjncy, d_cadjwgt, maxdeg, ncp, false, pre)
                 : contract_impl(ctx, W, d_match, d_cmap, nc, d_rep, (const int32_t *)d_segoff, row_order,
                                 (int32_t *)d_cxadj, d_cadjncy, d_cadjwgt, maxdeg, ncp, false, pre);
}

// --------------------------------------------------------- (d) projection
// where_f[i] = where_c[cmap[i]] (cu:3067-3075).
__global__ __launch_bounds__(256) void k_project(int64_t n, const int32_t *__restrict__ cmap,
                                                 const int32_t *__restrict__ where_c,
                                                 int32_t *__restrict__ where_f)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) where_f[i] = where_c[cmap[i]];
}

void project(cm_ctx *ctx, int64_t n, const int32_t *d_cmap, const int32_t *d_where_c,
             int32_t *d_where_f)
{
  launch(ctx, "project", 12.0 * n, k_project, dim3(nblocks(n, 256)), dim3(256), 0, n, d_cmap,
         d_where_c, d_where_f);
}

}  // namespace cm
