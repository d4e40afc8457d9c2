/*
 * cm_oracle.c -- CPU restatement of the reference's coarsening path, plus
 * the host initial partition and a serialised k-way refinement, used as the
 * checker for the HIP engine.  TEST INFRASTRUCTURE ONLY (see cm_oracle.h).
 *
 * Citations: "cu:" = cuMetis.cu of the reference.
 */
#include "cm_oracle.h"
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ rand */

/* cu:287-300: the reference draws from glibc rand() (never seeded => seed 1). */
void orc_srand(unsigned seed) { srand(seed); }
int orc_rand(void) { return (int)(uint32_t)rand(); }
int orc_randinrange(int max) { return (int)(orc_rand() % max); }

/* cu:639-671 cuMetis_int_randarrayofp */
void orc_randarrayofp(int n, int *p, int m, int flag)
{
  int i, u, v, t;
  if (flag == 1)
    for (i = 0; i < n; i++) p[i] = i;
  if (n <= 0) return;
  if (n < 10) {
    for (i = 0; i < n; i++) {
      v = orc_randinrange(n);
      u = orc_randinrange(n);
      t = p[v]; p[v] = p[u]; p[u] = t;
    }
  } else {
    for (i = 0; i < m; i++) {
      v = orc_randinrange(n - 3);
      u = orc_randinrange(n - 3);
      t = p[v + 0]; p[v + 0] = p[u + 2]; p[u + 2] = t;
      t = p[v + 1]; p[v + 1] = p[u + 3]; p[u + 3] = t;
      t = p[v + 2]; p[v + 2] = p[u + 0]; p[u + 0] = t;
      t = p[v + 3]; p[v + 3] = p[u + 1]; p[u + 1] = t;
    }
  }
}

/* ------------------------------------------------------------ thresholds */

/* cu:279-284: floor(log2(a)) */
int orc_log2(int a)
{
  int i;
  for (i = 1; a > 1; i++, a = a >> 1) {}
  return i - 1;
}

/* cu:3634: Coarsen_threshold = max(n/(20*log2 k), 30k) */
int orc_threshold(int n, int nparts)
{
  int l = orc_log2(nparts);
  int a = l > 0 ? n / (20 * l) : n;
  int b = 30 * nparts;
  return a >= b ? a : b;
}

/* cu:1828 / cu:1858: maxvwgt = (int)(1.5*tvwgt/T) */
int orc_maxvwgt(long long tvwgt, int threshold)
{
  return (int)(1.5 * (double)tvwgt / threshold);
}

/* cu:1843-1845 / cu:1870-1872: keep coarsening while this holds. */
int orc_coarsen_continue(int nc, int nf, int64_t mc, int threshold)
{
  return nc > threshold && nc < 0.75 * nf && mc > nc / 2;
}

/* --------------------------------------------------------- cpu matching */

/* cu:830-903 cuMetis_cpu_match (matching + canonical cmap); returns cnvtxs.
 * perm_out (optional) receives the visit order. */
int orc_cpu_match(int n, const int64_t *xadj, const int *adjncy, const int *adjwgt,
                  const int *vwgt, int maxvwgt, int *match, int *cmap, int *perm_out)
{
  int i, pi, k, cnvtxs = 0, maxidx, maxwgt, aved;
  int64_t j;
  int *perm = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
  int *tperm = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
  int *d = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
  int *counts;

  for (i = 0; i < n; i++) match[i] = -1;
  orc_randarrayofp(n, tperm, n / 8, 1);                           /* cu:853 */
  aved = (int)(0.7 * (double)(int)(xadj[n] / n));                 /* cu:854 */
  for (i = 0; i < n; i++) {                                       /* cu:856-858 */
    int deg = (int)(xadj[i + 1] - xadj[i]);
    d[i] = deg > aved ? aved : deg;
  }
  /* cu:675-692 cuMetis_matching_sort: stable counting sort of tperm by d */
  counts = (int *)calloc((size_t)aved + 2, sizeof(int));
  for (i = 0; i < n; i++) counts[d[i]]++;
  for (i = 1; i < aved + 1; i++) counts[i] += counts[i - 1];
  for (i = aved + 1; i > 0; i--) counts[i] = counts[i - 1];
  counts[0] = 0;
  for (pi = 0; pi < n; pi++) {
    i = tperm[pi];
    perm[counts[d[i]]++] = i;
  }
  free(counts);

  for (pi = 0; pi < n; pi++) {                                    /* cu:862-888 */
    i = perm[pi];
    if (match[i] != -1) continue;
    maxidx = i;
    maxwgt = -1;
    for (j = xadj[i]; j < xadj[i + 1]; j++) {
      k = adjncy[j];
      if (match[k] == -1 && maxwgt < adjwgt[j] && vwgt[i] + vwgt[k] <= maxvwgt) {
        maxidx = k;
        maxwgt = adjwgt[j];
      }
      if (maxidx == i && 3 * vwgt[i] < maxvwgt) maxidx = -1;
    }
    if (maxidx != -1) {
      cmap[i] = cmap[maxidx] = cnvtxs++;
      match[i] = maxidx;
      match[maxidx] = i;
    }
  }
  for (cnvtxs = 0, i = 0; i < n; i++) {                           /* cu:892-902 */
    if (match[i] == -1) {
      match[i] = i;
      cmap[i] = cnvtxs++;
    } else if (i <= match[i]) {
      cmap[i] = cmap[match[i]] = cnvtxs++;
    }
  }
  if (perm_out) memcpy(perm_out, perm, sizeof(int) * n);
  free(perm); free(tperm); free(d);
  return cnvtxs;
}

/* cu:892-902 and cu:1634-1694: cmap[i] = rank of min(i, match[i]) among
 * representatives (j is one iff j <= match[j]).  match[i]==-1 means self. */
int orc_number(int n, const int *match, int *cmap)
{
  int i, c = 0;
  for (i = 0; i < n; i++) {
    int u = match[i] < 0 ? i : match[i];
    if (i <= u) cmap[i] = c++;
  }
  for (i = 0; i < n; i++) {
    int u = match[i] < 0 ? i : match[i];
    if (u < i) cmap[i] = cmap[u];
  }
  return c;
}

/* ------------------------------------------------------ cpu contraction */

/* cu:737-825 cuMetis_cpu_create_cgraph.  Coarse row c = first-seen distinct
 * coarse neighbours of row(v) then row(u) (v = min endpoint), weights summed;
 * for a matched pair the self entry is replaced by the last entry.
 * cadjncy/cadjwgt need room for xadj[n] entries.  Returns cnedges. */
int64_t orc_cpu_contract(int n, const int64_t *xadj, const int *adjncy, const int *adjwgt,
                         const int *vwgt, const int *match, const int *cmap, int nc,
                         int64_t *cxadj, int *cadjncy, int *cadjwgt, int *cvwgt)
{
  int v, u, k;
  int64_t j, m, nedges, cnedges = 0;
  int cnvtxs = 0;
  int64_t *htable = (int64_t *)malloc(sizeof(int64_t) * (nc > 0 ? nc : 1));
  for (v = 0; v < nc; v++) htable[v] = -1;
  cxadj[0] = 0;
  for (v = 0; v < n; v++) {
    int *ca = cadjncy + cnedges, *cw = cadjwgt + cnedges;
    if ((u = match[v]) < v) continue;
    cvwgt[cnvtxs] = vwgt[v];
    nedges = 0;
    for (j = xadj[v]; j < xadj[v + 1]; j++) {
      k = cmap[adjncy[j]];
      if ((m = htable[k]) == -1) {
        ca[nedges] = k; cw[nedges] = adjwgt[j]; htable[k] = nedges++;
      } else {
        cw[m] += adjwgt[j];
      }
    }
    if (v != u) {
      cvwgt[cnvtxs] += vwgt[u];
      for (j = xadj[u]; j < xadj[u + 1]; j++) {
        k = cmap[adjncy[j]];
        if ((m = htable[k]) == -1) {
          ca[nedges] = k; cw[nedges] = adjwgt[j]; htable[k] = nedges++;
        } else {
          cw[m] += adjwgt[j];
        }
      }
      if ((j = htable[cnvtxs]) != -1) {                              /* cu:804-808 */
        ca[j] = ca[--nedges];
        cw[j] = cw[nedges];
        htable[cnvtxs] = -1;
      }
    }
    for (j = 0; j < nedges; j++) htable[ca[j]] = -1;
    cnedges += nedges;
    cxadj[++cnvtxs] = cnedges;
  }
  free(htable);
  return cnedges;
}

/* -------------------------------------------- handshake matching model */

/* splitmix64 finaliser (Steele, Lea, Flood 2014) -> 32 bits. */
static uint64_t orc_mix64(uint64_t z)
{
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

uint32_t orc_edge_hash(uint64_t seed, uint32_t a, uint32_t b)
{
  return (uint32_t)(orc_mix64(seed ^ (((uint64_t)a << 32) | b)) >> 32);
}

/* Model of the engine's seeded handshake matching (no reference counterpart;
 * the reference's own GPU matching, cu:1563-1630, is racy).  Synchronous
 * rounds over the round-start state:
 *   propose: every unmatched v keeps its previous proposal if that vertex is
 *            still unmatched, else scans its row for the maximum edge among
 *            unmatched neighbours u with vwgt[v]+vwgt[u] <= maxvwgt, ordered
 *            by (adjwgt, hash(seed,min,max), min, max); none => dead (-2).
 *   accept:  dead v => match[v]=v;  prop[prop[v]]==v => match[v]=prop[v].
 * A round that changes nothing, or max_rounds, ends the loop; the remaining
 * unmatched vertices match themselves.  Returns the number of rounds. */
int orc_handshake_match(int n, const int64_t *xadj, const int *adjncy, const int *adjwgt,
                        const int *vwgt, int maxvwgt, uint64_t seed, int max_rounds,
                        int *match)
{
  int v, round = 0;
  int *prop = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
  int *nprop = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
  for (v = 0; v < n; v++) { match[v] = -1; prop[v] = -1; }
  for (;;) {
    int active = 0, changed = 0;
    for (v = 0; v < n; v++) active += match[v] == -1;
    if (active == 0 || round >= max_rounds) break;
    round++;
    for (v = 0; v < n; v++) {
      int p, best = -2;
      uint64_t bk1 = 0, bk2 = 0;
      int64_t j;
      nprop[v] = prop[v];
      if (match[v] != -1) continue;
      p = prop[v];
      if (p >= 0 && match[p] == -1) continue;
      for (j = xadj[v]; j < xadj[v + 1]; j++) {
        int u = adjncy[j];
        uint32_t a, b;
        uint64_t k1, k2;
        if (u == v || match[u] != -1 || vwgt[v] + vwgt[u] > maxvwgt) continue;
        a = (uint32_t)(u < v ? u : v);
        b = (uint32_t)(u < v ? v : u);
        k1 = ((uint64_t)(uint32_t)adjwgt[j] << 32) | orc_edge_hash(seed, a, b);
        k2 = ((uint64_t)a << 32) | b;
        if (best == -2 || k1 > bk1 || (k1 == bk1 && k2 > bk2)) {
          best = u; bk1 = k1; bk2 = k2;
        }
      }
      nprop[v] = best;
    }
    memcpy(prop, nprop, sizeof(int) * n);
    for (v = 0; v < n; v++) {
      int p;
      if (match[v] != -1) continue;
      p = prop[v];
      if (p == -2) { match[v] = v; changed++; }
      else if (p >= 0 && prop[p] == v) { match[v] = p; changed++; }
    }
    if (changed == 0) break;
  }
  for (v = 0; v < n; v++)
    if (match[v] == -1) match[v] = v;
  free(prop); free(nprop);
  return round;
}

/* ------------------------------------------------------ projection, cut */

/* cu:3067-3075 projectback */
void orc_project(int n, const int *cmap, const int *where_c, int *where_f)
{
  int i;
  for (i = 0; i < n; i++) where_f[i] = where_c[cmap[i]];
}

/* cu:356-366 cuMetis_computecut */
int64_t orc_edgecut(int n, const int64_t *xadj, const int *adjncy, const int *adjwgt,
                    const int *where)
{
  int i;
  int64_t j, cut = 0;
  for (i = 0; i < n; i++)
    for (j = xadj[i]; j < xadj[i + 1]; j++)
      if (where[i] != where[adjncy[j]]) cut += adjwgt[j];
  return cut / 2;
}

void orc_pwgts(int n, const int *vwgt, const int *where, int nparts, int64_t *pwgts)
{
  int i;
  for (i = 0; i < nparts; i++) pwgts[i] = 0;
  for (i = 0; i < n; i++) pwgts[where[i]] += vwgt[i];
}

/* ------------------------------------------------------------- segsort */

static int orc_cmp_kv(const void *a, const void *b)
{
  const uint64_t *x = (const uint64_t *)a, *y = (const uint64_t *)b;
  if (x[0] != y[0]) return x[0] < y[0] ? -1 : 1;
  return x[1] < y[1] ? -1 : (x[1] > y[1]);
}

/* Per-segment ascending sort of 64-bit keys carrying 32-bit values (the
 * bb_segsort role, include/bb_segsort.h:35); ties ordered by value. */
void orc_segsort_u64(int64_t nseg, const int64_t *segoff, uint64_t *keys, uint32_t *vals)
{
  int64_t s, i;
  for (s = 0; s < nseg; s++) {
    int64_t a = segoff[s], len = segoff[s + 1] - a;
    uint64_t *tmp;
    if (len < 2) continue;
    tmp = (uint64_t *)malloc(sizeof(uint64_t) * 2 * len);
    for (i = 0; i < len; i++) { tmp[2 * i] = keys[a + i]; tmp[2 * i + 1] = vals[a + i]; }
    qsort(tmp, (size_t)len, 2 * sizeof(uint64_t), orc_cmp_kv);
    for (i = 0; i < len; i++) { keys[a + i] = tmp[2 * i]; vals[a + i] = (uint32_t)tmp[2 * i + 1]; }
    free(tmp);
  }
}
