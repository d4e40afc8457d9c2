/*
 * cm_oracle_part.c -- CPU restatement of the reference's host initial
 * partition (recursive multilevel bisection, cu:1879-3062, 3091-3108) and a
 * serialised restatement of its k-way refinement pass (cu:3244-3515).
 * TEST INFRASTRUCTURE ONLY (see cm_oracle.h).
 *
 * Random choices use glibc rand() in the reference's call order, so a caller
 * that wants the reference's stream calls orc_srand(1) first.
 */
#include "cm_oracle.h"
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef struct og {
  int n;
  int64_t m;
  int64_t *xadj;
  int *adjncy, *adjwgt, *vwgt, *label, *cmap, *where, *bndptr, *bndlist, *id, *ed;
  int pwgts[2];
  int tvwgt, mincut, nbnd;
  struct og *coarser, *finer;
} og_t;

static og_t *og_new(void)
{
  og_t *g = (og_t *)calloc(1, sizeof(og_t));
  return g;
}

static void og_free(og_t *g)
{
  if (!g) return;
  free(g->xadj); free(g->adjncy); free(g->adjwgt); free(g->vwgt); free(g->label);
  free(g->cmap); free(g->where); free(g->bndptr); free(g->bndlist); free(g->id); free(g->ed);
  free(g);
}

static void og_settvwgt(og_t *g)
{
  int i, s = 0;
  for (i = 0; i < g->n; i++) s += g->vwgt[i];
  g->tvwgt = s;
}

/* ------------------------------------------------ max-heap priority queue */
/* cu:1976-2242: binary max-heap with float keys and a locator array. */
typedef struct { float key; int val; } rkv_t;
typedef struct { long nnodes, maxnodes; rkv_t *heap; long *locator; } pq_t;

static pq_t *pq_create(long maxnodes)
{
  long i;
  pq_t *q = (pq_t *)malloc(sizeof(pq_t));
  q->nnodes = 0;
  q->maxnodes = maxnodes;
  q->heap = (rkv_t *)malloc(sizeof(rkv_t) * (maxnodes > 0 ? maxnodes : 1));
  q->locator = (long *)malloc(sizeof(long) * (maxnodes > 0 ? maxnodes : 1));
  for (i = 0; i < maxnodes; i++) q->locator[i] = -1;
  return q;
}

static void pq_free(pq_t *q) { free(q->heap); free(q->locator); free(q); }

static void pq_reset(pq_t *q)
{
  long i;
  for (i = q->nnodes - 1; i >= 0; i--) q->locator[q->heap[i].val] = -1;
  q->nnodes = 0;
}

static void pq_insert(pq_t *q, int node, int key)                /* cu:2010-2036 */
{
  long i = q->nnodes++, j;
  while (i > 0) {
    j = (i - 1) >> 1;
    if ((float)key > q->heap[j].key) {
      q->heap[i] = q->heap[j];
      q->locator[q->heap[i].val] = i;
      i = j;
    } else break;
  }
  q->heap[i].key = (float)key;
  q->heap[i].val = node;
  q->locator[node] = i;
}

static int pq_top(pq_t *q)                                       /* cu:2041-2088 */
{
  long i, j;
  int vtx, node;
  float key;
  if (q->nnodes == 0) return -1;
  q->nnodes--;
  vtx = q->heap[0].val;
  q->locator[vtx] = -1;
  if ((i = q->nnodes) > 0) {
    key = q->heap[i].key;
    node = q->heap[i].val;
    i = 0;
    while ((j = 2 * i + 1) < q->nnodes) {
      if (q->heap[j].key > key) {
        if (j + 1 < q->nnodes && q->heap[j + 1].key > q->heap[j].key) j = j + 1;
        q->heap[i] = q->heap[j];
        q->locator[q->heap[i].val] = i;
        i = j;
      } else if (j + 1 < q->nnodes && q->heap[j + 1].key > key) {
        j = j + 1;
        q->heap[i] = q->heap[j];
        q->locator[q->heap[i].val] = i;
        i = j;
      } else break;
    }
    q->heap[i].key = key;
    q->heap[i].val = node;
    q->locator[node] = i;
  }
  return vtx;
}

static void pq_sift(pq_t *q, long i, float newkey, float oldkey, int node)
{
  long j, nnodes;
  if (newkey > oldkey) {
    while (i > 0) {
      j = (i - 1) >> 1;
      if (newkey > q->heap[j].key) {
        q->heap[i] = q->heap[j];
        q->locator[q->heap[i].val] = i;
        i = j;
      } else break;
    }
  } else {
    nnodes = q->nnodes;
    while ((j = (i << 1) + 1) < nnodes) {
      if (q->heap[j].key > newkey) {
        if (j + 1 < nnodes && q->heap[j + 1].key > q->heap[j].key) j++;
        q->heap[i] = q->heap[j];
        q->locator[q->heap[i].val] = i;
        i = j;
      } else if (j + 1 < nnodes && q->heap[j + 1].key > newkey) {
        j++;
        q->heap[i] = q->heap[j];
        q->locator[q->heap[i].val] = i;
        i = j;
      } else break;
    }
  }
  q->heap[i].key = newkey;
  q->heap[i].val = node;
  q->locator[node] = i;
}

static void pq_delete(pq_t *q, int node)                         /* cu:2093-2152 */
{
  long i = q->locator[node];
  q->locator[node] = -1;
  if (--q->nnodes > 0 && q->heap[q->nnodes].val != node) {
    int last = q->heap[q->nnodes].val;
    float newkey = q->heap[q->nnodes].key, oldkey = q->heap[i].key;
    pq_sift(q, i, newkey, oldkey, last);
  }
}

static void pq_update(pq_t *q, int node, int newkey)            /* cu:2157-2202 */
{
  long i = q->locator[node];
  pq_sift(q, i, (float)newkey, q->heap[i].key, node);
}

/* ------------------------------------------------ boundary list helpers */
#define LISTINSERT(n, list, lptr, i) do { list[n] = i; lptr[i] = (n)++; } while (0)
#define LISTDELETE(n, list, lptr, i) do { list[lptr[i]] = list[--(n)]; lptr[list[n]] = lptr[i]; lptr[i] = -1; } while (0)

typedef struct {
  float ub;            /* cu:390 ubfactors[0] = 1.03 */
  int niter;           /* cu:377 iteration_num = 10 */
  int threshold;       /* cu:378 Coarsen_threshold = 200 */
  int maxvwgt;
  float part_balance[2];
} padmin_t;

static void alloc2way(og_t *g)                                   /* cu:1880-1890 */
{
  int n = g->n > 0 ? g->n : 1;
  free(g->where); free(g->bndptr); free(g->bndlist); free(g->id); free(g->ed);
  g->where = (int *)malloc(sizeof(int) * n);
  g->bndptr = (int *)malloc(sizeof(int) * n);
  g->bndlist = (int *)malloc(sizeof(int) * n);
  g->id = (int *)malloc(sizeof(int) * n);
  g->ed = (int *)malloc(sizeof(int) * n);
}

static void compute2wayparam(og_t *g)                           /* cu:1895-1944 */
{
  int i, nbnd = 0, mincut = 0;
  int64_t j;
  g->pwgts[0] = g->pwgts[1] = 0;
  for (i = 0; i < g->n; i++) g->bndptr[i] = -1;
  for (i = 0; i < g->n; i++) g->pwgts[g->where[i]] += g->vwgt[i];
  for (i = 0; i < g->n; i++) {
    int tid = 0, ted = 0, me = g->where[i];
    for (j = g->xadj[i]; j < g->xadj[i + 1]; j++) {
      if (me == g->where[g->adjncy[j]]) tid += g->adjwgt[j];
      else ted += g->adjwgt[j];
    }
    g->id[i] = tid;
    g->ed[i] = ted;
    if (ted > 0 || g->xadj[i] == g->xadj[i + 1]) {
      LISTINSERT(nbnd, g->bndlist, g->bndptr, i);
      mincut += ted;
    }
  }
  g->mincut = mincut / 2;
  g->nbnd = nbnd;
}

static float imbal2(og_t *g, padmin_t *ad)                      /* cu:1953-1970 */
{
  int j;
  float max = -1.0f, cur;
  for (j = 0; j < 2; j++) {
    cur = g->pwgts[j] * ad->part_balance[j] - ad->ub;
    if (cur > max) max = cur;
  }
  return max;
}

static void bnd_2way_bal(og_t *g, float *ntpwgts)               /* cu:2247-2362 */
{
  int i, ii, k, kwgt, nbnd, nswaps, from, to, t, higain, mincut, mindiff;
  int64_t j;
  int n = g->n;
  int tp[2];
  int *moved = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
  int *perm = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
  pq_t *q;
  tp[0] = (int)(g->tvwgt * ntpwgts[0]);
  tp[1] = g->tvwgt - tp[0];
  mindiff = abs(tp[0] - g->pwgts[0]);
  from = (g->pwgts[0] < tp[0] ? 1 : 0);
  to = (from + 1) % 2;
  q = pq_create(n);
  for (i = 0; i < n; i++) moved[i] = -1;
  nbnd = g->nbnd;
  orc_randarrayofp(nbnd, perm, nbnd / 5, 1);
  for (ii = 0; ii < nbnd; ii++) {
    int v = g->bndlist[perm[ii]];
    if (g->where[v] == from && g->vwgt[v] <= mindiff) pq_insert(q, v, g->ed[v] - g->id[v]);
  }
  mincut = g->mincut;
  for (nswaps = 0; nswaps < n; nswaps++) {
    if ((higain = pq_top(q)) == -1) break;
    if (g->pwgts[to] + g->vwgt[higain] > tp[to]) break;
    mincut -= (g->ed[higain] - g->id[higain]);
    g->pwgts[to] += g->vwgt[higain];
    g->pwgts[from] -= g->vwgt[higain];
    g->where[higain] = to;
    moved[higain] = nswaps;
    t = g->id[higain]; g->id[higain] = g->ed[higain]; g->ed[higain] = t;
    if (g->ed[higain] == 0 && g->xadj[higain] < g->xadj[higain + 1])
      LISTDELETE(nbnd, g->bndlist, g->bndptr, higain);
    for (j = g->xadj[higain]; j < g->xadj[higain + 1]; j++) {
      k = g->adjncy[j];
      kwgt = (to == g->where[k] ? g->adjwgt[j] : -g->adjwgt[j]);
      g->id[k] += kwgt;
      g->ed[k] -= kwgt;
      if (g->bndptr[k] != -1) {
        if (g->ed[k] == 0) {
          LISTDELETE(nbnd, g->bndlist, g->bndptr, k);
          if (moved[k] == -1 && g->where[k] == from && g->vwgt[k] <= mindiff) pq_delete(q, k);
        } else if (moved[k] == -1 && g->where[k] == from && g->vwgt[k] <= mindiff) {
          pq_update(q, k, g->ed[k] - g->id[k]);
        }
      } else if (g->ed[k] > 0) {
        LISTINSERT(nbnd, g->bndlist, g->bndptr, k);
        if (moved[k] == -1 && g->where[k] == from && g->vwgt[k] <= mindiff)
          pq_insert(q, k, g->ed[k] - g->id[k]);
      }
    }
  }
  g->mincut = mincut;
  g->nbnd = nbnd;
  pq_free(q);
  free(moved); free(perm);
}

static void balance2way(og_t *g, padmin_t *ad, float *ntpwgts)   /* cu:2367-2378 */
{
  if (imbal2(g, ad) <= 0) return;
  if (fabsf(ntpwgts[0] * g->tvwgt - g->pwgts[0]) < 3 * g->tvwgt / g->n) return;
  bnd_2way_bal(g, ntpwgts);
}

static void fm_2way_refine(og_t *g, float *ntpwgts, int niter)   /* cu:2367-2531 */
{
  int i, ii, k, kwgt, nbnd, nswaps, from, to, pass, limit, t;
  int higain, mincut, mindiff, origdiff, initcut, newcut, mincutorder, avgvwgt;
  int64_t j;
  int n = g->n;
  int tp[2];
  int *moved = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
  int *swaps = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
  int *perm = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
  pq_t *qs[2];
  double lim;
  tp[0] = (int)(g->tvwgt * ntpwgts[0]);
  tp[1] = g->tvwgt - tp[0];
  lim = 0.01 * n >= 15 ? 0.01 * n : 15;
  limit = (int)(lim >= 100 ? 100 : lim);
  {
    int s = g->pwgts[0] + g->pwgts[1];
    int a = s / 20, b = 2 * s / n;
    avgvwgt = a >= b ? b : a;
  }
  qs[0] = pq_create(n);
  qs[1] = pq_create(n);
  origdiff = abs(tp[0] - g->pwgts[0]);
  for (i = 0; i < n; i++) moved[i] = -1;
  for (pass = 0; pass < niter; pass++) {
    pq_reset(qs[0]);
    pq_reset(qs[1]);
    mincutorder = -1;
    newcut = mincut = initcut = g->mincut;
    mindiff = abs(tp[0] - g->pwgts[0]);
    nbnd = g->nbnd;
    orc_randarrayofp(nbnd, perm, nbnd, 1);
    for (ii = 0; ii < nbnd; ii++) {
      int v = g->bndlist[perm[ii]];
      pq_insert(qs[g->where[v]], v, g->ed[v] - g->id[v]);
    }
    for (nswaps = 0; nswaps < n; nswaps++) {
      from = (tp[0] - g->pwgts[0] < tp[1] - g->pwgts[1] ? 0 : 1);
      to = (from + 1) % 2;
      if ((higain = pq_top(qs[from])) == -1) break;
      newcut -= (g->ed[higain] - g->id[higain]);
      g->pwgts[to] += g->vwgt[higain];
      g->pwgts[from] -= g->vwgt[higain];
      if ((newcut < mincut && abs(tp[0] - g->pwgts[0]) <= origdiff + avgvwgt) ||
          (newcut == mincut && abs(tp[0] - g->pwgts[0]) < mindiff)) {
        mincut = newcut;
        mindiff = abs(tp[0] - g->pwgts[0]);
        mincutorder = nswaps;
      } else if (nswaps - mincutorder > limit) {
        newcut += (g->ed[higain] - g->id[higain]);
        g->pwgts[from] += g->vwgt[higain];
        g->pwgts[to] -= g->vwgt[higain];
        break;
      }
      g->where[higain] = to;
      moved[higain] = nswaps;
      swaps[nswaps] = higain;
      t = g->id[higain]; g->id[higain] = g->ed[higain]; g->ed[higain] = t;
      if (g->ed[higain] == 0 && g->xadj[higain] < g->xadj[higain + 1])
        LISTDELETE(nbnd, g->bndlist, g->bndptr, higain);
      for (j = g->xadj[higain]; j < g->xadj[higain + 1]; j++) {
        k = g->adjncy[j];
        kwgt = (to == g->where[k] ? g->adjwgt[j] : -g->adjwgt[j]);
        g->id[k] += kwgt;
        g->ed[k] -= kwgt;
        if (g->bndptr[k] != -1) {
          if (g->ed[k] == 0) {
            LISTDELETE(nbnd, g->bndlist, g->bndptr, k);
            if (moved[k] == -1) pq_delete(qs[g->where[k]], k);
          } else if (moved[k] == -1) {
            pq_update(qs[g->where[k]], k, g->ed[k] - g->id[k]);
          }
        } else if (g->ed[k] > 0) {
          LISTINSERT(nbnd, g->bndlist, g->bndptr, k);
          if (moved[k] == -1) pq_insert(qs[g->where[k]], k, g->ed[k] - g->id[k]);
        }
      }
    }
    for (i = 0; i < nswaps; i++) moved[swaps[i]] = -1;
    for (nswaps--; nswaps > mincutorder; nswaps--) {
      higain = swaps[nswaps];
      to = g->where[higain] = (g->where[higain] + 1) % 2;
      t = g->id[higain]; g->id[higain] = g->ed[higain]; g->ed[higain] = t;
      if (g->ed[higain] == 0 && g->bndptr[higain] != -1 && g->xadj[higain] < g->xadj[higain + 1])
        LISTDELETE(nbnd, g->bndlist, g->bndptr, higain);
      else if (g->ed[higain] > 0 && g->bndptr[higain] == -1)
        LISTINSERT(nbnd, g->bndlist, g->bndptr, higain);
      g->pwgts[to] += g->vwgt[higain];
      g->pwgts[(to + 1) % 2] -= g->vwgt[higain];
      for (j = g->xadj[higain]; j < g->xadj[higain + 1]; j++) {
        k = g->adjncy[j];
        kwgt = (to == g->where[k] ? g->adjwgt[j] : -g->adjwgt[j]);
        g->id[k] += kwgt;
        g->ed[k] -= kwgt;
        if (g->bndptr[k] != -1 && g->ed[k] == 0) LISTDELETE(nbnd, g->bndlist, g->bndptr, k);
        if (g->bndptr[k] == -1 && g->ed[k] > 0) LISTINSERT(nbnd, g->bndlist, g->bndptr, k);
      }
    }
    g->mincut = mincut;
    g->nbnd = nbnd;
    if (mincutorder <= 0 || mincut == initcut) break;
  }
  pq_free(qs[0]);
  pq_free(qs[1]);
  free(moved); free(swaps); free(perm);
}

static void growbisection(og_t *g, padmin_t *ad, float *ntpwgts, int niparts) /* cu:2536-2661 */
{
  int i, k, n = g->n, dd, nleft, first, last, pw[2], oneminpwgt, onemaxpwgt, bestcut = 0, iter;
  int64_t j;
  int *queue, *tra, *bestwhere;
  alloc2way(g);
  bestwhere = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
  queue = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
  tra = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
  onemaxpwgt = (int)(ad->ub * g->tvwgt * ntpwgts[1]);
  oneminpwgt = (int)((1.0 / ad->ub) * g->tvwgt * ntpwgts[1]);
  for (iter = 0; iter < niparts; iter++) {
    for (i = 0; i < n; i++) { g->where[i] = 1; tra[i] = 0; }
    pw[1] = g->tvwgt;
    pw[0] = 0;
    queue[0] = orc_randinrange(n);
    tra[queue[0]] = 1;
    first = 0; last = 1; nleft = n - 1; dd = 0;
    for (;;) {
      if (first == last) {
        if (nleft == 0 || dd) break;
        k = orc_randinrange(nleft);
        for (i = 0; i < n; i++) {
          if (tra[i] == 0) {
            if (k == 0) break;
            else k--;
          }
        }
        queue[0] = i;
        tra[i] = 1;
        first = 0; last = 1; nleft--;
      }
      i = queue[first++];
      if (pw[0] > 0 && pw[1] - g->vwgt[i] < oneminpwgt) { dd = 1; continue; }
      g->where[i] = 0;
      pw[0] += g->vwgt[i];
      pw[1] -= g->vwgt[i];
      if (pw[1] <= onemaxpwgt) break;
      dd = 0;
      for (j = g->xadj[i]; j < g->xadj[i + 1]; j++) {
        k = g->adjncy[j];
        if (tra[k] == 0) { queue[last++] = k; tra[k] = 1; nleft--; }
      }
    }
    compute2wayparam(g);
    balance2way(g, ad, ntpwgts);
    fm_2way_refine(g, ntpwgts, ad->niter);
    if (iter == 0 || bestcut > g->mincut) {
      bestcut = g->mincut;
      memcpy(bestwhere, g->where, sizeof(int) * n);
      if (bestcut == 0) break;
    }
  }
  g->mincut = bestcut;
  memcpy(g->where, bestwhere, sizeof(int) * n);
  free(bestwhere); free(queue); free(tra);
}

/* cu:1854-1875 cuMetis_cpu_coarsen with the T=200 admin of the initial partition */
static og_t *cpu_coarsen(og_t *g, padmin_t *ad)
{
  ad->maxvwgt = (int)(1.5 * g->tvwgt / ad->threshold);
  do {
    og_t *c = og_new();
    int nc;
    int64_t mc;
    int *match = (int *)malloc(sizeof(int) * (g->n > 0 ? g->n : 1));
    if (!g->cmap) g->cmap = (int *)malloc(sizeof(int) * (g->n > 0 ? g->n : 1));
    nc = orc_cpu_match(g->n, g->xadj, g->adjncy, g->adjwgt, g->vwgt, ad->maxvwgt, match, g->cmap, NULL);
    c->n = nc;
    c->xadj = (int64_t *)malloc(sizeof(int64_t) * (nc + 1));
    c->adjncy = (int *)malloc(sizeof(int) * (g->m > 0 ? g->m : 1));
    c->adjwgt = (int *)malloc(sizeof(int) * (g->m > 0 ? g->m : 1));
    c->vwgt = (int *)malloc(sizeof(int) * (nc > 0 ? nc : 1));
    mc = orc_cpu_contract(g->n, g->xadj, g->adjncy, g->adjwgt, g->vwgt, match, g->cmap, nc,
                          c->xadj, c->adjncy, c->adjwgt, c->vwgt);
    c->m = mc;
    c->tvwgt = g->tvwgt;  /* sum of coarse weights equals the finer sum */
    og_settvwgt(c);
    free(match);
    c->finer = g;
    g->coarser = c;
    g = c;
  } while (g->n > ad->threshold && g->n < 0.75 * g->finer->n && g->m > g->n / 2);
  return g;
}

static void project2way(og_t *g)                                 /* cu:2693-2764 */
{
  int i, nbnd = 0;
  int64_t j;
  og_t *c = g->coarser;
  alloc2way(g);
  for (i = 0; i < g->n; i++) g->bndptr[i] = -1;
  for (i = 0; i < g->n; i++) {
    int cj = g->cmap[i];
    g->where[i] = c->where[cj];
    g->cmap[i] = c->bndptr[cj];
  }
  for (i = 0; i < g->n; i++) {
    int tid = 0, ted = 0, me;
    if (g->cmap[i] == -1) {
      for (j = g->xadj[i]; j < g->xadj[i + 1]; j++) tid += g->adjwgt[j];
    } else {
      me = g->where[i];
      for (j = g->xadj[i]; j < g->xadj[i + 1]; j++) {
        if (me == g->where[g->adjncy[j]]) tid += g->adjwgt[j];
        else ted += g->adjwgt[j];
      }
    }
    g->id[i] = tid;
    g->ed[i] = ted;
    if (ted > 0 || g->xadj[i] == g->xadj[i + 1]) LISTINSERT(nbnd, g->bndlist, g->bndptr, i);
  }
  g->mincut = c->mincut;
  g->nbnd = nbnd;
  g->pwgts[0] = c->pwgts[0];
  g->pwgts[1] = c->pwgts[1];
  og_free(c);
  g->coarser = NULL;
}

static int mlevelbisect(og_t *g, padmin_t *ad, float *tpwgts)    /* cu:2806-2829 */
{
  og_t *c, *x;
  int i;
  float rev = (float)(1.0 / (g->tvwgt > 0 ? g->tvwgt : 1));   /* cu:476-477 tvwgt_reverse */
  for (i = 0; i < 2; i++) ad->part_balance[i] = rev / tpwgts[i];
  c = cpu_coarsen(g, ad);
  growbisection(c, ad, tpwgts, 5);
  /* cu:2769-2802 cuMetis_cpu_refinement */
  x = c;
  compute2wayparam(x);
  for (;;) {
    balance2way(x, ad, tpwgts);
    fm_2way_refine(x, tpwgts, ad->niter);
    if (x == g) break;
    x = x->finer;
    project2way(x);
  }
  return g->mincut;
}

static void splitgraph(og_t *g, og_t **rl, og_t **rr)            /* cu:2834-2959 */
{
  int i, p, sn[2] = {0, 0};
  int64_t j, se[2] = {0, 0};
  int *rename = (int *)malloc(sizeof(int) * (g->n > 0 ? g->n : 1));
  og_t *s[2];
  for (i = 0; i < g->n; i++) {
    p = g->where[i];
    rename[i] = sn[p]++;
    se[p] += g->xadj[i + 1] - g->xadj[i];
  }
  for (p = 0; p < 2; p++) {
    s[p] = og_new();
    s[p]->n = sn[p];
    s[p]->xadj = (int64_t *)malloc(sizeof(int64_t) * (sn[p] + 1));
    s[p]->vwgt = (int *)malloc(sizeof(int) * (sn[p] + 1));
    s[p]->adjncy = (int *)malloc(sizeof(int) * (se[p] > 0 ? se[p] : 1));
    s[p]->adjwgt = (int *)malloc(sizeof(int) * (se[p] > 0 ? se[p] : 1));
    s[p]->label = (int *)malloc(sizeof(int) * (sn[p] > 0 ? sn[p] : 1));
    s[p]->xadj[0] = 0;
    sn[p] = 0;
    se[p] = 0;
  }
  for (i = 0; i < g->n; i++) {
    p = g->where[i];
    if (g->bndptr[i] == -1) {
      for (j = g->xadj[i]; j < g->xadj[i + 1]; j++) {
        s[p]->adjncy[se[p]] = g->adjncy[j];
        s[p]->adjwgt[se[p]++] = g->adjwgt[j];
      }
    } else {
      for (j = g->xadj[i]; j < g->xadj[i + 1]; j++) {
        int k = g->adjncy[j];
        if (g->where[k] == p) {
          s[p]->adjncy[se[p]] = k;
          s[p]->adjwgt[se[p]++] = g->adjwgt[j];
        }
      }
    }
    s[p]->vwgt[sn[p]] = g->vwgt[i];
    s[p]->label[sn[p]] = g->label[i];
    s[p]->xadj[++sn[p]] = se[p];
  }
  for (p = 0; p < 2; p++) {
    for (j = 0; j < se[p]; j++) s[p]->adjncy[j] = rename[s[p]->adjncy[j]];
    s[p]->m = se[p];
    og_settvwgt(s[p]);
  }
  free(rename);
  *rl = s[0];
  *rr = s[1];
}

static float fsum(int n, const float *a)
{
  int i;
  float s = 0;
  for (i = 0; i < n; i++) s += a[i];
  return s;
}

static int mlevel_rb(og_t *g, padmin_t *ad, int nparts, int *part, float *tpwgts, int fpart) /* cu:2964-3025 */
{
  int i, objval;
  og_t *l = NULL, *r = NULL;
  float tp2[2], wsum;
  if (g->n == 0) { og_free(g); return 0; }
  tp2[0] = fsum(nparts >> 1, tpwgts);
  tp2[1] = 1.0f - tp2[0];
  objval = mlevelbisect(g, ad, tp2);
  for (i = 0; i < g->n; i++) part[g->label[i]] = g->where[i] + fpart;
  if (nparts > 2) splitgraph(g, &l, &r);
  og_free(g);
  wsum = fsum(nparts >> 1, tpwgts);
  {
    float f = (float)(1.0 / wsum);
    for (i = 0; i < (nparts >> 1); i++) tpwgts[i] *= f;
    f = (float)(1.0 / (1.0 - wsum));
    for (i = nparts >> 1; i < nparts; i++) tpwgts[i] *= f;
  }
  if (nparts > 3) {
    objval += mlevel_rb(l, ad, nparts >> 1, part, tpwgts, fpart);
    objval += mlevel_rb(r, ad, nparts - (nparts >> 1), part, tpwgts + (nparts >> 1), fpart + (nparts >> 1));
  } else if (nparts == 3) {
    og_free(l);
    objval += mlevel_rb(r, ad, nparts - (nparts >> 1), part, tpwgts + (nparts >> 1), fpart + (nparts >> 1));
  }
  return objval;
}

/* cu:3046-3062 cuMetis_rbbisection (called from cu:3091-3108).  Returns the
 * summed bisection cuts; where[] receives the k-way assignment. */
int orc_initpart(int n, const int64_t *xadj, const int *adjncy, const int *adjwgt,
                 const int *vwgt, int nparts, int *where)
{
  int i;
  padmin_t ad;
  og_t *g = og_new();
  float *tpwgts = (float *)malloc(sizeof(float) * nparts);
  int64_t m = xadj[n];
  ad.ub = 1.03f;
  ad.niter = 10;
  ad.threshold = 200;
  ad.maxvwgt = 0;
  for (i = 0; i < nparts; i++) tpwgts[i] = (float)(1.0 / nparts);
  g->n = n;
  g->m = m;
  g->xadj = (int64_t *)malloc(sizeof(int64_t) * (n + 1));
  memcpy(g->xadj, xadj, sizeof(int64_t) * (n + 1));
  g->adjncy = (int *)malloc(sizeof(int) * (m > 0 ? m : 1));
  memcpy(g->adjncy, adjncy, sizeof(int) * m);
  g->adjwgt = (int *)malloc(sizeof(int) * (m > 0 ? m : 1));
  memcpy(g->adjwgt, adjwgt, sizeof(int) * m);
  g->vwgt = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
  memcpy(g->vwgt, vwgt, sizeof(int) * n);
  g->label = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
  for (i = 0; i < n; i++) g->label[i] = i;
  og_settvwgt(g);
  i = mlevel_rb(g, &ad, nparts, where, tpwgts, 0);
  free(tpwgts);
  return i;
}

/* Serialised restatement of one k-way refinement pass (cu:3244-3515):
 *  - targets tpwgts = 1/k (float), max = int(tp*tvwgt*1.03f), min = int(tp*tvwgt/1.03f)
 *    (cu:3401-3412);
 *  - candidates: vertices whose external weight exceeds internal (cu:3263),
 *    destination = first part with the strictly largest connectivity, own part
 *    when none beats it (cu:3292-3302);
 *  - moves: the reference runs one thread per destination part with atomics
 *    (cu:3417-3488); here destinations are taken in ascending order and, within
 *    one destination, candidates in ascending vertex id, first for from<=to,
 *    then for from>to. */
void orc_kway_refine(int n, const int64_t *xadj, const int *adjncy, const int *adjwgt,
                     const int *vwgt, int nparts, int *where)
{
  int i, p, pass;
  int64_t j, tvwgt = 0;
  int64_t *pw = (int64_t *)calloc(nparts, sizeof(int64_t));
  int *maxw = (int *)malloc(sizeof(int) * nparts), *minw = (int *)malloc(sizeof(int) * nparts);
  int *cand = (int *)malloc(sizeof(int) * (n > 0 ? n : 1)), *to = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
  int64_t *conn = (int64_t *)calloc(nparts, sizeof(int64_t));
  int *cnt = (int *)calloc(nparts + 1, sizeof(int));
  int *order = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
  int nc = 0;
  float tp = (float)(1.0 / nparts);
  for (i = 0; i < n; i++) { tvwgt += vwgt[i]; pw[where[i]] += vwgt[i]; }
  for (p = 0; p < nparts; p++) {
    maxw[p] = (int)(tp * (float)tvwgt * 1.03f);
    minw[p] = (int)(tp * (float)tvwgt / 1.03f);
  }
  for (i = 0; i < n; i++) {
    int64_t me = 0, oth = 0;
    for (j = xadj[i]; j < xadj[i + 1]; j++) {
      if (where[adjncy[j]] == where[i]) me += adjwgt[j];
      else oth += adjwgt[j];
    }
    if (oth > me) cand[nc++] = i;
  }
  for (i = 0; i < nc; i++) {
    int v = cand[i], best = where[v];
    int64_t bw;
    for (j = xadj[v]; j < xadj[v + 1]; j++) conn[where[adjncy[j]]] += adjwgt[j];
    bw = conn[best];
    for (p = 0; p < nparts; p++)
      if (conn[p] > bw) { bw = conn[p]; best = p; }
    for (j = xadj[v]; j < xadj[v + 1]; j++) conn[where[adjncy[j]]] = 0;
    to[i] = best;
    cnt[best + 1]++;
  }
  for (p = 0; p < nparts; p++) cnt[p + 1] += cnt[p];
  for (i = 0; i < nc; i++) order[cnt[to[i]]++] = i;   /* ascending vertex id per destination */
  for (p = nparts; p > 0; p--) cnt[p] = cnt[p - 1];
  cnt[0] = 0;
  for (pass = 0; pass < 2; pass++) {
    int t;
    for (t = 0; t < nparts; t++) {
      int q;
      for (q = cnt[t]; q < cnt[t + 1]; q++) {
        int v = cand[order[q]], me = where[v], vw = vwgt[v];
        if (pass == 0 ? me > t : me <= t) continue;
        if (pw[t] + vw >= minw[t] && pw[t] + vw <= maxw[t] && pw[me] - vw >= minw[me] &&
            pw[me] - vw <= maxw[me]) {
          pw[t] += vw;
          pw[me] -= vw;
          where[v] = t;
        }
      }
    }
  }
  free(pw); free(maxw); free(minw); free(cand); free(to); free(conn); free(cnt); free(order);
}

/* Balancing pass run before each level's refinement by the engine's
 * pipeline; it has no counterpart in the reference.  Up to four sweeps:
 *  - parts above maxw = int(tp*tvwgt*1.03f) are "over", parts below it have
 *    room; the fallback destination is the lightest part (lowest id);
 *  - every vertex of an over part picks the adjacent part with room and the
 *    largest connectivity (lowest id on ties), else the fallback; its gain
 *    is that connectivity (0 for the fallback) minus its own part's;
 *  - vertices are taken in order of gain descending, then vertex id, and a
 *    move is accepted while its source part is still over and the
 *    destination stays within maxw.
 * A sweep with no accepted move ends the pass. */
typedef struct { uint64_t key; int to; } bal_t;

static int bal_cmp(const void *a, const void *b)
{
  uint64_t x = ((const bal_t *)a)->key, y = ((const bal_t *)b)->key;
  return x < y ? -1 : x > y;
}

int orc_kway_balance(int n, const int64_t *xadj, const int *adjncy, const int *adjwgt,
                     const int *vwgt, int nparts, int *where)
{
  int i, p, it, total = 0;
  int64_t j, tvwgt = 0, maxw;
  int64_t *pw = (int64_t *)calloc(nparts, sizeof(int64_t));
  int64_t *conn = (int64_t *)calloc(nparts, sizeof(int64_t));
  bal_t *cand = (bal_t *)malloc(sizeof(bal_t) * (n > 0 ? n : 1));
  float tp = (float)(1.0 / nparts);
  for (i = 0; i < n; i++) { tvwgt += vwgt[i]; pw[where[i]] += vwgt[i]; }
  maxw = (int64_t)(tp * (float)tvwgt * 1.03f);
  for (it = 0; it < 4; it++) {
    int any = 0, fallback = 0, nc = 0, moved = 0;
    for (p = 0; p < nparts; p++) {
      if (pw[p] > maxw) any = 1;
      if (pw[p] < pw[fallback]) fallback = p;
    }
    if (!any) break;
    for (i = 0; i < n; i++) {
      int me = where[i], bp = -1, to;
      int64_t best = -1, gain;
      if (pw[me] <= maxw) continue;
      for (j = xadj[i]; j < xadj[i + 1]; j++) conn[where[adjncy[j]]] += adjwgt[j];
      for (p = 0; p < nparts; p++)
        if (pw[p] < maxw && conn[p] > 0 && conn[p] > best) { best = conn[p]; bp = p; }
      to = bp >= 0 ? bp : fallback;
      gain = (bp >= 0 ? best : 0) - conn[me];
      for (j = xadj[i]; j < xadj[i + 1]; j++) conn[where[adjncy[j]]] = 0;
      if (to != me) {
        if (gain > 0x7fffffffLL) gain = 0x7fffffffLL;
        if (gain < -0x7fffffffLL) gain = -0x7fffffffLL;
        cand[nc].key = ((uint64_t)(uint32_t)(0x80000000LL - gain) << 32) | (uint32_t)i;
        cand[nc].to = to;
        nc++;
      }
    }
    qsort(cand, nc, sizeof(bal_t), bal_cmp);
    for (i = 0; i < nc; i++) {
      int v = (int)(uint32_t)cand[i].key, from = where[v], t = cand[i].to;
      if (pw[from] > maxw && pw[t] + vwgt[v] <= maxw) {
        pw[from] -= vwgt[v];
        pw[t] += vwgt[v];
        where[v] = t;
        moved++;
      }
    }
    total += moved;
    if (!moved) break;
  }
  free(pw); free(conn); free(cand);
  return total;
}
