// cm_metis_io.cpp -- METIS .graph reader and partition writer with the
// reference's checks (cu:3712-3851, cu:3856-3870), returning errors instead
// of exiting, and rejecting self loops (the reference accepts them, but its
// two contractions then disagree).
#include <errno.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <limits.h>

#include <algorithm>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../../include/cumetis_amd.h"
#include "cm_host.hpp"

namespace cmh {

namespace {
std::string fmt(const char *f, long a = 0, long b = 0, long c = 0)
{
  char buf[512];
  snprintf(buf, sizeof(buf), f, a, b, c);
  return buf;
}
}  // namespace

namespace {

// strtol over [cur, end): leading white space, optional sign, decimal
// digits.  Returns false (nothing consumed) when no digit follows, as strtol
// leaves its end pointer at the start then.  Lines never contain '\n', so a
// number cannot run into the next line.
inline bool next_long(const char *&cur, const char *end, long &out)
{
  const char *p = cur;
  while (p < end && (*p == ' ' || *p == '\t' || *p == '\r' || *p == '\v' || *p == '\f')) ++p;
  bool neg = false;
  if (p < end && (*p == '+' || *p == '-')) neg = *p++ == '-';
  if (p >= end || *p < '0' || *p > '9') return false;
  unsigned long v = 0;
  const unsigned long lim = (unsigned long)LONG_MAX;
  for (; p < end && *p >= '0' && *p <= '9'; ++p) {
    const unsigned long d = (unsigned long)(*p - '0');
    v = v > (lim - d) / 10 ? lim : v * 10 + d;  // strtol saturates
  }
  out = neg ? -(long)v : (long)v;
  cur = p;
  return true;
}

struct Line {
  const char *b, *e;
};

// First error of one vertex line in file order (vertex index i, 0-based), or
// an empty string.  cap: entries the graph may still take before this line
// (only binding on the line where the file exceeds the declared count).
struct LineResult {
  std::string err;
  int64_t count = 0;  // neighbours on the line (valid when err is empty)
};

LineResult scan_line(const Line &ln, long i, long nv, long ne, bool readvw, bool readew, int64_t cap,
                     int *adj, int *wgt, int *vw)
{
  LineResult r;
  const char *cur = ln.b;
  long w;
  if (readvw) {
    if (!next_long(cur, ln.e, w)) {
      r.err = fmt("The line for vertex %ld does not have enough weights for the %ld constraints.", i + 1, 1);
      return r;
    }
    if (w < 0) {
      r.err = fmt("The weight vertex %ld and constraint %ld must be >= 0", i + 1, 0);
      return r;
    }
    if (vw) *vw = (int)w;
  }
  long u;
  while (next_long(cur, ln.e, u)) {
    if (u < 1 || u > nv) {
      r.err = fmt("Edge %ld for vertex %ld is out of bounds", u, i + 1);
      return r;
    }
    if (u == i + 1) {
      r.err = fmt("Vertex %ld has a self loop", i + 1);
      return r;
    }
    w = 1;
    if (readew) {
      if (!next_long(cur, ln.e, w)) {
        r.err = fmt("Premature end of line for vertex %ld", i + 1);
        return r;
      }
      if (w <= 0) {
        r.err = fmt("The weight (%ld) for edge (%ld, %ld) must be positive.", w, i + 1, u);
        return r;
      }
    }
    if (r.count == cap) {
      r.err = fmt("There are more edges in the file than the %ld specified.", ne);
      return r;
    }
    if (adj) {
      adj[r.count] = (int)(u - 1);
      wgt[r.count] = (int)w;
    }
    ++r.count;
  }
  return r;
}

int reader_threads(size_t bytes)
{
  if (const char *e = getenv("CM_READ_THREADS")) {
    const int t = atoi(e);
    if (t >= 1) return std::min(t, 256);
  }
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  const size_t by_size = bytes / (1 << 20) + 1;  // about a MiB of text per thread at least
  return (int)std::min<size_t>({(size_t)hw, (size_t)16, by_size});
}

}  // namespace

// The reference reads line by line (cu:3712-3851).  Here the text is read
// once, cut into T chunks at line starts; each thread finds its vertex lines
// (comment lines, '%' first, are skipped as the reference does), the vertex
// numbers follow from a scan of the per-chunk line counts, a first parallel
// pass counts every line's neighbours, a scan gives xadj, and a second
// parallel pass writes the rows.  Errors are those of the sequential reader:
// the first in file order wins.
HostGraph read_metis(const std::string &path)
{
  FILE *fp = fopen(path.c_str(), "rb");
  if (!fp) throw std::runtime_error("Failed on file fopen(): " + path + ": " + strerror(errno));
  std::string data;
  {
    if (fseek(fp, 0, SEEK_END) == 0) {
      const long sz = ftell(fp);
      if (sz > 0) data.reserve((size_t)sz);
      fseek(fp, 0, SEEK_SET);
    }
    char buf[1 << 16];
    size_t r;
    while ((r = fread(buf, 1, sizeof(buf), fp)) > 0) data.append(buf, r);
    fclose(fp);
  }
  const char *const D = data.data(), *const DE = D + data.size();
  // header: the first line that is not a comment
  const char *p = D;
  Line head{nullptr, nullptr};
  while (p < DE) {
    const char *nl = (const char *)memchr(p, '\n', DE - p);
    const char *e = nl ? nl : DE;
    if (!(p < e && *p == '%')) {
      head = Line{p, e};
      p = nl ? nl + 1 : DE;
      break;
    }
    p = nl ? nl + 1 : DE;
  }
  if (!head.b) throw std::runtime_error("Premature end of input file: file: " + path);
  long nv = 0, ne = 0, f = 0;
  {
    const char *c = head.b;
    int k = 0;
    if (next_long(c, head.e, nv)) {
      ++k;
      if (next_long(c, head.e, ne)) {
        ++k;
        if (next_long(c, head.e, f)) ++k;
      }
    }
    if (k < 2) throw std::runtime_error("The input file does not specify the number of vertices and edges.");
  }
  if (nv <= 0 || ne <= 0)
    throw std::runtime_error(fmt("The supplied nvtxs:%ld and nedges:%ld must be positive.", nv, ne));
  if (f > 111) throw std::runtime_error(fmt("Cannot read this type of file format [fmt=%ld]!", f));
  char fs[8];
  snprintf(fs, sizeof(fs), "%03ld", f % 1000);
  const bool readvw = fs[1] == '1', readew = fs[2] == '1';
  const int64_t m2 = 2 * (int64_t)ne;

  // chunks of the body, each starting at a line start
  const char *body = p;
  const int T = reader_threads((size_t)(DE - body));
  std::vector<const char *> cut(T + 1);
  cut[0] = body;
  cut[T] = DE;
  for (int t = 1; t < T; ++t) {
    const char *q = body + (size_t)(DE - body) * t / T;
    if (q < cut[t - 1]) q = cut[t - 1];
    const char *nl = q > body ? (const char *)memchr(q - 1, '\n', DE - (q - 1)) : q - 1;
    cut[t] = q == body ? body : (nl ? nl + 1 : DE);
  }
  auto run = [&](auto &&fn) {
    std::vector<std::thread> th;
    for (int t = 1; t < T; ++t) th.emplace_back(fn, t);
    fn(0);
    for (auto &x : th) x.join();
  };
  // pass 0: the vertex lines of each chunk
  std::vector<std::vector<Line>> lines(T);
  run([&](int t) {
    const char *q = cut[t];
    while (q < cut[t + 1]) {
      const char *nl = (const char *)memchr(q, '\n', cut[t + 1] - q);
      const char *e = nl ? nl : cut[t + 1];
      if (!(q < e && *q == '%')) lines[t].push_back(Line{q, e});
      q = nl ? nl + 1 : cut[t + 1];
    }
  });
  std::vector<int64_t> first(T + 1, 0);
  for (int t = 0; t < T; ++t) first[t + 1] = first[t] + (int64_t)lines[t].size();
  const int64_t nlines = std::min<int64_t>(first[T], nv);  // lines past the last vertex are ignored

  HostGraph g;
  g.n = (int)nv;
  g.xadj.assign(nv + 1, 0);
  g.vwgt.assign(nv, 1);
  // pass 1: counts, and the first error of each chunk (no capacity yet)
  std::vector<int64_t> err_at(T, INT64_MAX);
  std::vector<std::string> err_msg(T);
  run([&](int t) {
    for (size_t k = 0; k < lines[t].size(); ++k) {
      const int64_t i = first[t] + (int64_t)k;
      if (i >= nv) break;
      LineResult r = scan_line(lines[t][k], (long)i, nv, ne, readvw, readew, INT64_MAX, nullptr, nullptr,
                               &g.vwgt[i]);
      if (!r.err.empty()) {
        err_at[t] = i;
        err_msg[t] = r.err;
        break;
      }
      g.xadj[i + 1] = r.count;
    }
  });
  int64_t bad = INT64_MAX;
  std::string bad_msg;
  for (int t = 0; t < T; ++t)
    if (err_at[t] < bad) {
      bad = err_at[t];
      bad_msg = err_msg[t];
    }
  // rows before the first bad line are complete: scan their counts, and find
  // the line (if any) where the file passes the declared count
  const int64_t good = std::min<int64_t>(bad, nlines);
  int64_t over = INT64_MAX;
  for (int64_t i = 0; i < good; ++i) {
    g.xadj[i + 1] += g.xadj[i];
    if (g.xadj[i + 1] > m2) {
      over = i;
      break;
    }
  }
  // the first line that passes the declared count or holds an error, again
  // sequentially with the room left: the error the sequential reader meets
  // first on it (an overflow can come before a bad token of the same line)
  const int64_t stop = over != INT64_MAX ? over : (bad < nlines ? bad : INT64_MAX);
  if (stop != INT64_MAX) {
    int t = 0;
    while (first[t + 1] <= stop) ++t;
    LineResult r = scan_line(lines[t][stop - first[t]], (long)stop, nv, ne, readvw, readew, m2 - g.xadj[stop],
                             nullptr, nullptr, nullptr);
    throw std::runtime_error(r.err.empty() ? bad_msg : r.err);
  }
  if (nlines < nv) throw std::runtime_error(fmt("Premature end of input file while reading vertex %ld.", nlines + 1));
  // pass 2: the rows
  g.adjncy.resize(std::max<int64_t>(g.xadj[nv], 0));
  g.adjwgt.resize(std::max<int64_t>(g.xadj[nv], 0));
  run([&](int t) {
    for (size_t k = 0; k < lines[t].size(); ++k) {
      const int64_t i = first[t] + (int64_t)k;
      if (i >= nv) break;
      scan_line(lines[t][k], (long)i, nv, ne, readvw, readew, INT64_MAX, g.adjncy.data() + g.xadj[i],
                g.adjwgt.data() + g.xadj[i], nullptr);
    }
  });
  if (g.xadj[nv] != m2)
    throw std::runtime_error(fmt("In the first line of the file, you specified that the graph contained %ld edges. "
                                 "However, I only found %ld edges in the file.",
                                 ne, (long)(g.xadj[nv] / 2)));
  return g;
}

void write_partition(const std::string &path, const int *part, int n)
{
  FILE *fp = fopen(path.c_str(), "w");
  if (!fp) throw std::runtime_error("Failed on file fopen(): " + path + ": " + strerror(errno));
  std::string out;
  out.reserve((size_t)n * 4);
  char buf[16];
  for (int i = 0; i < n; ++i) {
    const int k = snprintf(buf, sizeof(buf), "%d\n", part[i]);
    out.append(buf, k);
  }
  fwrite(out.data(), 1, out.size(), fp);
  fclose(fp);
}

}  // namespace cmh

extern "C" {

int cm_read_metis(const char *path, int64_t *n, int64_t *m, int64_t **xadj, int32_t **adjncy,
                  int32_t **adjwgt, int32_t **vwgt, char *err, int errcap)
{
  if (!path || !n || !m || !xadj || !adjncy || !adjwgt || !vwgt) return CM_ERR_INVALID;
  try {
    cmh::HostGraph g = cmh::read_metis(path);
    *n = g.n;
    *m = g.m();
    *xadj = (int64_t *)malloc(sizeof(int64_t) * (g.n + 1));
    *adjncy = (int32_t *)malloc(sizeof(int32_t) * (g.m() > 0 ? g.m() : 1));
    *adjwgt = (int32_t *)malloc(sizeof(int32_t) * (g.m() > 0 ? g.m() : 1));
    *vwgt = (int32_t *)malloc(sizeof(int32_t) * (g.n > 0 ? g.n : 1));
    memcpy(*xadj, g.xadj.data(), sizeof(int64_t) * (g.n + 1));
    memcpy(*adjncy, g.adjncy.data(), sizeof(int32_t) * g.m());
    memcpy(*adjwgt, g.adjwgt.data(), sizeof(int32_t) * g.m());
    memcpy(*vwgt, g.vwgt.data(), sizeof(int32_t) * g.n);
  } catch (const std::exception &e) {
    if (err && errcap > 0) {
      strncpy(err, e.what(), errcap - 1);
      err[errcap - 1] = 0;
    }
    return CM_ERR_INPUT;
  }
  return CM_OK;
}

void cm_free_host(void *p) { free(p); }

int cm_write_partition(const char *path, const int32_t *part, int64_t n)
{
  if (!path || (!part && n)) return CM_ERR_INVALID;
  try {
    cmh::write_partition(path, part, (int)n);
  } catch (const std::exception &) {
    return CM_ERR_INPUT;
  }
  return CM_OK;
}

static int initial_partition_call(int64_t n, const int64_t *xadj, const int32_t *adjncy, const int32_t *adjwgt,
                                  const int32_t *vwgt, int32_t nparts, int32_t *where, int64_t *objval,
                                  cmh::InitForm form)
{
  if (n < 0 || n >= INT32_MAX || !xadj || (!where && n) || nparts < 1) return CM_ERR_INVALID;
  cmh::HostGraph g;
  g.n = (int)n;
  g.xadj.assign(xadj, xadj + n + 1);
  const int64_t m = xadj[n];
  g.adjncy.assign(adjncy, adjncy + m);
  if (adjwgt) g.adjwgt.assign(adjwgt, adjwgt + m);
  else g.adjwgt.assign(m, 1);
  if (vwgt) g.vwgt.assign(vwgt, vwgt + n);
  else g.vwgt.assign(n, 1);
  std::vector<int> w;
  const int obj = cmh::initial_partition(g, nparts, w, form);
  for (int64_t i = 0; i < n; ++i) where[i] = w[i];
  if (objval) *objval = obj;
  return CM_OK;
}

int cm_initial_partition(int64_t n, const int64_t *xadj, const int32_t *adjncy, const int32_t *adjwgt,
                         const int32_t *vwgt, int32_t nparts, int32_t *where, int64_t *objval)
{
  return initial_partition_call(n, xadj, adjncy, adjwgt, vwgt, nparts, where, objval, cmh::InitForm::EXACT);
}

int cm_initial_partition_par(int64_t n, const int64_t *xadj, const int32_t *adjncy, const int32_t *adjwgt,
                             const int32_t *vwgt, int32_t nparts, int32_t *where, int64_t *objval)
{
  return initial_partition_call(n, xadj, adjncy, adjwgt, vwgt, nparts, where, objval, cmh::InitForm::PARALLEL);
}

}  // extern "C"
