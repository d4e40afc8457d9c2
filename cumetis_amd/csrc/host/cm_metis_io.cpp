// cm_metis_io.cpp -- METIS .graph reader and partition writer with the
// reference's checks (cu:3712-3851, cu:3856-3870), returning errors instead
// of exiting, and rejecting self loops (the reference accepts them, but its
// two contractions then disagree).
#include <errno.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <stdexcept>
#include <string>
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

HostGraph read_metis(const std::string &path)
{
  FILE *fp = fopen(path.c_str(), "rb");
  if (!fp) throw std::runtime_error("Failed on file fopen(): " + path + ": " + strerror(errno));
  std::string data;
  {
    char buf[1 << 16];
    size_t r;
    while ((r = fread(buf, 1, sizeof(buf), fp)) > 0) data.append(buf, r);
    fclose(fp);
  }
  size_t pos = 0;
  auto next_line = [&](const char *&b, const char *&e) -> bool {
    while (pos < data.size()) {
      size_t nl = data.find('\n', pos);
      if (nl == std::string::npos) nl = data.size();
      b = data.data() + pos;
      e = data.data() + nl;
      pos = nl + 1;
      if (b < e && *b == '%') continue;  // comment line
      return true;
    }
    return false;
  };
  const char *b, *e;
  if (!next_line(b, e)) throw std::runtime_error("Premature end of input file: file: " + path);
  long nv = 0, ne = 0, f = 0;
  {
    std::string head(b, e);
    const int k = sscanf(head.c_str(), "%ld %ld %ld", &nv, &ne, &f);
    if (k < 2) throw std::runtime_error("The input file does not specify the number of vertices and edges.");
  }
  if (nv <= 0 || ne <= 0)
    throw std::runtime_error(fmt("The supplied nvtxs:%ld and nedges:%ld must be positive.", nv, ne));
  if (f > 111) throw std::runtime_error(fmt("Cannot read this type of file format [fmt=%ld]!", f));
  char fs[8];
  snprintf(fs, sizeof(fs), "%03ld", f % 1000);
  const bool readvw = fs[1] == '1', readew = fs[2] == '1';
  HostGraph g;
  g.n = (int)nv;
  const int64_t m2 = 2 * (int64_t)ne;
  g.xadj.assign(nv + 1, 0);
  g.adjncy.reserve(m2);
  g.adjwgt.reserve(m2);
  g.vwgt.assign(nv, 1);
  for (long i = 0; i < nv; ++i) {
    if (!next_line(b, e))
      throw std::runtime_error(fmt("Premature end of input file while reading vertex %ld.", i + 1));
    std::string line(b, e);
    const char *cur = line.c_str();
    char *nxt = nullptr;
    if (readvw) {
      const long w = strtol(cur, &nxt, 10);
      if (nxt == cur)
        throw std::runtime_error(fmt("The line for vertex %ld does not have enough weights for the %ld constraints.",
                                     i + 1, 1));
      if (w < 0) throw std::runtime_error(fmt("The weight vertex %ld and constraint %ld must be >= 0", i + 1, 0));
      g.vwgt[i] = (int)w;
      cur = nxt;
    }
    for (;;) {
      const long u = strtol(cur, &nxt, 10);
      if (nxt == cur) break;
      cur = nxt;
      if (u < 1 || u > nv) throw std::runtime_error(fmt("Edge %ld for vertex %ld is out of bounds", u, i + 1));
      if (u == i + 1) throw std::runtime_error(fmt("Vertex %ld has a self loop", i + 1));
      long w = 1;
      if (readew) {
        w = strtol(cur, &nxt, 10);
        if (nxt == cur) throw std::runtime_error(fmt("Premature end of line for vertex %ld", i + 1));
        if (w <= 0)
          throw std::runtime_error(fmt("The weight (%ld) for edge (%ld, %ld) must be positive.", w, i + 1, u));
        cur = nxt;
      }
      if ((int64_t)g.adjncy.size() == m2)
        throw std::runtime_error(fmt("There are more edges in the file than the %ld specified.", ne));
      g.adjncy.push_back((int)(u - 1));
      g.adjwgt.push_back((int)w);
    }
    g.xadj[i + 1] = (int64_t)g.adjncy.size();
  }
  if ((int64_t)g.adjncy.size() != m2)
    throw std::runtime_error(fmt("In the first line of the file, you specified that the graph contained %ld edges. "
                                 "However, I only found %ld edges in the file.",
                                 ne, (long)(g.adjncy.size() / 2)));
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

int cm_initial_partition(int64_t n, const int64_t *xadj, const int32_t *adjncy, const int32_t *adjwgt,
                         const int32_t *vwgt, int32_t nparts, int32_t *where, int64_t *objval)
{
  if (n < 0 || !xadj || (!where && n) || nparts < 1) return CM_ERR_INVALID;
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
  const int obj = cmh::initial_partition(g, nparts, w);
  for (int64_t i = 0; i < n; ++i) where[i] = w[i];
  if (objval) *objval = obj;
  return CM_OK;
}

}  // extern "C"
