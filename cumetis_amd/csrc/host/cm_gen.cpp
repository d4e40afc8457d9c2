// cm_gen.cpp -- seeded synthetic graph generators for the benchmark
// configurations (multithreaded host code): R-MAT (Chakrabarti, Zhan,
// Faloutsos 2004) and 2D random geometric graphs.  Output: symmetric CSR,
// rows ascending, no self loops or duplicates; edge weights drawn once per
// undirected edge and mirrored.  The stream depends only on the seed (fixed
// chunking), not on the thread count.
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <thread>
#include <vector>

#include "../../../include/cumetis_amd.h"

namespace {

inline uint64_t mix64(uint64_t z)
{
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

struct Rng {  // splitmix64 stream
  uint64_t s;
  explicit Rng(uint64_t seed) : s(seed) {}
  uint64_t next()
  {
    s += 0x9E3779B97F4A7C15ull;
    uint64_t z = s;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
};

int nthreads()
{
  unsigned h = std::thread::hardware_concurrency();
  return (int)std::max(1u, std::min(h ? h : 1u, 16u));
}

template <typename F>
void parallel_for(int64_t n, F &&f)
{
  const int t = nthreads();
  std::vector<std::thread> th;
  const int64_t chunk = (n + t - 1) / t;
  for (int i = 0; i < t; ++i) {
    const int64_t a = i * chunk, b = std::min<int64_t>(n, a + chunk);
    if (a >= b) break;
    th.emplace_back([=, &f] { f(a, b); });
  }
  for (auto &x : th) x.join();
}

// Undirected edge list (a < b) with weights -> CSR (rows sorted, duplicates
// dropped keeping the smallest weight of each duplicate, so the result does
// not depend on list order).
void build_csr(int64_t n, const std::vector<int32_t> &ea, const std::vector<int32_t> &eb,
               const std::vector<int32_t> &ew, int64_t **xadj, int32_t **adjncy, int32_t **adjwgt)
{
  const int64_t E = (int64_t)ea.size();
  std::vector<int64_t> deg(n + 1, 0);
  for (int64_t e = 0; e < E; ++e) {
    deg[ea[e] + 1]++;
    deg[eb[e] + 1]++;
  }
  for (int64_t i = 0; i < n; ++i) deg[i + 1] += deg[i];
  std::vector<int64_t> pos(deg.begin(), deg.end() - 1);
  std::vector<uint64_t> ent(2 * E);  // (neighbour << 32 | weight)
  for (int64_t e = 0; e < E; ++e) {
    ent[pos[ea[e]]++] = ((uint64_t)(uint32_t)eb[e] << 32) | (uint32_t)ew[e];
    ent[pos[eb[e]]++] = ((uint64_t)(uint32_t)ea[e] << 32) | (uint32_t)ew[e];
  }
  std::vector<int64_t> cnt(n + 1, 0);
  parallel_for(n, [&](int64_t a, int64_t b) {
    for (int64_t i = a; i < b; ++i) {
      uint64_t *r = ent.data() + deg[i];
      const int64_t len = deg[i + 1] - deg[i];
      std::sort(r, r + len);
      int64_t k = 0;
      for (int64_t j = 0; j < len; ++j)
        if (k == 0 || (r[j] >> 32) != (r[k - 1] >> 32)) r[k++] = r[j];
      cnt[i + 1] = k;
    }
  });
  for (int64_t i = 0; i < n; ++i) cnt[i + 1] += cnt[i];
  *xadj = (int64_t *)malloc(sizeof(int64_t) * (n + 1));
  *adjncy = (int32_t *)malloc(sizeof(int32_t) * std::max<int64_t>(cnt[n], 1));
  *adjwgt = (int32_t *)malloc(sizeof(int32_t) * std::max<int64_t>(cnt[n], 1));
  memcpy(*xadj, cnt.data(), sizeof(int64_t) * (n + 1));
  parallel_for(n, [&](int64_t a, int64_t b) {
    for (int64_t i = a; i < b; ++i) {
      const uint64_t *r = ent.data() + deg[i];
      for (int64_t j = 0; j < cnt[i + 1] - cnt[i]; ++j) {
        (*adjncy)[cnt[i] + j] = (int32_t)(r[j] >> 32);
        (*adjwgt)[cnt[i] + j] = (int32_t)(uint32_t)r[j];
      }
    }
  });
}

}  // namespace

extern "C" {

int cm_gen_rmat(int32_t scale, int32_t edge_factor, double a, double b, double c, uint64_t seed,
                int64_t *n, int64_t **xadj, int32_t **adjncy, int32_t **adjwgt)
{
  if (scale < 1 || scale > 30 || edge_factor < 1 || !n || !xadj || !adjncy || !adjwgt) return CM_ERR_INVALID;
  const int64_t nv = (int64_t)1 << scale, E = (int64_t)edge_factor * nv;
  const int64_t CH = 1 << 20;
  const int64_t nch = (E + CH - 1) / CH;
  std::vector<int32_t> ea(E), eb(E);
  const double ab = a + b, abc = a + b + c;
  parallel_for(nch, [&](int64_t c0, int64_t c1) {
    for (int64_t ch = c0; ch < c1; ++ch) {
      Rng r(mix64(seed ^ mix64((uint64_t)ch)));
      for (int64_t e = ch * CH; e < std::min(E, (ch + 1) * CH); ++e) {
        int64_t u = 0, v = 0;
        for (int bit = 0; bit < scale; ++bit) {
          const double x = r.uniform();
          const int down = x >= ab, right = (x >= a && x < ab) || x >= abc;
          u |= (int64_t)down << bit;
          v |= (int64_t)right << bit;
        }
        ea[e] = (int32_t)std::min(u, v);
        eb[e] = (int32_t)std::max(u, v);
      }
    }
  });
  // drop self loops
  int64_t k = 0;
  for (int64_t e = 0; e < E; ++e)
    if (ea[e] != eb[e]) {
      ea[k] = ea[e];
      eb[k++] = eb[e];
    }
  ea.resize(k);
  eb.resize(k);
  std::vector<int32_t> ew(k, 1);
  *n = nv;
  build_csr(nv, ea, eb, ew, xadj, adjncy, adjwgt);
  return CM_OK;
}

int cm_gen_rgg2d(int64_t nv, double mean_degree, int32_t wmax, uint64_t seed, int64_t **xadj,
                 int32_t **adjncy, int32_t **adjwgt)
{
  if (nv < 1 || mean_degree <= 0 || !xadj || !adjncy || !adjwgt) return CM_ERR_INVALID;
  Rng r(mix64(seed));
  std::vector<double> x(nv), y(nv);
  for (int64_t i = 0; i < nv; ++i) {
    x[i] = r.uniform();
    y[i] = r.uniform();
  }
  const double rad = std::sqrt(mean_degree / (M_PI * (double)nv));
  const int64_t g = std::max<int64_t>(1, (int64_t)(1.0 / rad));
  std::vector<int64_t> cell(g * g + 1, 0);
  auto cid = [&](int64_t i) {
    const int64_t cx = std::min<int64_t>(g - 1, (int64_t)(x[i] * g)), cy = std::min<int64_t>(g - 1, (int64_t)(y[i] * g));
    return cy * g + cx;
  };
  for (int64_t i = 0; i < nv; ++i) cell[cid(i) + 1]++;
  for (int64_t i = 0; i < g * g; ++i) cell[i + 1] += cell[i];
  std::vector<int64_t> fill(cell.begin(), cell.end() - 1);
  std::vector<int32_t> pts(nv);
  for (int64_t i = 0; i < nv; ++i) pts[fill[cid(i)]++] = (int32_t)i;
  const double r2 = rad * rad;
  std::vector<int32_t> ea, eb, ew;
  for (int64_t i = 0; i < nv; ++i) {
    const int64_t c0 = cid(i), cx = c0 % g, cy = c0 / g;
    for (int64_t dy = -1; dy <= 1; ++dy)
      for (int64_t dx = -1; dx <= 1; ++dx) {
        const int64_t ux = cx + dx, uy = cy + dy;
        if (ux < 0 || uy < 0 || ux >= g || uy >= g) continue;
        const int64_t cc = uy * g + ux;
        for (int64_t q = cell[cc]; q < cell[cc + 1]; ++q) {
          const int32_t j = pts[q];
          if (j <= i) continue;
          const double ddx = x[i] - x[j], ddy = y[i] - y[j];
          if (ddx * ddx + ddy * ddy <= r2) {
            ea.push_back((int32_t)i);
            eb.push_back(j);
          }
        }
      }
  }
  ew.resize(ea.size());
  for (size_t e = 0; e < ea.size(); ++e)
    ew[e] = wmax > 0 ? 1 + (int32_t)(mix64(seed * 0x632BE59BD9B4E019ull ^ mix64(((uint64_t)ea[e] << 32) | (uint32_t)eb[e])) %
                                      (uint64_t)wmax)
                     : 1;
  build_csr(nv, ea, eb, ew, xadj, adjncy, adjwgt);
  return CM_OK;
}

}  // extern "C"
