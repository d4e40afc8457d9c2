// cm_cli.cpp -- `cuMetis <graph-path> <part-num>`: the reference's command
// line (cu:3874-3920).  Reads a METIS graph, partitions it with the GPU
// coarsening pipeline, prints the reference's report lines plus imbalance,
// level count and coarsening throughput, and writes <graph>.part.<k>.
#include <stdio.h>
#include <stdlib.h>

#include <string>
#include <vector>

#include "../../../include/cumetis_amd.h"

int main(int argc, char **argv)
{
  if (argc < 3) {
    fprintf(stderr, "usage: %s <graph-path> <part-num> [--no-write] [--ascending] [--seed S]\n", argv[0]);
    return 2;
  }
  const char *path = argv[1];
  const int nparts = atoi(argv[2]);
  bool write = true;
  int order = CM_ROW_FIRST_SEEN;
  unsigned long long seed = 1;
  for (int i = 3; i < argc; ++i) {
    std::string a = argv[i];
    if (a == "--no-write") write = false;
    else if (a == "--ascending") order = CM_ROW_ASCENDING;
    else if (a == "--seed" && i + 1 < argc) seed = strtoull(argv[++i], nullptr, 10);
  }
  if (nparts < 1) {
    fprintf(stderr, "part-num must be >= 1\n");
    return 2;
  }
  int64_t n = 0, m = 0;
  int64_t *xadj = nullptr;
  int32_t *adjncy = nullptr, *adjwgt = nullptr, *vwgt = nullptr;
  char err[1024] = {0};
  if (cm_read_metis(path, &n, &m, &xadj, &adjncy, &adjwgt, &vwgt, err, sizeof(err)) != CM_OK) {
    fprintf(stderr, "%s\n", err);
    return 2;
  }
  cm_ctx *ctx = nullptr;
  int rc = cm_ctx_create(0, nullptr, &ctx);
  if (rc != CM_OK) {
    fprintf(stderr, "cannot create a device context (%d)\n", rc);
    return 3;
  }
  cm_graph_view g{n, m, CM_OFF64, 0, xadj, adjncy, adjwgt, vwgt};
  cm_coarsen_params p{};
  p.nparts = nparts;
  p.row_order = order;
  p.seed = seed;
  std::vector<int32_t> part(n > 0 ? n : 1);
  cm_part_report rep{};
  rc = cm_partition(ctx, &g, &p, part.data(), &rep);
  if (rc != CM_OK) {
    fprintf(stderr, "partition failed (%d): %s\n", rc, cm_ctx_last_error(ctx));
    return 4;
  }
  std::vector<int64_t> pw(nparts, 0);
  int64_t tv = 0;
  for (int64_t i = 0; i < n; ++i) {
    pw[part[i]] += vwgt[i];
    tv += vwgt[i];
  }
  int64_t mx = 0;
  for (int k = 0; k < nparts; ++k) mx = pw[k] > mx ? pw[k] : mx;
  printf("Graph: %s \nVertex: %lld Edge: %lld\n", path, (long long)n, (long long)(m / 2));
  printf("cuMetis-Partition-end\n");
  printf("cuMetis_Partition_time= %lf     ms\n", rep.total_ms);
  printf("------Coarsen_time=          %lf     ms\n", rep.coarsen_ms);
  printf("------Init_time=             %lf     ms\n", rep.init_ms);
  printf("------Uncoarsen_time=        %lf     ms\n", rep.uncoarsen_ms);
  printf("------------------------------------------\n");
  printf("Levels=                      %d (coarsest %d vertices)\n", rep.levels, rep.coarsest_n);
  printf("Imbalance=                   %.4f\n", tv ? (double)mx * nparts / (double)tv : 0.0);
  printf("Edge-cut=                    %lld\n", (long long)rep.edgecut);
  if (write) {
    std::string out = std::string(path) + ".part." + std::to_string(nparts);
    if (cm_write_partition(out.c_str(), part.data(), n) != CM_OK) {
      fprintf(stderr, "cannot write %s\n", out.c_str());
      return 5;
    }
  }
  cm_ctx_destroy(ctx);
  cm_free_host(xadj);
  cm_free_host(adjncy);
  cm_free_host(adjwgt);
  cm_free_host(vwgt);
  return 0;
}
