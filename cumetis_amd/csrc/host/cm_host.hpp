// cm_host.hpp -- host-side parts of the partitioning pipeline around the
// GPU coarsening: METIS I/O, the recursive-bisection initial partition and
// the k-way refinement move scheduler.
#pragma once
#include <stdint.h>

#include <memory>
#include <string>
#include <utility>
#include <vector>

namespace cmh {

// A vector whose resize leaves new elements uninitialised: the large graph
// arrays of the initial partition are written whole after they are sized,
// and a zero fill would touch every fresh page once more, on one thread.
template <class T>
struct DefaultInit : std::allocator<T> {
  using value_type = T;
  template <class U>
  struct rebind {
    using other = DefaultInit<U>;
  };
  DefaultInit() noexcept = default;
  template <class U>
  DefaultInit(const DefaultInit<U> &) noexcept
  {
  }
  template <class U>
  void construct(U *p) noexcept(noexcept(::new((void *)p) U))
  {
    ::new ((void *)p) U;
  }
  template <class U, class... A>
  void construct(U *p, A &&...a)
  {
    ::new ((void *)p) U(std::forward<A>(a)...);
  }
};
template <class T>
using hvec = std::vector<T, DefaultInit<T>>;

struct HostGraph {
  int n = 0;
  hvec<int64_t> xadj;
  hvec<int> adjncy, adjwgt, vwgt;
  int64_t m() const { return xadj.empty() ? 0 : xadj.back(); }
};

// METIS .graph reader with the checks of cu:3712-3851; throws
// std::runtime_error with the reference's message on malformed input.
HostGraph read_metis(const std::string &path);
// "<graph>.part.<k>" writer, one part id per line (cu:3856-3870).
void write_partition(const std::string &path, const int *part, int n);

// Recursive multilevel bisection initial partition (cu:3091-3108 ->
// cu:3046-3062 -> cu:2964-3025).  EXACT: glibc rand() in the reference's
// order.  PARALLEL: isolated vertices set aside and a stream per recursion
// step (cm_initpart.cpp).  AUTO: PARALLEL on graphs of 2^14 vertices or
// more, EXACT below.  Returns the summed bisection cuts, which is the edge
// cut of the result.
enum class InitForm { AUTO, EXACT, PARALLEL };

// One coarser level of a recursion step's chain: its graph, and cmap, the
// map from the next finer level's vertices to this level's.
struct StepLevel {
  int n = 0;
  hvec<int64_t> xadj;
  hvec<int> adjncy, adjwgt, vwgt, cmap;
};
// A coarsener the GPU pipeline lends the PARALLEL form for its large
// recursion steps (graphs of STEP_DEVICE_MIN_M entries or more): it builds
// the step's chain with the same threshold, maxvwgt and stop rule as the host
// chain (cu:1854-1875) and returns the coarser levels, finest first.  It
// throws on a device error.
struct StepCoarsener {
  virtual ~StepCoarsener() = default;
  virtual void coarsen(int n, const int64_t *xadj, const int *adjncy, const int *adjwgt, const int *vwgt,
                       int threshold, int maxvwgt, uint64_t seed, std::vector<StepLevel> &levels) = 0;
};
constexpr int64_t STEP_DEVICE_MIN_M = 1 << 20;

// dev (PARALLEL form only; may be null): the coarsener of the large steps.
int initial_partition(const HostGraph &g, int nparts, std::vector<int> &where, InitForm form = InitForm::AUTO,
                      StepCoarsener *dev = nullptr);

}  // namespace cmh
