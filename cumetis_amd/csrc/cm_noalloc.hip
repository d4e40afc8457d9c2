// cm_noalloc.hip -- the segmented sort for callers that cannot allocate.
//
// cm_segsort_i32_noalloc runs the kernels of cm_segsort_i32 with all of its
// scratch (packed keys, bin lists, scan status words, merge ping-pong
// buffers, merge descriptors) carved from one static device array, so the
// call itself performs no hipMalloc/hipFree and never touches the context's
// pool.  Each device the library runs on holds its own copy of the array
// (CM_NOALLOC_SCRATCH_BYTES) until the process ends.  Its device address is
// looked up once per device, when the first context on that device is
// created (cm_noalloc_init), and only that address is handed to kernels.
#include <mutex>

#include "cm_internal.hpp"
#include "cm_segsort.hpp"

static __device__ __align__(256) unsigned char scratch[CM_NOALLOC_SCRATCH_BYTES];

namespace cm {

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
void *g_ptr[MAX_DEVICES] = {};
}  // namespace

// With `device` current: the device address of this device's scratch array
// (hipGetSymbolAddress once per device, then the kept address).
void *noalloc_init(int device)
{
  CM_REQUIRE(device >= 0 && device < MAX_DEVICES, CM_ERR_UNSUPPORTED, "device index beyond 64");
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      throw Error(CM_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_ptr[device] = p;
  }
  return g_ptr[device];
}

}  // namespace cm

extern "C" int cm_segsort_i32_noalloc(cm_ctx *ctx, int32_t *d_keys, int32_t *d_vals, int32_t n,
                                      const int32_t *d_segs, int32_t length)
{
  if (!ctx || n < 0 || length < 0) return CM_ERR_INVALID;
  try {
    CM_REQUIRE(n == 0 || length == 0 || (d_keys && d_vals && d_segs), CM_ERR_INVALID, "NULL array");
    const size_t need = cm::segsort_i32_arena_bytes(n, length);
    CM_REQUIRE(need <= (size_t)CM_NOALLOC_SCRATCH_BYTES, CM_ERR_UNSUPPORTED,
               "cm_segsort_i32_noalloc: " + std::to_string(n) + " keys in " + std::to_string(length) +
                   " segments need " + std::to_string(need) + " bytes of scratch, more than " +
                   std::to_string((size_t)CM_NOALLOC_SCRATCH_BYTES));
    CM_REQUIRE(ctx->noalloc_scratch != nullptr, CM_ERR_INVALID, "context has no scratch address");
    CM_HIP(hipSetDevice(ctx->device));
    cm::SegArena ar;
    ar.base = (char *)ctx->noalloc_scratch;
    ar.cap = CM_NOALLOC_SCRATCH_BYTES;
    cm::segsort_i32(ctx, d_keys, d_vals, n, d_segs, length, &ar);
  } catch (const cm::Error &e) {
    ctx->err = e.what();
    return e.code;
  } catch (const std::exception &e) {
    ctx->err = e.what();
    return CM_ERR_INVALID;
  }
  return CM_OK;
}
