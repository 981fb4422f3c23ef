// noalloc.hip -- the entry point that allocates no device memory: the
// Gram-Schmidt multi-dot of GMRES without an operator handle, for callers
// that cannot call hipMalloc (code under HIP graph capture, tight serving
// loops, harnesses that forbid allocation).  Its per-block partial sums live
// in a file-scope device array of fixed size; the results go to device
// memory on the caller's stream, with no host synchronisation.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <string>

#include "../../include/hpsint_amd.h"
#include "blas.hpp"

int hps_internal_fail(int code, const char *msg);

namespace
{
  // one partial row of k <= MAXK doubles per reduction block
  constexpr int    NOALLOC_MAX_K   = hps::MAXK;
  constexpr size_t SCRATCH_DOUBLES = (size_t)hps::RED_BLOCKS * NOALLOC_MAX_K; // 512 KiB

  // every GPU the library runs on holds its own copy until the process ends
  __device__ __align__(256) double scratch[SCRATCH_DOUBLES];

  // device address of `scratch` per device (the host-side symbol is only a
  // stand-in and must never be handed to a kernel or a copy)
  constexpr int MAX_DEVICES = 64;
  double       *g_scratch[MAX_DEVICES];
  std::mutex    g_scratch_mutex;

  int scratch_of_current_device(double **out)
  {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess)
      return hps_internal_fail(HPS_ERR_HIP, (std::string("hipGetDevice: ") +
                                             hipGetErrorString(e)).c_str());
    if (dev < 0 || dev >= MAX_DEVICES)
      return hps_internal_fail(HPS_ERR_UNSUPPORTED, "device index beyond 63");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch[dev])
      {
        void *p = nullptr;
        e       = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
          return hps_internal_fail(HPS_ERR_HIP, (std::string("hipGetSymbolAddress: ") +
                                                 hipGetErrorString(e)).c_str());
        g_scratch[dev] = static_cast<double *>(p);
      }
    *out = g_scratch[dev];
    return HPS_OK;
  }
} // namespace

extern "C" int hps_multi_dot_noalloc(void *stream, int64_t n, int k, const double *V,
                                     const double *w, double *h)
{
  if (k < 1 || k > NOALLOC_MAX_K)
    return hps_internal_fail(HPS_ERR_INVALID_ARGUMENT,
                             "multi_dot_noalloc: k must be 1..32 (the scratch array holds "
                             "2048 x 32 partial sums)");
  if (n < 0 || !h || (n > 0 && (!V || !w)))
    return hps_internal_fail(HPS_ERR_INVALID_ARGUMENT, "multi_dot_noalloc: bad arguments");
  double *ptr = nullptr;
  if (int r = scratch_of_current_device(&ptr))
    return r;
  // the partial rows are written in full by every call (each reduction
  // block stores its k sums, empty blocks store zeros): nothing to clear
  hps::blas_multi_dot(static_cast<hipStream_t>(stream), n, k, V, w, ptr, h);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    return hps_internal_fail(HPS_ERR_HIP,
                             (std::string("kernel launch: ") + hipGetErrorString(e)).c_str());
  return HPS_OK;
}
