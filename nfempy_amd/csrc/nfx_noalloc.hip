// Allocation-free variant of the FP64 refinement residual (nfx_residual64_noalloc).
//
// nfx_residual64 obtains the residual kernel's per-block partial sums from the stream-ordered
// allocator (hipMallocAsync / hipFreeAsync) on every call. Callers that cannot allocate
// (code under graph capture, tight serving loops, harnesses that forbid allocation) use this
// entry point instead: the partial sums live in `scratch`, a file-scope device array of
// 2 x kResMaxBlocks doubles (32 KiB) that every GPU the library runs on holds until the process
// ends. The residual kernel's grid is capped at kResMaxBlocks blocks, so the array covers any
// n_rows. Each block overwrites its two partial sums, so nothing needs clearing between calls;
// calls that overlap on one GPU share the array and race.
//
// In host code `scratch` names a host-side stand-in, not a device address: the device address
// of each GPU's copy comes from hipGetSymbolAddress, taken once per device on the first call
// with that device current, and only that address is passed to the kernel.
#include <mutex>

#include "nfx_common.h"
#include "nfx_residual.h"

namespace nfx {

static __device__ __align__(256) double scratch[2 * kResMaxBlocks];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mutex;
static double* g_scratch_ptr[kMaxDevices] = {};

static int scratch_address(double** out) {
  int dev = -1;
  NFX_CHECK_HIP(hipGetDevice(&dev));
  NFX_REQUIRE(dev >= 0 && dev < kMaxDevices, "nfx_residual64_noalloc: device index beyond 64");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    NFX_CHECK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    NFX_REQUIRE(p != nullptr, "nfx_residual64_noalloc: no device address for the scratch array");
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[dev];
  return NFX_OK;
}

}  // namespace nfx

using namespace nfx;

extern "C" int nfx_residual64_noalloc(int n_rows, const int32_t* d_rowptr, const int32_t* d_col,
                                      const double* d_val, const double* d_x, const double* d_b,
                                      double* d_r, void* stream) {
  NFX_REQUIRE(n_rows > 0, "nfx_residual64_noalloc: n_rows");
  NFX_REQUIRE(d_rowptr && d_col && d_val && d_x && d_b && d_r, "nfx_residual64_noalloc: null pointer");
  const int blocks = residual_blocks(n_rows, kResidualLanesDefault);
  NFX_REQUIRE(blocks <= kResMaxBlocks, "nfx_residual64_noalloc: grid beyond the scratch array");
  double* partial = nullptr;
  const int rc = scratch_address(&partial);
  if (rc != NFX_OK) return rc;
  NFX_CHECK_HIP(launch_residual64(kResidualLanesDefault, blocks, static_cast<hipStream_t>(stream), n_rows,
                                  d_rowptr, d_col, d_val, d_x, d_b, d_r, partial));
  return NFX_OK;
}
