// Library-level entry points: version, thread-local error text, device probe.
#include <hip/hip_runtime.h>

#include <string>

#include "nfx_common.h"

namespace nfx {
static thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }
}  // namespace nfx

extern "C" int nfx_version(void) { return 500; }

extern "C" const char* nfx_last_error(void) { return nfx::g_last_error.c_str(); }

extern "C" int nfx_device_count(int* h_count) {
  NFX_REQUIRE(h_count, "nfx_device_count: null pointer");
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess) n = 0;
  *h_count = n;
  return NFX_OK;
}
