// Sort-based numbering on the device: the stable sort of int64 keys behind the pattern
// build (the union of element (row, col) entries in row-major order) and the
// first-appearance numbering behind entity ids (edges, faces, volumes in the order the
// reference's mesh builder meets them) and the tag-filtered entity orders.
//
// First-appearance ranks from a stable radix sort: in the sorted keys every run of equal
// keys starts with its first input position (the sort is stable); marking those positions
// in input order and scanning the marks gives each distinct key its rank by first
// appearance, with no second sort.
#include <hipcub/hipcub.hpp>

#include "nfx_common.h"

namespace nfx {
namespace {

constexpr int kSB = 256;

inline size_t align_up(size_t x) { return (x + 255) & ~size_t(255); }

__global__ __launch_bounds__(kSB) void k_iota(int n, int32_t* __restrict__ a) {
  const int i = blockIdx.x * kSB + threadIdx.x;
  if (i < n) a[i] = i;
}

__global__ __launch_bounds__(kSB) void k_zero(int n, int32_t* __restrict__ a) {
  const int i = blockIdx.x * kSB + threadIdx.x;
  if (i < n) a[i] = 0;
}

// head[i] = 1 where a run of equal sorted keys starts; mark[perm[i]] = 1 there
__global__ __launch_bounds__(kSB) void k_heads(int n, const int64_t* __restrict__ ks, const int32_t* __restrict__ perm,
                                               int32_t* __restrict__ head, int32_t* __restrict__ mark) {
  const int i = blockIdx.x * kSB + threadIdx.x;
  if (i >= n) return;
  const int h = (i == 0 || ks[i] != ks[i - 1]) ? 1 : 0;
  head[i] = h;
  if (h) mark[perm[i]] = 1;
}

// seg[i] = 1 + run index of sorted position i (inclusive scan of head); run r starts at
// input position p = perm[start]: rank = scan of marks before p
__global__ __launch_bounds__(kSB) void k_first_pos(int n, const int32_t* __restrict__ head,
                                                   const int32_t* __restrict__ seg, const int32_t* __restrict__ perm,
                                                   const int32_t* __restrict__ mscan, int32_t* __restrict__ run_rank,
                                                   int32_t* __restrict__ first) {
  const int i = blockIdx.x * kSB + threadIdx.x;
  if (i >= n || !head[i]) return;
  const int p = perm[i];
  const int r = mscan[p];
  run_rank[seg[i] - 1] = r;
  first[r] = p;
}

__global__ __launch_bounds__(kSB) void k_ids(int n, const int32_t* __restrict__ seg, const int32_t* __restrict__ perm,
                                             const int32_t* __restrict__ run_rank, int32_t* __restrict__ ids) {
  const int i = blockIdx.x * kSB + threadIdx.x;
  if (i >= n) return;
  ids[perm[i]] = run_rank[seg[i] - 1];
}

__global__ void k_count(int n, const int32_t* __restrict__ seg, int32_t* __restrict__ count) {
  if (threadIdx.x == 0 && blockIdx.x == 0) count[0] = n > 0 ? seg[n - 1] : 0;
}

struct Ws {
  int64_t* ks;
  int32_t* pin;
  int32_t* perm;
  int32_t* head;
  int32_t* seg;
  int32_t* mark;
  int32_t* mscan;
  int32_t* rrank;
  void* temp;
  size_t temp_bytes;
};

size_t cub_temp(int n) {
  size_t a = 0, b = 0, c = 0;
  (void)hipcub::DeviceRadixSort::SortPairs(nullptr, a, (const int64_t*)nullptr, (int64_t*)nullptr,
                                           (const int32_t*)nullptr, (int32_t*)nullptr, n, 0, 64);
  (void)hipcub::DeviceScan::InclusiveSum(nullptr, b, (const int32_t*)nullptr, (int32_t*)nullptr, n);
  (void)hipcub::DeviceScan::ExclusiveSum(nullptr, c, (const int32_t*)nullptr, (int32_t*)nullptr, n);
  return std::max(a, std::max(b, c));
}

size_t ws_bytes(int n) {
  return align_up(8 * (size_t)n) + 7 * align_up(4 * (size_t)n) + align_up(cub_temp(n));
}

Ws carve(void* w, int n) {
  char* p = static_cast<char*>(w);
  Ws s;
  s.ks = reinterpret_cast<int64_t*>(p);
  p += align_up(8 * (size_t)n);
  int32_t** arr[7] = {&s.pin, &s.perm, &s.head, &s.seg, &s.mark, &s.mscan, &s.rrank};
  for (int k = 0; k < 7; ++k) {
    *arr[k] = reinterpret_cast<int32_t*>(p);
    p += align_up(4 * (size_t)n);
  }
  s.temp = p;
  s.temp_bytes = align_up(cub_temp(n));
  return s;
}

int sort_into(int n, int key_bits, const int64_t* keys, Ws& s, hipStream_t st) {
  hipLaunchKernelGGL(k_iota, dim3(grid_for(n, kSB)), dim3(kSB), 0, st, n, s.pin);
  size_t tb = s.temp_bytes;
  NFX_CHECK_HIP(hipcub::DeviceRadixSort::SortPairs(s.temp, tb, keys, s.ks, s.pin, s.perm, n, 0, key_bits, st));
  NFX_LAUNCH_CHECK();
  return NFX_OK;
}

}  // namespace
}  // namespace nfx

using namespace nfx;

extern "C" {

int nfx_sort_workspace(int64_t K, size_t* h_bytes) {
  NFX_REQUIRE(h_bytes && K >= 0 && K < (int64_t(1) << 31), "nfx_sort_workspace: bad arguments");
  *h_bytes = ws_bytes((int)K);
  return NFX_OK;
}

int nfx_stable_sort64(int64_t K, int key_bits, const int64_t* d_keys, int64_t* d_sorted_keys, int32_t* d_perm,
                      void* d_work, size_t work_bytes, void* stream) {
  NFX_REQUIRE(K >= 0 && K < (int64_t(1) << 31), "nfx_stable_sort64: K out of range");
  NFX_REQUIRE(key_bits > 0 && key_bits <= 64, "nfx_stable_sort64: key_bits");
  if (K == 0) return NFX_OK;
  NFX_REQUIRE(d_keys && d_sorted_keys && d_perm && d_work, "nfx_stable_sort64: null pointer");
  const int n = (int)K;
  NFX_REQUIRE(work_bytes >= ws_bytes(n), "nfx_stable_sort64: workspace too small");
  hipStream_t st = (hipStream_t)stream;
  Ws s = carve(d_work, n);
  const int rc = sort_into(n, key_bits, d_keys, s, st);
  if (rc) return rc;
  NFX_CHECK_HIP(hipMemcpyAsync(d_sorted_keys, s.ks, 8 * (size_t)n, hipMemcpyDeviceToDevice, st));
  NFX_CHECK_HIP(hipMemcpyAsync(d_perm, s.perm, 4 * (size_t)n, hipMemcpyDeviceToDevice, st));
  return NFX_OK;
}

int nfx_first_appearance(int64_t K, int key_bits, const int64_t* d_keys, int32_t* d_ids, int32_t* d_first,
                         int32_t* d_count, void* d_work, size_t work_bytes, void* stream) {
  NFX_REQUIRE(K >= 0 && K < (int64_t(1) << 31), "nfx_first_appearance: K out of range");
  NFX_REQUIRE(key_bits > 0 && key_bits <= 64, "nfx_first_appearance: key_bits");
  NFX_REQUIRE(d_count, "nfx_first_appearance: null count");
  hipStream_t st = (hipStream_t)stream;
  if (K == 0) {
    hipLaunchKernelGGL(k_zero, dim3(1), dim3(kSB), 0, st, 1, d_count);
    NFX_LAUNCH_CHECK();
    return NFX_OK;
  }
  NFX_REQUIRE(d_keys && d_ids && d_first && d_work, "nfx_first_appearance: null pointer");
  const int n = (int)K;
  NFX_REQUIRE(work_bytes >= ws_bytes(n), "nfx_first_appearance: workspace too small");
  Ws s = carve(d_work, n);
  int rc = sort_into(n, key_bits, d_keys, s, st);
  if (rc) return rc;
  const dim3 g(grid_for(n, kSB)), b(kSB);
  hipLaunchKernelGGL(k_zero, g, b, 0, st, n, s.mark);
  hipLaunchKernelGGL(k_heads, g, b, 0, st, n, s.ks, s.perm, s.head, s.mark);
  size_t tb = s.temp_bytes;
  NFX_CHECK_HIP(hipcub::DeviceScan::InclusiveSum(s.temp, tb, s.head, s.seg, n, st));
  tb = s.temp_bytes;
  NFX_CHECK_HIP(hipcub::DeviceScan::ExclusiveSum(s.temp, tb, s.mark, s.mscan, n, st));
  hipLaunchKernelGGL(k_first_pos, g, b, 0, st, n, s.head, s.seg, s.perm, s.mscan, s.rrank, d_first);
  hipLaunchKernelGGL(k_ids, g, b, 0, st, n, s.seg, s.perm, s.rrank, d_ids);
  hipLaunchKernelGGL(k_count, dim3(1), dim3(64), 0, st, n, s.seg, d_count);
  NFX_LAUNCH_CHECK();
  return NFX_OK;
}

}  // extern "C"
