// FP64 refinement residual r = b - A x on a CSR matrix, shared by the MINRES and FGMRES
// refinement loops and the nfx_residual64 entry points.
//
// LANES consecutive lanes own one row (4, 8 or 16, from the mean row length): each lane
// issues the column and value loads of three entries before their gathers of x, the
// group reduces with a fixed butterfly, so a row's sum has the same order on every run
// (bitwise reproducible). Partial sums of r.r and b.b are written per block and summed by
// the caller in block order.
#pragma once

#include "nfx_common.h"

namespace nfx {

constexpr int kResBlock = 256;
// grid cap of the residual kernels (each block writes 2 partial sums)
constexpr int kResMaxBlocks = 2048;

// lane count of the standalone entry points, which do not read the row lengths
constexpr int kResidualLanesDefault = 8;

inline int residual_lanes(double mean_row) { return mean_row > 24 ? 16 : (mean_row > 10 ? 8 : 4); }

inline int residual_blocks(int n_rows, int lanes) {
  const int b = grid_for(n_rows, kResBlock / lanes);
  return b < kResMaxBlocks ? (b > 0 ? b : 1) : kResMaxBlocks;
}

template <int LANES>
__device__ __forceinline__ double res_group_sum(double v) {
#pragma unroll
  for (int off = LANES / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, LANES);
  return v;
}

__device__ __forceinline__ double res_wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

template <int LANES>
__global__ __launch_bounds__(kResBlock) void k_residual64_rows(
    int n, const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
    const double* __restrict__ va, const double* __restrict__ x, const double* __restrict__ b,
    double* __restrict__ r, double* __restrict__ partial) {
  constexpr int kRows = kResBlock / LANES;
  const int lane = threadIdx.x % LANES;
  double rr = 0, bb = 0;
  for (int row = blockIdx.x * kRows + threadIdx.x / LANES; row < n; row += gridDim.x * kRows) {
    const int j1 = rp[row + 1];
    double s = 0;
    for (int j = rp[row] + lane; j < j1; j += 3 * LANES) {
      const bool pb = j + LANES < j1, pc = j + 2 * LANES < j1;
      const int ca = ci[j];
      const int cb = pb ? ci[j + LANES] : 0;
      const int cc = pc ? ci[j + 2 * LANES] : 0;
      const double a0 = va[j];
      const double a1 = pb ? va[j + LANES] : 0.0;
      const double a2 = pc ? va[j + 2 * LANES] : 0.0;
      const double x0 = x[ca];
      const double x1 = pb ? x[cb] : 0.0;
      const double x2 = pc ? x[cc] : 0.0;
      s += a0 * x0;
      s += a1 * x1;
      s += a2 * x2;
    }
    s = res_group_sum<LANES>(s);
    if (lane == 0) {
      const double bi = b[row];
      const double ri = bi - s;
      r[row] = ri;
      rr += ri * ri;
      bb += bi * bi;
    }
  }
  __shared__ double sh[2][kResBlock / 64];
  rr = res_wave_sum(rr);
  bb = res_wave_sum(bb);
  if ((threadIdx.x & 63) == 0) {
    sh[0][threadIdx.x >> 6] = rr;
    sh[1][threadIdx.x >> 6] = bb;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0, c = 0;
#pragma unroll
    for (int w = 0; w < kResBlock / 64; ++w) {
      a += sh[0][w];
      c += sh[1][w];
    }
    partial[2 * blockIdx.x] = a;
    partial[2 * blockIdx.x + 1] = c;
  }
}

// Launch with the lane count of the matrix; blocks from residual_blocks(n, lanes).
inline hipError_t launch_residual64(int lanes, int blocks, hipStream_t stream, int n,
                                    const int32_t* rp, const int32_t* ci, const double* va,
                                    const double* x, const double* b, double* r, double* partial) {
  if (lanes == 16)
    hipLaunchKernelGGL(k_residual64_rows<16>, dim3(blocks), dim3(kResBlock), 0, stream, n, rp, ci,
                       va, x, b, r, partial);
  else if (lanes == 8)
    hipLaunchKernelGGL(k_residual64_rows<8>, dim3(blocks), dim3(kResBlock), 0, stream, n, rp, ci,
                       va, x, b, r, partial);
  else
    hipLaunchKernelGGL(k_residual64_rows<4>, dim3(blocks), dim3(kResBlock), 0, stream, n, rp, ci,
                       va, x, b, r, partial);
  return hipGetLastError();
}

}  // namespace nfx
