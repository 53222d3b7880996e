// Device helpers of the elasticity two-grid block preconditioner (nfx_elasticity_tg),
// shared by its standalone apply (nfx_el_twogrid.hip) and the MINRES kernels that fuse
// its first and last steps with the Lanczos updates (nfx_krylov.hip).
//
// Cell data live in aggregate order: position m holds cell agg_mem[m], the positions of
// one aggregate are contiguous (agg_ptr), nbr / agg_of / x1 / S / W are indexed by
// position, S and W entry-major (entry e of block b at (36 b + e) n_cells + m), so the
// loads of the lanes of one aggregate coalesce.
#pragma once

#include "nfx_common.h"

namespace nfx {

constexpr int kTgBlock = 256;

// the 6 cell unknowns (u_0..2, t_0..2) of cell c in a full vector
__device__ __forceinline__ void el_gather(int c, int ns, int nu, const float* __restrict__ r, float (&v)[6]) {
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    v[i] = r[ns + 3 * (int64_t)c + i];
    v[3 + i] = r[ns + nu + 3 * (int64_t)c + i];
  }
}

__device__ __forceinline__ int64_t el_row(int c, int ns, int nu, int i) {
  return i < 3 ? ns + 3 * (int64_t)c + i : ns + nu + 3 * (int64_t)c + (i - 3);
}

// row `row` of the face-block product D^-1 r (r read through a functor giving r[j])
template <typename F>
__device__ __forceinline__ float el_stress_row(int row, const float* __restrict__ dinv, F r) {
  const int f = row / 9, a = row - 9 * f;
  const float* D = dinv + (int64_t)f * 81 + a * 9;
  float v = 0.f;
#pragma unroll
  for (int b = 0; b < 9; ++b) v += D[b] * r(9 * f + b);
  return v;
}

// x1[m] = W_m rc
__device__ __forceinline__ void el_smooth(int m, int n_cells, const float* __restrict__ W, const float (&rc)[6],
                                          float* __restrict__ x1) {
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    float v = 0.f;
#pragma unroll
    for (int j = 0; j < 6; ++j) v += W[(int64_t)(i * 6 + j) * n_cells + m] * rc[j];
    x1[(int64_t)m * 6 + i] = v;
  }
}

// x at position m: x1, plus the coarse correction y[aggregate] when y is not null
__device__ __forceinline__ void el_load_x(int m, const float* __restrict__ x1, const int32_t* __restrict__ agg_of,
                                          const float* __restrict__ y, float (&v)[6]) {
#pragma unroll
  for (int i = 0; i < 6; ++i) v[i] = x1[(int64_t)m * 6 + i];
  if (y) {
    const int64_t a = 6 * (int64_t)agg_of[m];
#pragma unroll
    for (int i = 0; i < 6; ++i) v[i] += y[a + i];
  }
}

// t = rc - S x at position m (rc the cell's right-hand side); xc gets the cell's own x
__device__ __forceinline__ void el_residual(int m, int n_cells, const int32_t* __restrict__ nbr,
                                            const float* __restrict__ S, const float (&rc)[6],
                                            const float* __restrict__ x1, const int32_t* __restrict__ agg_of,
                                            const float* __restrict__ y, float (&t)[6], float (&xc)[6]) {
  int nb[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) nb[k] = nbr[(int64_t)m * 4 + k];
#pragma unroll
  for (int i = 0; i < 6; ++i) t[i] = rc[i];
  el_load_x(m, x1, agg_of, y, xc);
#pragma unroll
  for (int i = 0; i < 6; ++i)
#pragma unroll
    for (int j = 0; j < 6; ++j) t[i] -= S[(int64_t)(i * 6 + j) * n_cells + m] * xc[j];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if (nb[k] < 0) continue;
    float xn[6];
    el_load_x(nb[k], x1, agg_of, y, xn);
#pragma unroll
    for (int i = 0; i < 6; ++i)
#pragma unroll
      for (int j = 0; j < 6; ++j) t[i] -= S[(int64_t)(36 * (1 + k) + i * 6 + j) * n_cells + m] * xn[j];
  }
}

// z = x2 + W (rc - S x2), x2 = x1 + P y, at position m (the post-smoothing step)
__device__ __forceinline__ void el_post(int m, int n_cells, const int32_t* __restrict__ nbr,
                                        const int32_t* __restrict__ agg_of, const float* __restrict__ S,
                                        const float* __restrict__ W, const float (&rc)[6],
                                        const float* __restrict__ x1, const float* __restrict__ y, float (&z)[6]) {
  float t2[6], x2[6];
  el_residual(m, n_cells, nbr, S, rc, x1, agg_of, y, t2, x2);
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    float v = x2[i];
#pragma unroll
    for (int j = 0; j < 6; ++j) v += W[(int64_t)(i * 6 + j) * n_cells + m] * t2[j];
    z[i] = v;
  }
}

// The middle of the cycle on the device work buffers: rc = P^T (r - S x1) and y = Ainv rc
// (x1 must be in the work buffer already).
int el_tg_launch_mid(const nfx_elasticity_tg* pc, const float* r, hipStream_t st);
// work buffer views: x1[6 n_cells] | rc[6 n_agg] | y[6 n_agg] (16-byte aligned offsets)
inline float* el_tg_x1(const nfx_elasticity_tg* pc) { return pc->work; }
inline float* el_tg_rc(const nfx_elasticity_tg* pc) {
  return pc->work + ((6 * (int64_t)pc->n_cells + 3) & ~(int64_t)3);
}
inline float* el_tg_y(const nfx_elasticity_tg* pc) {
  return el_tg_rc(pc) + ((6 * (int64_t)pc->n_agg + 3) & ~(int64_t)3);
}

}  // namespace nfx
