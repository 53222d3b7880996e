// Two-grid block preconditioner of the dual-mixed elasticity saddle point
// [[M, B^T], [B, 0]] (stress rows first, then u and t per cell), for MINRES.
//
// P^-1 = diag(D^-1, psign TG(S)): D the 9 x 9 face blocks of the compliance block M,
// S = B D^-1 B^T on the cell unknowns (u_0..2, t_0..2 per cell: 6 x 6 blocks, the cell's
// own block and one per face neighbour), TG(S) one symmetric two-grid cycle on S:
//   x1 = W r,  x2 = x1 + P Ainv P^T (r - S x1),  z = x2 + W (r - S x2),
// W = omega S_cc^-1 (block Jacobi), P piecewise constant per aggregate of cells for each
// of the 6 components, Ainv = (P^T S P)^-1. With an exact coarse solve and a convergent
// smoother TG(S) is SPD, so psign = +1 gives MINRES the SPD preconditioner it needs.
//
// Build (FP64, from the FP64 face-block inverses and the coefficient-independent B blocks):
// nfx_elasticity_tg_blocks forms S and W; nfx_elasticity_tg_coarse accumulates P^T S P
// (deterministic: one thread per coarse row entry, members in aggregate order), whose
// inverse the caller forms. Apply (FP32): nfx_elasticity_tg_apply.
#include "nfx_blockinv.h"
#include "nfx_common.h"
#include "nfx_el_twogrid.h"

namespace nfx {

// Thread per (cell c, local face k): T = B_ck D_f^-1 (6 x 9), the cell-block part
// T B_ck^T into part[c][k] and the neighbour block T B_{n k'}^T into S[c][1 + k].
__global__ __launch_bounds__(kTgBlock) void k_el_tg_blocks(int n_cells, const int32_t* __restrict__ efac,
                                                          const int32_t* __restrict__ nbr,
                                                          const int32_t* __restrict__ nbrk,
                                                          const double* __restrict__ bblk,
                                                          const double* __restrict__ dinv,
                                                          double* __restrict__ part, double* __restrict__ S) {
  const int t = blockIdx.x * kTgBlock + threadIdx.x;
  if (t >= 4 * n_cells) return;
  const int c = t >> 2, k = t & 3;
  const int f = efac[t], n = nbr[t];
  const double* B = bblk + (int64_t)t * 54;
  const double* D = dinv + (int64_t)f * 81;
  const double* Bn = n >= 0 ? bblk + ((int64_t)n * 4 + nbrk[t]) * 54 : nullptr;
  double* P = part + (int64_t)t * 36;
  double* Sn = S + ((int64_t)c * 5 + 1 + k) * 36;
  for (int i = 0; i < 6; ++i) {
    double Ti[9];
#pragma unroll
    for (int j = 0; j < 9; ++j) {
      double v = 0;
#pragma unroll
      for (int l = 0; l < 9; ++l) v += B[i * 9 + l] * D[l * 9 + j];
      Ti[j] = v;
    }
#pragma unroll
    for (int m = 0; m < 6; ++m) {
      double a = 0, b = 0;
#pragma unroll
      for (int j = 0; j < 9; ++j) {
        a += Ti[j] * B[m * 9 + j];
        if (Bn) b += Ti[j] * Bn[m * 9 + j];
      }
      P[i * 6 + m] = a;
      Sn[i * 6 + m] = b;
    }
  }
}

// Thread per cell: S_cc = sum over the 4 faces (in local order), its inverse, the FP32
// copies of the 5 blocks and of W = omega S_cc^-1 at the cell's aggregate-order position
// pos[c], entry-major (entry e of block b at (36 b + e) n_cells + pos; W: e n_cells + pos).
__global__ __launch_bounds__(kTgBlock) void k_el_tg_diag(int n_cells, const double* __restrict__ part,
                                                        const int32_t* __restrict__ pos, double* __restrict__ S,
                                                        float* __restrict__ S32, float* __restrict__ W32,
                                                        double omega, int* __restrict__ status) {
  const int c = blockIdx.x * kTgBlock + threadIdx.x;
  if (c >= n_cells) return;
  const int m = pos[c];
  double A[6][6];
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) {
      double v = 0;
      for (int k = 0; k < 4; ++k) v += part[((int64_t)c * 4 + k) * 36 + i * 6 + j];
      A[i][j] = v;
      S[(int64_t)c * 180 + i * 6 + j] = v;
    }
  for (int b = 0; b < 5; ++b)
    for (int e = 0; e < 36; ++e)
      S32[(int64_t)(36 * b + e) * n_cells + m] = (float)S[(int64_t)c * 180 + b * 36 + e];
  if (!spd_block_inverse<6>(A)) {
    atomicOr(status, 1);
    return;
  }
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j < 6; ++j) W32[(int64_t)(i * 6 + j) * n_cells + m] = (float)(omega * A[i][j]);
}

// Block per aggregate, thread per (i, j) < 36: row 6 A + i of P^T S P, column offset j of
// every aggregate block; the members and their neighbours are visited in a fixed order,
// so the sums are reproducible. Ac (nc x nc, nc = 6 n_agg) must be zero on entry.
__global__ __launch_bounds__(64) void k_el_tg_coarse(int n_agg, const int32_t* __restrict__ agg_ptr,
                                                    const int32_t* __restrict__ agg_mem,
                                                    const int32_t* __restrict__ agg_of,
                                                    const int32_t* __restrict__ nbr, const double* __restrict__ S,
                                                    double* __restrict__ Ac) {
  const int A = blockIdx.x, tid = threadIdx.x;
  if (A >= n_agg || tid >= 36) return;
  const int i = tid / 6, j = tid - 6 * (tid / 6);
  const int64_t nc = 6 * (int64_t)n_agg;
  double* row = Ac + (6 * (int64_t)A + i) * nc + j;
  for (int m = agg_ptr[A]; m < agg_ptr[A + 1]; ++m) {
    const int c = agg_mem[m];
    const double* Sc = S + (int64_t)c * 180 + i * 6 + j;
    row[6 * (int64_t)A] += Sc[0];
    for (int k = 0; k < 4; ++k) {
      const int n = nbr[(int64_t)c * 4 + k];
      if (n >= 0) row[6 * (int64_t)agg_of[n]] += Sc[36 * (1 + k)];
    }
  }
}

// ---- apply (FP32), cell data in aggregate order -------------------------------

// stress rows: z = D^-1 r per face block (optionally scaled by 1 / *dscale)
__global__ __launch_bounds__(kTgBlock) void k_el_tg_stress(int ns, const float* __restrict__ dinv,
                                                          const float* __restrict__ r,
                                                          const double* __restrict__ dscale,
                                                          float* __restrict__ z) {
  const int row = blockIdx.x * kTgBlock + threadIdx.x;
  if (row >= ns) return;
  const float v = el_stress_row(row, dinv, [&](int j) { return r[j]; });
  z[row] = dscale ? v * (float)(1.0 / *dscale) : v;
}

// x1 = W r on the cells
__global__ __launch_bounds__(kTgBlock) void k_el_tg_pre(int n_cells, int ns, int nu,
                                                       const int32_t* __restrict__ agg_mem,
                                                       const float* __restrict__ W, const float* __restrict__ r,
                                                       float* __restrict__ x1) {
  const int m = blockIdx.x * kTgBlock + threadIdx.x;
  if (m >= n_cells) return;
  float rc[6];
  el_gather(agg_mem[m], ns, nu, r, rc);
  el_smooth(m, n_cells, W, rc, x1);
}

// coarse right-hand side rc = P^T (r - S x1): one block of 6 wavefronts per aggregate, wave
// i computes component i of the residual of its positions (lanes over the positions) and
// sums it with a fixed butterfly
__global__ __launch_bounds__(384) void k_el_tg_resid_restrict(int n_agg, int n_cells, int ns, int nu,
                                                             const int32_t* __restrict__ agg_ptr,
                                                             const int32_t* __restrict__ agg_mem,
                                                             const int32_t* __restrict__ nbr,
                                                             const float* __restrict__ S, const float* __restrict__ r,
                                                             const float* __restrict__ x1, float* __restrict__ rc) {
  const int A = blockIdx.x, i = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (A >= n_agg) return;
  float acc = 0.f;
  for (int m = agg_ptr[A] + lane; m < agg_ptr[A + 1]; m += 64) {
    const int c = agg_mem[m];
    int nb[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) nb[k] = nbr[(int64_t)m * 4 + k];
    float t = r[el_row(c, ns, nu, i)];
#pragma unroll
    for (int j = 0; j < 6; ++j) t -= S[(int64_t)(i * 6 + j) * n_cells + m] * x1[(int64_t)m * 6 + j];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if (nb[k] < 0) continue;
#pragma unroll
      for (int j = 0; j < 6; ++j)
        t -= S[(int64_t)(36 * (1 + k) + i * 6 + j) * n_cells + m] * x1[(int64_t)nb[k] * 6 + j];
    }
    acc += t;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
  if (lane == 0) rc[6 * (int64_t)A + i] = acc;
}

// y = Ainv rc: one 256-thread block per row, each thread four-wide loads at a stride of
// the block, a fixed reduction order (waves by butterfly, then the 4 waves in order)
__global__ __launch_bounds__(kTgBlock) void k_el_tg_gemv(int nc, const float* __restrict__ A,
                                                        const float* __restrict__ x, float* __restrict__ y) {
  const int row = blockIdx.x;
  if (row >= nc) return;
  const float* a = A + (int64_t)row * nc;
  float acc = 0.f;
  if ((nc & 3) == 0) {
    const float4* a4 = reinterpret_cast<const float4*>(a);
    const float4* x4 = reinterpret_cast<const float4*>(x);
    const int n4 = nc >> 2;
    int k = threadIdx.x;
    for (; k + 3 * kTgBlock < n4; k += 4 * kTgBlock) {
      const float4 p0 = a4[k], p1 = a4[k + kTgBlock], p2 = a4[k + 2 * kTgBlock], p3 = a4[k + 3 * kTgBlock];
      const float4 q0 = x4[k], q1 = x4[k + kTgBlock], q2 = x4[k + 2 * kTgBlock], q3 = x4[k + 3 * kTgBlock];
      acc += p0.x * q0.x + p0.y * q0.y + p0.z * q0.z + p0.w * q0.w;
      acc += p1.x * q1.x + p1.y * q1.y + p1.z * q1.z + p1.w * q1.w;
      acc += p2.x * q2.x + p2.y * q2.y + p2.z * q2.z + p2.w * q2.w;
      acc += p3.x * q3.x + p3.y * q3.y + p3.z * q3.z + p3.w * q3.w;
    }
    for (; k < n4; k += kTgBlock) {
      const float4 p = a4[k], q = x4[k];
      acc += p.x * q.x + p.y * q.y + p.z * q.z + p.w * q.w;
    }
  } else {
    for (int k = threadIdx.x; k < nc; k += kTgBlock) acc += a[k] * x[k];
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
  __shared__ float part[kTgBlock / 64];
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    float v = 0.f;
#pragma unroll
    for (int w = 0; w < kTgBlock / 64; ++w) v += part[w];
    y[row] = v;
  }
}

// z = psign (x2 + W (r - S x2)) on the cell rows
__global__ __launch_bounds__(kTgBlock) void k_el_tg_post(int n_cells, int ns, int nu,
                                                        const int32_t* __restrict__ agg_mem,
                                                        const int32_t* __restrict__ nbr,
                                                        const int32_t* __restrict__ agg_of,
                                                        const float* __restrict__ S, const float* __restrict__ W,
                                                        const float* __restrict__ r, const float* __restrict__ x1,
                                                        const float* __restrict__ y, float psign,
                                                        const double* __restrict__ dscale, float* __restrict__ z) {
  const int m = blockIdx.x * kTgBlock + threadIdx.x;
  if (m >= n_cells) return;
  const int c = agg_mem[m];
  float rv[6], zc[6];
  el_gather(c, ns, nu, r, rv);
  el_post(m, n_cells, nbr, agg_of, S, W, rv, x1, y, zc);
  const float s = psign * (dscale ? (float)(1.0 / *dscale) : 1.f);
#pragma unroll
  for (int i = 0; i < 6; ++i) z[el_row(c, ns, nu, i)] = zc[i] * s;
}

}  // namespace nfx

using namespace nfx;

static int el_tg_check(const nfx_elasticity_tg* pc) {
  NFX_REQUIRE(pc, "nfx_elasticity_tg: null descriptor");
  const int ns = pc->n_s, nu = pc->n_u, E = pc->n_cells, n_agg = pc->n_agg;
  NFX_REQUIRE(ns >= 0 && ns % 9 == 0 && nu == 3 * E && E > 0 && n_agg > 0, "nfx_elasticity_tg: sizes");
  NFX_REQUIRE(pc->dinv && pc->nbr && pc->sblk && pc->winv && pc->agg_ptr && pc->agg_mem && pc->agg_of && pc->ainv &&
                  pc->work,
              "nfx_elasticity_tg: incomplete descriptor");
  return NFX_OK;
}

int nfx::el_tg_launch_mid(const nfx_elasticity_tg* pc, const float* r, hipStream_t st) {
  const int nc = 6 * pc->n_agg;
  hipLaunchKernelGGL(k_el_tg_resid_restrict, dim3(pc->n_agg), dim3(384), 0, st, pc->n_agg, pc->n_cells, pc->n_s,
                     pc->n_u, pc->agg_ptr, pc->agg_mem, pc->nbr, pc->sblk, r, el_tg_x1(pc), el_tg_rc(pc));
  hipLaunchKernelGGL(k_el_tg_gemv, dim3(nc), dim3(kTgBlock), 0, st, nc, pc->ainv, el_tg_rc(pc), el_tg_y(pc));
  NFX_LAUNCH_CHECK();
  return NFX_OK;
}

int nfx::el_twogrid_apply_scaled(const nfx_elasticity_tg* pc, const float* r, float* z, const double* dscale,
                                 hipStream_t st) {
  int rc = el_tg_check(pc);
  if (rc) return rc;
  NFX_REQUIRE(r && z, "nfx_elasticity_tg_apply: null vector");
  const int ns = pc->n_s, nu = pc->n_u, E = pc->n_cells;
  if (ns)
    hipLaunchKernelGGL(k_el_tg_stress, dim3(grid_for(ns, kTgBlock)), dim3(kTgBlock), 0, st, ns, pc->dinv, r, dscale,
                       z);
  hipLaunchKernelGGL(k_el_tg_pre, dim3(grid_for(E, kTgBlock)), dim3(kTgBlock), 0, st, E, ns, nu, pc->agg_mem,
                     pc->winv, r, el_tg_x1(pc));
  rc = el_tg_launch_mid(pc, r, st);
  if (rc) return rc;
  hipLaunchKernelGGL(k_el_tg_post, dim3(grid_for(E, kTgBlock)), dim3(kTgBlock), 0, st, E, ns, nu, pc->agg_mem,
                     pc->nbr, pc->agg_of, pc->sblk, pc->winv, r, el_tg_x1(pc), el_tg_y(pc), pc->psign, dscale, z);
  NFX_LAUNCH_CHECK();
  return NFX_OK;
}

extern "C" int nfx_elasticity_tg_apply(const nfx_elasticity_tg* pc, const float* d_r, float* d_z, void* stream) {
  return nfx::el_twogrid_apply_scaled(pc, d_r, d_z, nullptr, static_cast<hipStream_t>(stream));
}

extern "C" int nfx_elasticity_tg_blocks(int n_cells, const int32_t* d_efac, const int32_t* d_nbr,
                                        const int32_t* d_nbrk, const double* d_bblk, const double* d_dinv,
                                        const int32_t* d_pos, double* d_part, double* d_S, float* d_S32,
                                        float* d_W32, double omega, int* d_status, void* stream) {
  NFX_REQUIRE(n_cells >= 0, "nfx_elasticity_tg_blocks: n_cells");
  if (n_cells == 0) return NFX_OK;
  NFX_REQUIRE(d_efac && d_nbr && d_nbrk && d_bblk && d_dinv && d_pos && d_part && d_S && d_S32 && d_W32 && d_status,
              "nfx_elasticity_tg_blocks: null pointer");
  NFX_REQUIRE(omega > 0 && omega < 2, "nfx_elasticity_tg_blocks: omega outside (0, 2)");
  hipStream_t st = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(k_el_tg_blocks, dim3(grid_for(4 * (int64_t)n_cells, kTgBlock)), dim3(kTgBlock), 0, st, n_cells,
                     d_efac, d_nbr, d_nbrk, d_bblk, d_dinv, d_part, d_S);
  hipLaunchKernelGGL(k_el_tg_diag, dim3(grid_for(n_cells, kTgBlock)), dim3(kTgBlock), 0, st, n_cells, d_part, d_pos,
                     d_S, d_S32, d_W32, omega, d_status);
  NFX_LAUNCH_CHECK();
  return NFX_OK;
}

extern "C" int nfx_elasticity_tg_coarse(int n_agg, const int32_t* d_agg_ptr, const int32_t* d_agg_mem,
                                        const int32_t* d_agg_of, const int32_t* d_nbr, const double* d_S,
                                        double* d_Ac, void* stream) {
  NFX_REQUIRE(n_agg >= 0, "nfx_elasticity_tg_coarse: n_agg");
  if (n_agg == 0) return NFX_OK;
  NFX_REQUIRE(d_agg_ptr && d_agg_mem && d_agg_of && d_nbr && d_S && d_Ac, "nfx_elasticity_tg_coarse: null pointer");
  hipLaunchKernelGGL(k_el_tg_coarse, dim3(n_agg), dim3(64), 0, static_cast<hipStream_t>(stream), n_agg, d_agg_ptr,
                     d_agg_mem, d_agg_of, d_nbr, d_S, d_Ac);
  NFX_LAUNCH_CHECK();
  return NFX_OK;
}
