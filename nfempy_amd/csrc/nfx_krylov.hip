// Sparse kernels and the FP32 MINRES solver inside FP64 iterative refinement.
//
// Reductions are deterministic: every block writes its partial sum, an integer
// ticket picks the last block to finish, and that block sums the partials in a
// fixed order (no floating-point atomics). Scalars of the Krylov recurrence live
// in device memory; the host only reads the residual estimate every few
// iterations, so an iteration is two kernel launches:
//   A(k): x/w update of iteration k-1, y = D A z / beta - (beta/oldb) r1, alpha = v.y
//   B(k): y -= (alpha/beta) r2, z = P^-1 y, beta' = sqrt(y.z), Givens update
#include <algorithm>
#include <cmath>
#include <vector>

#include "nfx_common.h"
#include "nfx_el_twogrid.h"
#include "nfx_residual.h"

namespace nfx {

constexpr int kBlock = 256;
constexpr int kMaxBlocks = 1024;
// Block size of the two plain MINRES kernels: with the grid capped at kMaxBlocks (every
// block sums the other kernel's block partials itself), 512 threads keep twice the waves in
// flight for the SpMV's gathers: the 32^3 x 6 Darcy solve (2,704 iterations) took 99 ms with
// 256 threads, 75 ms with 512 in the A kernel and 72 ms in both; 1,024 threads took 105 ms.
constexpr int kBlockMinres = 512;

struct MinresState {
  double beta, oldb, alfa, beta1;
  double dbar, epsln, oldeps, phibar, cs, sn;
  double delta, denom, phi;
  double s, ratio;  // 1/beta and beta/oldb for the next A kernel
  double rnorm2;    // scratch for norms
  double bnorm2;
  int stop;
  int pad;
};

__device__ __forceinline__ double warp_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// Deterministic block sum; result valid in thread 0.
__device__ __forceinline__ double block_sum(double v, double* sh) {
  v = warp_sum(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double r = 0;
  if (threadIdx.x == 0) {
    const int nw = (blockDim.x + 63) >> 6;
    for (int i = 0; i < nw; ++i) r += sh[i];
  }
  __syncthreads();
  return r;
}

template <int LANES>
__device__ __forceinline__ float group_sum(float v) {
#pragma unroll
  for (int off = LANES / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, LANES);
  return v;
}

// ---------------------------------------------------------------------------
// plain SpMV: y = alpha A x + beta y (FP32, LANES lanes per row)
template <int LANES, int BLK = kBlock>
__global__ __launch_bounds__(BLK) void k_spmv(int n, const int32_t* __restrict__ rp,
                                                 const int32_t* __restrict__ ci,
                                                 const float* __restrict__ va, float alpha,
                                                 const float* __restrict__ x, float beta,
                                                 float* __restrict__ y) {
  const int lane = threadIdx.x % LANES;
  const int rows_per_grid = gridDim.x * (BLK / LANES);
  for (int row = blockIdx.x * (BLK / LANES) + threadIdx.x / LANES; row < n;
       row += rows_per_grid) {
    // three entries per lane per pass, predicated: the column / value loads of a pass are
    // issued together, then the gathers of x (one pass covers rows of up to 3 LANES entries)
    const int j1 = rp[row + 1];
    float s = 0.f;
    for (int j = rp[row] + lane; j < j1; j += 3 * LANES) {
      const bool pb = j + LANES < j1, pc = j + 2 * LANES < j1;
      const int ca = ci[j], cb = pb ? ci[j + LANES] : 0, cc = pc ? ci[j + 2 * LANES] : 0;
      const float va_ = va[j], vb = pb ? va[j + LANES] : 0.f, vc = pc ? va[j + 2 * LANES] : 0.f;
      const float xa = x[ca], xb = pb ? x[cb] : 0.f, xc = pc ? x[cc] : 0.f;
      s += va_ * xa + vb * xb + vc * xc;
    }
    s = group_sum<LANES>(s);
    if (lane == 0) y[row] = alpha * s + (beta == 0.f ? 0.f : beta * y[row]);
  }
}

__global__ __launch_bounds__(kBlock) void k_diag64(int n, const int32_t* __restrict__ rp,
                                                   const int32_t* __restrict__ ci,
                                                   const double* __restrict__ va,
                                                   double* __restrict__ d) {
  const int row = blockIdx.x * kBlock + threadIdx.x;
  if (row >= n) return;
  double v = 0;
  for (int j = rp[row]; j < rp[row + 1]; ++j)
    if (ci[j] == row) v = va[j];
  d[row] = v;
}

// S_kk = sum_{j < nu} A[k, j] * A[j, k] / A[j, j]; A[j, k] found by a search in row j.
__global__ __launch_bounds__(kBlock) void k_schur64(int n, int nu, const int32_t* __restrict__ rp,
                                                    const int32_t* __restrict__ ci,
                                                    const double* __restrict__ va,
                                                    const double* __restrict__ d,
                                                    double* __restrict__ out) {
  const int k = nu + blockIdx.x * kBlock + threadIdx.x;
  if (k >= n) return;
  double s = 0;
  for (int jj = rp[k]; jj < rp[k + 1]; ++jj) {
    const int j = ci[jj];
    if (j >= nu) continue;
    double akj = va[jj], ajk = 0;
    int lo = rp[j], hi = rp[j + 1];
    while (lo < hi) {  // columns are sorted
      const int mid = (lo + hi) >> 1;
      if (ci[mid] < k) lo = mid + 1; else hi = mid;
    }
    if (lo < rp[j + 1] && ci[lo] == k) ajk = va[lo];
    s += akj * ajk / d[j];
  }
  out[k - nu] = s;
}

// ---------------------------------------------------------------------------
// MINRES kernels (FP32 vectors, FP64 scalars)
//
// Reductions: a producer kernel writes one partial per block; every block of the
// consumer kernel sums all partials in the same fixed order in its prologue, so
// all blocks derive identical scalars without atomics. Scalar state is double
// buffered: A(k) reads S[(k+1)&1] and writes S[k&1]; B(k) adds alpha_k to S[k&1].

// Fixed-order sum of partial[0..np) (every block gets the same value).
__device__ __forceinline__ double sum_partials(const double* __restrict__ partial, int np,
                                               double* sh) {
  double v = 0;
  for (int i = threadIdx.x; i < np; i += blockDim.x) v += partial[i];
  v = block_sum(v, sh);
  __shared__ double bc;
  if (threadIdx.x == 0) bc = v;
  __syncthreads();
  return bc;
}

// Initial step: r1 = r2 = rhs, z = P^-1 rhs, x = w* = 0, partials of rhs . z.
__global__ __launch_bounds__(kBlock) void k_minres_init(int n, const float* __restrict__ rhs,
                                                        const int32_t* __restrict__ prp,
                                                        const int32_t* __restrict__ pci,
                                                        const float* __restrict__ pva,
                                                        float* r1, float* r2, float* z,
                                                        float* x, float* w0, float* w1, float* w2,
                                                        double* partial_b) {
  __shared__ double sh[kBlock / 64];
  double acc = 0;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    const float b = rhs[i];
    float zz = 0.f;
    for (int j = prp[i]; j < prp[i + 1]; ++j) zz += pva[j] * rhs[pci[j]];
    r1[i] = b;
    r2[i] = b;
    z[i] = zz;
    x[i] = 0.f;
    w0[i] = 0.f;
    w1[i] = 0.f;
    w2[i] = 0.f;
    acc += (double)b * (double)zz;
  }
  const double bs = block_sum(acc, sh);
  if (threadIdx.x == 0) partial_b[blockIdx.x] = bs;
}

// Scalar step of A(k): beta_k from B(k-1)'s partials, Givens rotation of
// iteration k-1 (Paige-Saunders, as in SciPy's minres).
__device__ __forceinline__ void minres_scalars(const MinresState& in, MinresState& out,
                                               double beta, int first) {
  out = in;
  if (first) {
    out.beta1 = beta;
    out.beta = beta;
    out.oldb = 0;
    out.alfa = 0;
    out.dbar = 0;
    out.epsln = 0;
    out.oldeps = 0;
    out.phibar = beta;
    out.cs = -1;
    out.sn = 0;
    out.delta = 0;
    out.denom = 0;
    out.phi = 0;
    out.s = beta > 0 ? 1.0 / beta : 0.0;
    out.ratio = 0;
    out.stop = beta > 0 ? 0 : 1;
    return;
  }
  const double alfa = in.alfa;
  const double oldb = in.beta;
  const double oldeps = in.epsln;
  const double delta = in.cs * in.dbar + in.sn * alfa;
  const double gbar = in.sn * in.dbar - in.cs * alfa;
  const double epsln = in.sn * beta;
  const double dbar = -in.cs * beta;
  double gamma = hypot(gbar, beta);
  if (gamma < 1e-300) gamma = 1e-300;
  out.oldb = oldb;
  out.beta = beta;
  out.oldeps = oldeps;
  out.delta = delta;
  out.epsln = epsln;
  out.dbar = dbar;
  out.cs = gbar / gamma;
  out.sn = beta / gamma;
  out.phi = out.cs * in.phibar;
  out.phibar = out.sn * in.phibar;
  out.denom = 1.0 / gamma;
  out.s = beta > 0 ? 1.0 / beta : 0.0;
  out.ratio = oldb > 0 ? beta / oldb : 0.0;
  if (!(beta > 0)) out.stop = 1;
}

// A(k): scalars, w/x update of iteration k-1 (unless first), then (unless
// final) y = s D A z - ratio r1 and partials of alpha = (s z) . y.
template <int LANES>
__global__ __launch_bounds__(kBlockMinres) void k_minres_a(
    int n, const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
    const float* __restrict__ va, const float* __restrict__ sign, const float* __restrict__ z,
    const float* __restrict__ zprev, const float* __restrict__ r1, float* __restrict__ y,
    float* __restrict__ x, float* __restrict__ wnew, const float* __restrict__ wa,
    const float* __restrict__ wb, int first, int final_only, const double* __restrict__ partial_b,
    int np_b, double* __restrict__ partial_a, const MinresState* __restrict__ sin,
    MinresState* __restrict__ sout) {
  __shared__ double sh[kBlockMinres / 64];
  const double bsum = sum_partials(partial_b, np_b, sh);
  MinresState st;
  minres_scalars(*sin, st, sqrt(bsum > 0 ? bsum : 0), first);
  if (blockIdx.x == 0 && threadIdx.x == 0) *sout = st;
  const float s = (float)st.s;
  const float ratio = first ? 0.f : (float)st.ratio;
  const float sprev = st.oldb > 0 ? (float)(1.0 / st.oldb) : 0.f;
  const float oldeps = (float)st.oldeps, delta = (float)st.delta, denom = (float)st.denom,
              phi = (float)st.phi;
  const int lane = threadIdx.x % LANES;
  double acc = 0;
  const int rows_per_grid = gridDim.x * (kBlockMinres / LANES);
  for (int row = blockIdx.x * (kBlockMinres / LANES) + threadIdx.x / LANES; row < n;
       row += rows_per_grid) {
    float sum = 0.f;
    if (!final_only)
      for (int j = rp[row] + lane, j1 = rp[row + 1]; j < j1; j += 3 * LANES) {
        // three entries per lane per pass, loads issued together (as k_spmv)
        const bool pb = j + LANES < j1, pc = j + 2 * LANES < j1;
        const int ca = ci[j], cb = pb ? ci[j + LANES] : 0, cc = pc ? ci[j + 2 * LANES] : 0;
        const float va_ = va[j], vb = pb ? va[j + LANES] : 0.f, vc = pc ? va[j + 2 * LANES] : 0.f;
        const float za = z[ca], zb = pb ? z[cb] : 0.f, zc = pc ? z[cc] : 0.f;
        sum += va_ * za + vb * zb + vc * zc;
      }
    sum = group_sum<LANES>(sum);
    if (lane == 0) {
      if (!first) {
        const float v = zprev[row] * sprev;
        const float w = (v - oldeps * wa[row] - delta * wb[row]) * denom;
        wnew[row] = w;
        x[row] += phi * w;
      }
      if (!final_only) {
        const float sg = sign ? sign[row] : 1.f;
        const float yi = s * sg * sum - ratio * r1[row];
        y[row] = yi;
        acc += (double)(s * z[row]) * (double)yi;
      }
    }
  }
  const double bs = block_sum(acc, sh);
  if (threadIdx.x == 0) partial_a[blockIdx.x] = bs;
}

// A(k) with the operator as a 3 x 3 block CSR copy (block row R = rows 3R .. 3R+2; block
// j's entry (a, b) at bva[(3a + b) * nblk + j], so that the lanes of a group load each
// entry of consecutive blocks with coalesced loads; three consecutive loads of z per block
// instead of nine scattered ones). Same scalar step and vector updates as k_minres_a.
template <int LANES>
__global__ __launch_bounds__(kBlock) void k_minres_a_b3(
    int nbr, int nblk, const int32_t* __restrict__ brp, const int32_t* __restrict__ bci,
    const float* __restrict__ bva, const float* __restrict__ sign, const float* __restrict__ z,
    const float* __restrict__ zprev, const float* __restrict__ r1, float* __restrict__ y,
    float* __restrict__ x, float* __restrict__ wnew, const float* __restrict__ wa,
    const float* __restrict__ wb, int first, int final_only, const double* __restrict__ partial_b,
    int np_b, double* __restrict__ partial_a, const MinresState* __restrict__ sin,
    MinresState* __restrict__ sout) {
  __shared__ double sh[kBlock / 64];
  const double bsum = sum_partials(partial_b, np_b, sh);
  MinresState st;
  minres_scalars(*sin, st, sqrt(bsum > 0 ? bsum : 0), first);
  if (blockIdx.x == 0 && threadIdx.x == 0) *sout = st;
  const float s = (float)st.s;
  const float ratio = first ? 0.f : (float)st.ratio;
  const float sprev = st.oldb > 0 ? (float)(1.0 / st.oldb) : 0.f;
  const float oldeps = (float)st.oldeps, delta = (float)st.delta, denom = (float)st.denom,
              phi = (float)st.phi;
  const int lane = threadIdx.x % LANES;
  double acc = 0;
  const int groups_per_grid = gridDim.x * (kBlock / LANES);
  for (int br = blockIdx.x * (kBlock / LANES) + threadIdx.x / LANES; br < nbr; br += groups_per_grid) {
    float s0 = 0.f, s1 = 0.f, s2 = 0.f;
    if (!final_only)
      for (int j = brp[br] + lane, j1 = brp[br + 1]; j < j1; j += 2 * LANES) {
        // two blocks per lane per pass, the second predicated (its loads repeat block j
        // when absent, with its input zeroed): a row of up to 2 LANES blocks in one pass
        const bool pb = j + LANES < j1;
        const int jb = pb ? j + LANES : j;
        const int ca = 3 * bci[j], cb = 3 * bci[jb];
        float a[9], b[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) {
          a[k] = bva[k * (int64_t)nblk + j];
          b[k] = bva[k * (int64_t)nblk + jb];
        }
        const float z0 = z[ca], z1 = z[ca + 1], z2 = z[ca + 2];
        const float w0 = pb ? z[cb] : 0.f, w1 = pb ? z[cb + 1] : 0.f, w2 = pb ? z[cb + 2] : 0.f;
        s0 += a[0] * z0 + a[1] * z1 + a[2] * z2 + (b[0] * w0 + b[1] * w1 + b[2] * w2);
        s1 += a[3] * z0 + a[4] * z1 + a[5] * z2 + (b[3] * w0 + b[4] * w1 + b[5] * w2);
        s2 += a[6] * z0 + a[7] * z1 + a[8] * z2 + (b[6] * w0 + b[7] * w1 + b[8] * w2);
      }
    s0 = group_sum<LANES>(s0);
    s1 = group_sum<LANES>(s1);
    s2 = group_sum<LANES>(s2);
    if (lane < 3) {
      const int row = 3 * br + lane;
      const float sum = lane == 0 ? s0 : (lane == 1 ? s1 : s2);
      if (!first) {
        const float v = zprev[row] * sprev;
        const float w = (v - oldeps * wa[row] - delta * wb[row]) * denom;
        wnew[row] = w;
        x[row] += phi * w;
      }
      if (!final_only) {
        const float sg = sign ? sign[row] : 1.f;
        const float yi = s * sg * sum - ratio * r1[row];
        y[row] = yi;
        acc += (double)(s * z[row]) * (double)yi;
      }
    }
  }
  const double bs = block_sum(acc, sh);
  if (threadIdx.x == 0) partial_a[blockIdx.x] = bs;
}

// B(k): alpha_k, y' = y - (alpha/beta) r2 (written into the free buffer ynew),
// z = P^-1 y' with P^-1 a sparse SPD matrix (CSR; y' of other rows is formed on
// the fly from the unmodified y and r2), partials of y' . z.
__global__ __launch_bounds__(kBlockMinres) void k_minres_b(
    int n, const float* __restrict__ y, const float* __restrict__ r2,
    const int32_t* __restrict__ prp, const int32_t* __restrict__ pci,
    const float* __restrict__ pva, float* __restrict__ ynew, float* __restrict__ znew,
    const double* __restrict__ partial_a, int np_a, double* __restrict__ partial_b,
    MinresState* __restrict__ st) {
  __shared__ double sh[kBlockMinres / 64];
  const double alfa = sum_partials(partial_a, np_a, sh);
  const double beta = st->beta;
  const float c = (float)(alfa / beta);
  double acc = 0;
  for (int i = blockIdx.x * kBlockMinres + threadIdx.x; i < n; i += gridDim.x * kBlockMinres) {
    const float yi = y[i] - c * r2[i];
    ynew[i] = yi;
    // three entries per pass, predicated: the loads of a pass are issued together
    float zi = 0.f;
    for (int j = prp[i], j1 = prp[i + 1]; j < j1; j += 3) {
      const bool pb = j + 1 < j1, pc = j + 2 < j1;
      const int ca = pci[j], cb = pb ? pci[j + 1] : ca, cc = pc ? pci[j + 2] : ca;
      const float wa = pva[j], wb = pb ? pva[j + 1] : 0.f, wc = pc ? pva[j + 2] : 0.f;
      const float ya = y[ca] - c * r2[ca], yb = y[cb] - c * r2[cb], yc = y[cc] - c * r2[cc];
      zi += wa * ya + wb * yb + wc * yc;
    }
    znew[i] = zi;
    acc += (double)yi * (double)zi;
  }
  const double bs = block_sum(acc, sh);
  if (threadIdx.x == 0) partial_b[blockIdx.x] = bs;
  // alfa is a field no block of this kernel reads (each sums the partials itself)
  if (blockIdx.x == 0 && threadIdx.x == 0) st->alfa = alfa;
}

// rf = sign * r / scale (FP64 -> FP32)
__global__ __launch_bounds__(kBlock) void k_to_f32(int n, const double* __restrict__ r,
                                                   const float* __restrict__ sign, double scale,
                                                   float* __restrict__ rf) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const double v = r[i] / scale;
  rf[i] = (float)(sign ? sign[i] * v : v);
}

// x += scale * d (FP32 -> FP64)
__global__ __launch_bounds__(kBlock) void k_axpy64(int n, double scale, const float* __restrict__ d,
                                                   double* __restrict__ x) {
  const int i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  x[i] += scale * (double)d[i];
}

// Variants for a preconditioner applied as a function (a sequence of launches between
// B1(k) and B2(k)): init without z, B1 = alpha and y' only, B2 = partials of y' . z.
__global__ __launch_bounds__(kBlock) void k_minres_init_fn(int n, const float* __restrict__ rhs, float* r1,
                                                           float* r2, float* x, float* w0, float* w1, float* w2) {
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
    const float b = rhs[i];
    r1[i] = b;
    r2[i] = b;
    x[i] = 0.f;
    w0[i] = 0.f;
    w1[i] = 0.f;
    w2[i] = 0.f;
  }
}

__global__ __launch_bounds__(kBlock) void k_minres_b1(int n, const float* __restrict__ y,
                                                      const float* __restrict__ r2, float* __restrict__ ynew,
                                                      const double* __restrict__ partial_a, int np_a,
                                                      MinresState* __restrict__ st) {
  __shared__ double sh[kBlock / 64];
  const double alfa = sum_partials(partial_a, np_a, sh);
  const float c = (float)(alfa / st->beta);
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) ynew[i] = y[i] - c * r2[i];
  if (blockIdx.x == 0 && threadIdx.x == 0) st->alfa = alfa;
}

__global__ __launch_bounds__(kBlock) void k_dot_partials(int n, const float* __restrict__ a,
                                                         const float* __restrict__ b, double* __restrict__ partial) {
  __shared__ double sh[kBlock / 64];
  double acc = 0;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) acc += (double)a[i] * (double)b[i];
  const double bs = block_sum(acc, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = bs;
}

// B(k) fused with the first steps of the elasticity two-grid preconditioner: alpha,
// y' = y - (alpha / beta) r2 into ynew, z = D^-1 y' on the stress rows, x1 = W y' on the
// cells (aggregate order), and the partials of y' . z over the stress rows (grid-stride,
// fixed order per thread).
__global__ __launch_bounds__(kBlock) void k_minres_tg_b(int ns, int nu, int n_cells, const float* __restrict__ y,
                                                        const float* __restrict__ r2, float* __restrict__ ynew,
                                                        float* __restrict__ z, const float* __restrict__ dinv,
                                                        const float* __restrict__ W,
                                                        const int32_t* __restrict__ agg_mem, float* __restrict__ x1,
                                                        const double* __restrict__ partial_a, int np_a,
                                                        double* __restrict__ partial_b,
                                                        MinresState* __restrict__ st) {
  __shared__ double sh[kBlock / 64];
  const double alfa = sum_partials(partial_a, np_a, sh);
  const float c = (float)(alfa / st->beta);
  double acc = 0;
  const int total = ns + n_cells;
  for (int g = blockIdx.x * kBlock + threadIdx.x; g < total; g += gridDim.x * kBlock) {
    if (g < ns) {
      const float yv = y[g] - c * r2[g];
      ynew[g] = yv;
      const float zv = el_stress_row(g, dinv, [&](int j) { return y[j] - c * r2[j]; });
      z[g] = zv;
      acc += (double)yv * (double)zv;
    } else {
      const int m = g - ns;
      const int cell = agg_mem[m];
      float rc[6];
#pragma unroll
      for (int i = 0; i < 6; ++i) {
        const int64_t row = el_row(cell, ns, nu, i);
        rc[i] = y[row] - c * r2[row];
        ynew[row] = rc[i];
      }
      el_smooth(m, n_cells, W, rc, x1);
    }
  }
  const double bs = block_sum(acc, sh);
  if (threadIdx.x == 0) partial_b[blockIdx.x] = bs;
  if (blockIdx.x == 0 && threadIdx.x == 0) st->alfa = alfa;
}

// The last step of the two-grid cycle on the cells (z = psign (x2 + W (r - S x2))) and the
// partials of y' . z over the cell rows.
__global__ __launch_bounds__(kBlock) void k_minres_tg_post(int ns, int nu, int n_cells,
                                                           const int32_t* __restrict__ agg_mem,
                                                           const int32_t* __restrict__ nbr,
                                                           const int32_t* __restrict__ agg_of,
                                                           const float* __restrict__ S, const float* __restrict__ W,
                                                           const float* __restrict__ r, const float* __restrict__ x1,
                                                           const float* __restrict__ yc, float psign,
                                                           float* __restrict__ z, double* __restrict__ partial_b) {
  __shared__ double sh[kBlock / 64];
  double acc = 0;
  for (int m = blockIdx.x * kBlock + threadIdx.x; m < n_cells; m += gridDim.x * kBlock) {
    const int cell = agg_mem[m];
    float rv[6], zc[6];
    el_gather(cell, ns, nu, r, rv);
    el_post(m, n_cells, nbr, agg_of, S, W, rv, x1, yc, zc);
#pragma unroll
    for (int i = 0; i < 6; ++i) {
      const float zv = psign * zc[i];
      z[el_row(cell, ns, nu, i)] = zv;
      acc += (double)rv[i] * (double)zv;
    }
  }
  const double bs = block_sum(acc, sh);
  if (threadIdx.x == 0) partial_b[blockIdx.x] = bs;
}

}  // namespace nfx

using namespace nfx;

struct nfx_solver {
  int n = 0;
  int nnz = 0;
  hipStream_t stream = nullptr;
  float* buf = nullptr;  // 10 vectors of n floats
  double* r64 = nullptr;
  double* partial_a = nullptr;
  double* partial_b = nullptr;
  double* partial_r = nullptr;
  double* h_partial = nullptr;  // pinned
  MinresState* st = nullptr;    // 2 states
  // optional 3 x 3 block CSR copy of the FP32 operator (nfx_solver_set_block3)
  int b3_nbr = 0, b3_nblk = 0, blocks_a3 = 1;
  const int32_t* b3_rp = nullptr;
  const int32_t* b3_ci = nullptr;
  const float* b3_va = nullptr;
  MinresState* h_st = nullptr;  // pinned
  int blocks_a = 1, blocks_b = 1;
  int lanes = 4;
  int blocks_r = 1;  // grid of the FP64 residual kernel
};

extern "C" int nfx_spmv(int n_rows, const int32_t* d_rowptr, const int32_t* d_col,
                        const float* d_val, float alpha, const float* d_x, float beta, float* d_y,
                        void* stream) {
  NFX_REQUIRE(n_rows >= 0, "nfx_spmv: n_rows < 0");
  if (n_rows == 0) return NFX_OK;
  NFX_REQUIRE(d_rowptr && d_col && d_val && d_x && d_y, "nfx_spmv: null pointer");
  // 512 threads per block over a grid of at most 1024 blocks (rows in a grid-stride loop):
  // the FGMRES SpMV at 32^3 x 6 (N = 595,968) took the learned solve from 13.8 to 13.55 ms
  // against 256 threads and up to 65,536 blocks
  const int rows_per_block = kBlockMinres / 4;
  int blocks = grid_for(n_rows, rows_per_block);
  if (blocks > 1024) blocks = 1024;
  hipLaunchKernelGGL((k_spmv<4, kBlockMinres>), dim3(blocks), dim3(kBlockMinres), 0, static_cast<hipStream_t>(stream),
                     n_rows, d_rowptr, d_col, d_val, alpha, d_x, beta, d_y);
  NFX_LAUNCH_CHECK();
  return NFX_OK;
}

extern "C" int nfx_csr_diag64(int n_rows, const int32_t* d_rowptr, const int32_t* d_col,
                              const double* d_val, double* d_diag, void* stream) {
  NFX_REQUIRE(n_rows >= 0, "nfx_csr_diag64: n_rows < 0");
  if (n_rows == 0) return NFX_OK;
  NFX_REQUIRE(d_rowptr && d_col && d_val && d_diag, "nfx_csr_diag64: null pointer");
  hipLaunchKernelGGL(k_diag64, dim3(grid_for(n_rows, kBlock)), dim3(kBlock), 0,
                     static_cast<hipStream_t>(stream), n_rows, d_rowptr, d_col, d_val, d_diag);
  NFX_LAUNCH_CHECK();
  return NFX_OK;
}

extern "C" int nfx_schur_diag64(int n_rows, int n_u, const int32_t* d_rowptr, const int32_t* d_col,
                                const double* d_val, const double* d_diag, double* d_schur,
                                void* stream) {
  NFX_REQUIRE(n_rows >= 0 && n_u >= 0 && n_u <= n_rows, "nfx_schur_diag64: sizes");
  if (n_rows == n_u) return NFX_OK;
  NFX_REQUIRE(d_rowptr && d_col && d_val && d_diag && d_schur, "nfx_schur_diag64: null pointer");
  hipLaunchKernelGGL(k_schur64, dim3(grid_for(n_rows - n_u, kBlock)), dim3(kBlock), 0,
                     static_cast<hipStream_t>(stream), n_rows, n_u, d_rowptr, d_col, d_val, d_diag,
                     d_schur);
  NFX_LAUNCH_CHECK();
  return NFX_OK;
}

extern "C" int nfx_solver_create(nfx_solver** out, int n_rows, int nnz, void* stream) {
  NFX_REQUIRE(out && n_rows > 0 && nnz >= 0, "nfx_solver_create: bad arguments");
  nfx_solver* s = new nfx_solver();
  s->n = n_rows;
  s->nnz = nnz;
  s->stream = static_cast<hipStream_t>(stream);
  const double avg = (double)nnz / n_rows;
  s->lanes = residual_lanes(avg);
  s->blocks_r = residual_blocks(n_rows, s->lanes);
  int ba = grid_for(n_rows, kBlockMinres / s->lanes);
  int bb = grid_for(n_rows, kBlockMinres);
  s->blocks_a = ba < kMaxBlocks ? ba : kMaxBlocks;
  s->blocks_b = bb < kMaxBlocks ? bb : kMaxBlocks;
  hipError_t e = hipMalloc(&s->buf, sizeof(float) * (size_t)n_rows * 10);
  if (e == hipSuccess) e = hipMalloc(&s->r64, sizeof(double) * (size_t)n_rows);
  if (e == hipSuccess) e = hipMalloc(&s->partial_a, sizeof(double) * kMaxBlocks);
  if (e == hipSuccess) e = hipMalloc(&s->partial_b, sizeof(double) * kMaxBlocks);
  if (e == hipSuccess) e = hipMalloc(&s->partial_r, sizeof(double) * 2 * kResMaxBlocks);
  if (e == hipSuccess) e = hipHostMalloc(&s->h_partial, sizeof(double) * 2 * kResMaxBlocks, hipHostMallocDefault);
  if (e == hipSuccess) e = hipMalloc(&s->st, 2 * sizeof(MinresState));
  if (e == hipSuccess) e = hipHostMalloc(&s->h_st, sizeof(MinresState), hipHostMallocDefault);
  if (e != hipSuccess) {
    set_error(std::string("nfx_solver_create: ") + hipGetErrorString(e));
    nfx_solver_destroy(s);
    return NFX_E_HIP;
  }
  *out = s;
  return NFX_OK;
}

extern "C" int nfx_solver_set_block3(nfx_solver* s, int n_blocks, const int32_t* d_browptr, const int32_t* d_bcol,
                                     const float* d_bval) {
  NFX_REQUIRE(s, "nfx_solver_set_block3: null solver");
  if (n_blocks <= 0 || !d_browptr || !d_bcol || !d_bval) {
    s->b3_va = nullptr;
    s->b3_rp = s->b3_ci = nullptr;
    s->b3_nbr = s->b3_nblk = 0;
    return NFX_OK;
  }
  NFX_REQUIRE(s->n % 3 == 0, "nfx_solver_set_block3: rows not a multiple of 3");
  s->b3_nbr = s->n / 3;
  s->b3_nblk = n_blocks;
  s->b3_rp = d_browptr;
  s->b3_ci = d_bcol;
  s->b3_va = d_bval;
  const int ba = grid_for(s->b3_nbr, kBlock / 16);
  s->blocks_a3 = ba < kMaxBlocks ? ba : kMaxBlocks;
  return NFX_OK;
}

extern "C" int nfx_solver_destroy(nfx_solver* s) {
  if (!s) return NFX_OK;
  if (s->buf) (void)hipFree(s->buf);
  if (s->r64) (void)hipFree(s->r64);
  if (s->partial_a) (void)hipFree(s->partial_a);
  if (s->partial_b) (void)hipFree(s->partial_b);
  if (s->partial_r) (void)hipFree(s->partial_r);
  if (s->h_partial) (void)hipHostFree(s->h_partial);
  if (s->st) (void)hipFree(s->st);
  if (s->h_st) (void)hipHostFree(s->h_st);
  delete s;
  return NFX_OK;
}

static int launch_a(nfx_solver* s, const int32_t* rp, const int32_t* ci, const float* va,
                    const float* sign, const float* z, const float* zprev, const float* r1,
                    float* y, float* x, float* wnew, const float* wa, const float* wb, int first,
                    int final_only, const MinresState* sin, MinresState* sout, int np_b = -1) {
  const int n = s->n;
  if (np_b < 0) np_b = s->blocks_b;
  if (s->b3_va) {
    hipLaunchKernelGGL(k_minres_a_b3<16>, dim3(s->blocks_a3), dim3(kBlock), 0, s->stream, s->b3_nbr, s->b3_nblk,
                       s->b3_rp, s->b3_ci, s->b3_va, sign, z, zprev, r1, y, x, wnew, wa, wb, first, final_only,
                       s->partial_b, np_b, s->partial_a, sin, sout);
    NFX_LAUNCH_CHECK();
    return NFX_OK;
  }
#define NFX_A(L)                                                                                  \
  hipLaunchKernelGGL(k_minres_a<L>, dim3(s->blocks_a), dim3(kBlockMinres), 0, s->stream, n, rp, ci, va, \
                     sign, z, zprev, r1, y, x, wnew, wa, wb, first, final_only, s->partial_b,     \
                     np_b, s->partial_a, sin, sout)
  if (s->lanes == 16)
    NFX_A(16);
  else if (s->lanes == 8)
    NFX_A(8);
  else
    NFX_A(4);
#undef NFX_A
  NFX_LAUNCH_CHECK();
  return NFX_OK;
}

// One FP32 MINRES solve of (D A) d = rhs (rhs already scaled), result in x.
// Buffers: r1/r2 swap every iteration (P), y lives in F, z swaps (period 2),
// w, w1, w2 rotate with period 3.
static int minres_inner(nfx_solver* s, const int32_t* rp, const int32_t* ci, const float* va,
                        const float* sign, const int32_t* prp, const int32_t* pci,
                        const float* pva, const float* rhs, float* x, double tol, int maxit,
                        int* iters) {
  const int n = s->n;
  float* V = s->buf;
  float* P[2] = {V + (size_t)n * 0, V + (size_t)n * 1};
  float* F = V + (size_t)n * 2;
  float* Z[2] = {V + (size_t)n * 3, V + (size_t)n * 4};
  float* W[3] = {V + (size_t)n * 5, V + (size_t)n * 6, V + (size_t)n * 7};
  MinresState* S[2] = {s->st, s->st + 1};
  hipLaunchKernelGGL(k_minres_init, dim3(s->blocks_b), dim3(kBlock), 0, s->stream, n, rhs, prp,
                     pci, pva, P[0], P[1], Z[1], x, W[0], W[1], W[2], s->partial_b);
  NFX_LAUNCH_CHECK();
  const int check = 16;
  int k = 1;
  for (; k <= maxit; ++k) {
    float* r1 = P[(k + 1) % 2];
    float* r2 = P[k % 2];
    float* z = Z[k % 2];
    float* zprev = Z[(k + 1) % 2];
    float* wnew = W[(k + 2) % 3];      // w_{k-1}
    const float* wa = W[k % 3];        // w_{k-3}
    const float* wb = W[(k + 1) % 3];  // w_{k-2}
    int rc = launch_a(s, rp, ci, va, sign, z, zprev, r1, F, x, wnew, wa, wb, k == 1, 0,
                      S[(k + 1) & 1], S[k & 1]);
    if (rc) return rc;
    // y' goes into r1's buffer (r1 is dead after A(k)); z_{k+1} into zprev's buffer
    hipLaunchKernelGGL(k_minres_b, dim3(s->blocks_b), dim3(kBlockMinres), 0, s->stream, n, F, r2, prp,
                       pci, pva, r1, zprev, s->partial_a, s->b3_va ? s->blocks_a3 : s->blocks_a, s->partial_b,
                       S[k & 1]);
    NFX_LAUNCH_CHECK();
    if (k % check == 0 || k == maxit) {
      NFX_CHECK_HIP(hipMemcpyAsync(s->h_st, S[k & 1], sizeof(MinresState), hipMemcpyDeviceToHost,
                                   s->stream));
      NFX_CHECK_HIP(hipStreamSynchronize(s->stream));
      const double beta1 = s->h_st->beta1;
      // phibar here is the estimate after iteration k-1
      if (s->h_st->stop || !(beta1 > 0) || s->h_st->phibar <= tol * beta1) break;
    }
  }
  if (k > maxit) k = maxit;
  // final: rotation of iteration k and its w/x update (an A kernel without SpMV)
  {
    const int kk = k + 1;
    float* z = Z[kk % 2];
    float* zprev = Z[(kk + 1) % 2];
    float* wnew = W[(kk + 2) % 3];
    const float* wa = W[kk % 3];
    const float* wb = W[(kk + 1) % 3];
    int rc = launch_a(s, rp, ci, va, sign, z, zprev, nullptr, nullptr, x, wnew, wa, wb, 0, 1,
                      S[(kk + 1) & 1], S[kk & 1]);
    if (rc) return rc;
  }
  *iters = k;
  return NFX_OK;
}

// r = b - A x in FP64; returns ||r||^2 and ||b||^2 (host sum of block partials).
static int residual64(nfx_solver* s, const int32_t* rp, const int32_t* ci, const double* va,
                      const double* x, const double* b, double* rr, double* bb) {
  const int blocks = s->blocks_r;
  NFX_CHECK_HIP(launch_residual64(s->lanes, blocks, s->stream, s->n, rp, ci, va, x, b, s->r64, s->partial_r));
  NFX_CHECK_HIP(hipMemcpyAsync(s->h_partial, s->partial_r, sizeof(double) * 2 * blocks,
                               hipMemcpyDeviceToHost, s->stream));
  NFX_CHECK_HIP(hipStreamSynchronize(s->stream));
  double a = 0, c = 0;
  for (int i = 0; i < blocks; ++i) {
    a += s->h_partial[2 * i];
    c += s->h_partial[2 * i + 1];
  }
  *rr = a;
  *bb = c;
  return NFX_OK;
}

extern "C" int nfx_solve_minres_ir(nfx_solver* s, const int32_t* d_rowptr, const int32_t* d_col,
                                   const float* d_val32, const double* d_val64,
                                   const float* d_sign, const int32_t* d_pc_rowptr,
                                   const int32_t* d_pc_col, const float* d_pc_val,
                                   const double* d_b, double* d_x, double rtol, double inner_rtol,
                                   int max_inner, int max_outer, int* h_info, double* h_resid,
                                   double* h_hist, int hist_cap) {
  NFX_REQUIRE(s && d_rowptr && d_col && d_val32 && d_val64 && d_pc_rowptr && d_pc_col &&
                  d_pc_val && d_b && d_x,
              "nfx_solve_minres_ir: null pointer");
  NFX_REQUIRE(max_inner > 0 && max_outer >= 0, "nfx_solve_minres_ir: iteration limits");
  const int n = s->n;
  float* rf = s->buf + (size_t)n * 8;
  float* dx = s->buf + (size_t)n * 9;
  int total = 0, outer = 0;
  double rel = 0;
  const int g = grid_for(n, kBlock);
  for (; outer <= max_outer; ++outer) {
    double rr = 0, bb = 0;
    int rc = residual64(s, d_rowptr, d_col, d_val64, d_x, d_b, &rr, &bb);
    if (rc) return rc;
    const double rn = sqrt(rr), bn = sqrt(bb);
    rel = bn > 0 ? rn / bn : rn;
    if (h_hist && outer < hist_cap) h_hist[outer] = rel;
    if (!(rel == rel) || rel <= rtol || outer == max_outer || rn == 0) break;
    hipLaunchKernelGGL(k_to_f32, dim3(g), dim3(kBlock), 0, s->stream, n, s->r64, d_sign, rn, rf);
    NFX_LAUNCH_CHECK();
    int it = 0;
    rc = minres_inner(s, d_rowptr, d_col, d_val32, d_sign, d_pc_rowptr, d_pc_col, d_pc_val, rf, dx,
                      inner_rtol, max_inner, &it);
    if (rc) return rc;
    total += it;
    hipLaunchKernelGGL(k_axpy64, dim3(g), dim3(kBlock), 0, s->stream, n, rn, dx, d_x);
    NFX_LAUNCH_CHECK();
  }
  NFX_CHECK_HIP(hipStreamSynchronize(s->stream));
  if (h_info) {
    h_info[0] = outer;
    h_info[1] = total;
  }
  if (h_resid) h_resid[0] = rel;
  return rel <= rtol ? NFX_OK : NFX_E_NOT_CONVERGED;
}

// minres_inner with the elasticity two-grid preconditioner: B(k) fused with its first
// steps (k_minres_tg_b), the coarse correction, then its last step fused with the second
// half of the y' . z partials (k_minres_tg_post); the partials of the two kernels follow
// each other in partial_b.
static int minres_inner_tg(nfx_solver* s, const int32_t* rp, const int32_t* ci, const float* va,
                           const nfx_elasticity_tg* pc, const float* rhs, float* x, double tol, int maxit,
                           int* iters) {
  const int n = s->n;
  float* V = s->buf;
  float* P[2] = {V + (size_t)n * 0, V + (size_t)n * 1};
  float* F = V + (size_t)n * 2;
  float* Z[2] = {V + (size_t)n * 3, V + (size_t)n * 4};
  float* W[3] = {V + (size_t)n * 5, V + (size_t)n * 6, V + (size_t)n * 7};
  MinresState* S[2] = {s->st, s->st + 1};
  const int ns = pc->n_s, nu = pc->n_u, E = pc->n_cells;
  const int b1 = std::min(grid_for(ns + E, kBlock), kMaxBlocks / 2);
  const int b2 = std::min(grid_for(E, kBlock), kMaxBlocks / 2);
  hipLaunchKernelGGL(k_minres_init_fn, dim3(s->blocks_b), dim3(kBlock), 0, s->stream, n, rhs, P[0], P[1], x, W[0],
                     W[1], W[2]);
  NFX_LAUNCH_CHECK();
  int rc = el_twogrid_apply_scaled(pc, rhs, Z[1], nullptr, s->stream);
  if (rc) return rc;
  hipLaunchKernelGGL(k_dot_partials, dim3(s->blocks_b), dim3(kBlock), 0, s->stream, n, rhs, Z[1], s->partial_b);
  NFX_LAUNCH_CHECK();
  int np_b = s->blocks_b;
  const int check = 16;
  int k = 1;
  for (; k <= maxit; ++k) {
    float* r1 = P[(k + 1) % 2];
    float* r2 = P[k % 2];
    float* z = Z[k % 2];
    float* zprev = Z[(k + 1) % 2];
    float* wnew = W[(k + 2) % 3];
    const float* wa = W[k % 3];
    const float* wb = W[(k + 1) % 3];
    rc = launch_a(s, rp, ci, va, nullptr, z, zprev, r1, F, x, wnew, wa, wb, k == 1, 0, S[(k + 1) & 1], S[k & 1],
                  np_b);
    if (rc) return rc;
    // y' into r1's buffer, z_{k+1} = P^-1 y' into zprev's buffer
    hipLaunchKernelGGL(k_minres_tg_b, dim3(b1), dim3(kBlock), 0, s->stream, ns, nu, E, F, r2, r1, zprev, pc->dinv,
                       pc->winv, pc->agg_mem, el_tg_x1(pc), s->partial_a, s->b3_va ? s->blocks_a3 : s->blocks_a,
                       s->partial_b, S[k & 1]);
    NFX_LAUNCH_CHECK();
    rc = el_tg_launch_mid(pc, r1, s->stream);
    if (rc) return rc;
    hipLaunchKernelGGL(k_minres_tg_post, dim3(b2), dim3(kBlock), 0, s->stream, ns, nu, E, pc->agg_mem, pc->nbr,
                       pc->agg_of, pc->sblk, pc->winv, r1, el_tg_x1(pc), el_tg_y(pc), pc->psign, zprev,
                       s->partial_b + b1);
    NFX_LAUNCH_CHECK();
    np_b = b1 + b2;
    if (k % check == 0 || k == maxit) {
      NFX_CHECK_HIP(hipMemcpyAsync(s->h_st, S[k & 1], sizeof(MinresState), hipMemcpyDeviceToHost, s->stream));
      NFX_CHECK_HIP(hipStreamSynchronize(s->stream));
      const double beta1 = s->h_st->beta1;
      if (s->h_st->stop || !(beta1 > 0) || s->h_st->phibar <= tol * beta1) break;
    }
  }
  if (k > maxit) k = maxit;
  {
    const int kk = k + 1;
    float* z = Z[kk % 2];
    float* zprev = Z[(kk + 1) % 2];
    float* wnew = W[(kk + 2) % 3];
    const float* wa = W[kk % 3];
    const float* wb = W[(kk + 1) % 3];
    rc = launch_a(s, rp, ci, va, nullptr, z, zprev, nullptr, nullptr, x, wnew, wa, wb, 0, 1, S[(kk + 1) & 1],
                  S[kk & 1], np_b);
    if (rc) return rc;
  }
  *iters = k;
  return NFX_OK;
}

extern "C" int nfx_solve_minres_ir_tg(nfx_solver* s, const int32_t* d_rowptr, const int32_t* d_col,
                                      const float* d_val32, const double* d_val64, const nfx_elasticity_tg* pc,
                                      const double* d_b, double* d_x, double rtol, double inner_rtol, int max_inner,
                                      int max_outer, int* h_info, double* h_resid, double* h_hist, int hist_cap) {
  NFX_REQUIRE(s && d_rowptr && d_col && d_val32 && d_val64 && pc && d_b && d_x, "nfx_solve_minres_ir_tg: null pointer");
  NFX_REQUIRE(max_inner > 0 && max_outer >= 0, "nfx_solve_minres_ir_tg: iteration limits");
  NFX_REQUIRE(pc->n_s + 2 * pc->n_u == s->n, "nfx_solve_minres_ir_tg: preconditioner size differs from the matrix");
  const int n = s->n;
  float* rf = s->buf + (size_t)n * 8;
  float* dx = s->buf + (size_t)n * 9;
  int total = 0, outer = 0;
  double rel = 0;
  const int g = grid_for(n, kBlock);
  for (; outer <= max_outer; ++outer) {
    double rr = 0, bb = 0;
    int rc = residual64(s, d_rowptr, d_col, d_val64, d_x, d_b, &rr, &bb);
    if (rc) return rc;
    const double rn = sqrt(rr), bn = sqrt(bb);
    rel = bn > 0 ? rn / bn : rn;
    if (h_hist && outer < hist_cap) h_hist[outer] = rel;
    if (!(rel == rel) || rel <= rtol || outer == max_outer || rn == 0) break;
    hipLaunchKernelGGL(k_to_f32, dim3(g), dim3(kBlock), 0, s->stream, n, s->r64, nullptr, rn, rf);
    NFX_LAUNCH_CHECK();
    int it = 0;
    rc = minres_inner_tg(s, d_rowptr, d_col, d_val32, pc, rf, dx, inner_rtol, max_inner, &it);
    if (rc) return rc;
    total += it;
    hipLaunchKernelGGL(k_axpy64, dim3(g), dim3(kBlock), 0, s->stream, n, rn, dx, d_x);
    NFX_LAUNCH_CHECK();
  }
  NFX_CHECK_HIP(hipStreamSynchronize(s->stream));
  if (h_info) {
    h_info[0] = outer;
    h_info[1] = total;
  }
  if (h_resid) h_resid[0] = rel;
  return rel <= rtol ? NFX_OK : NFX_E_NOT_CONVERGED;
}

extern "C" int nfx_residual64(int n_rows, const int32_t* d_rowptr, const int32_t* d_col,
                              const double* d_val, const double* d_x, const double* d_b,
                              double* d_r, void* stream) {
  NFX_REQUIRE(n_rows > 0, "nfx_residual64: n_rows");
  NFX_REQUIRE(d_rowptr && d_col && d_val && d_x && d_b && d_r, "nfx_residual64: null pointer");
  hipStream_t sm = static_cast<hipStream_t>(stream);
  double* partial = nullptr;
  // 8 lanes per row: the row lengths are not read here (no synchronisation)
  const int blocks = residual_blocks(n_rows, kResidualLanesDefault);
  NFX_CHECK_HIP(hipMallocAsync(reinterpret_cast<void**>(&partial), sizeof(double) * 2 * blocks, sm));
  NFX_CHECK_HIP(launch_residual64(kResidualLanesDefault, blocks, sm, n_rows, d_rowptr, d_col, d_val, d_x,
                                  d_b, d_r, partial));
  NFX_CHECK_HIP(hipFreeAsync(partial, sm));
  return NFX_OK;
}

// ---------------------------------------------------------------------------
// Batched MINRES: many independent systems, one set of launches per iteration.
//
// Every instance keeps its own MINRES recurrence (scalar state, stopping test, iteration
// count); the vectors of all instances are concatenated (instance i's rows at row_off[i]).
// Work is cut into tiles of kBmTile rows of one instance, tiles of an instance contiguous;
// a kernel over the tiles writes one partial per tile, and a reduction kernel (one block
// per instance) sums an instance's partials in tile order, so every result is reproducible
// and no block sums more than its instance's partials once. Per iteration: A (w/x update,
// y = s D A z - ratio r1, partials of alpha), reduce + alpha, B (y', z = P^-1 y' with a
// diagonal P^-1, partials of y'.z), reduce + the scalar step of the next A. An instance's
// mode: 0 iterating, 1 final update pending (its next A applies the last rotation only),
// 2 frozen. The host checks the states every kBmCheck iterations; the tile kernels run over
// the list of tiles of unfrozen instances (rebuilt when the set changes).

namespace nfx {

// rows per tile (one thread per row; 4, 8 or 16 lanes per row in A). 256 against 64 and 128
// on a 32-instance batch (2.66 M rows): plain batched MINRES 346 / 323 / 321 ms, two-grid
// 89.6 / 87.9 / 86.6 ms (fewer blocks, each with its own state loads and block sum)
constexpr int kBmTile = 256;
constexpr int kBmCheck = 16;  // iterations between host checks

__global__ __launch_bounds__(kBlock) void k_bm_init(const nfx_minres_batch_item* __restrict__ items,
                                                    const int32_t* __restrict__ tiles,
                                                    const int32_t* __restrict__ act,
                                                    const int32_t* __restrict__ mode,
                                                    const float* __restrict__ rhs, float* r1, float* r2, float* z,
                                                    float* x, float* w0, float* w1, float* w2,
                                                    double* __restrict__ partial_b) {
  __shared__ double sh[kBlock / 64];
  const int t = act[blockIdx.x], inst = tiles[2 * t], row0 = tiles[2 * t + 1];
  double acc = 0;
  if (mode[inst] == 0) {
    const nfx_minres_batch_item it = items[inst];
    const int row = row0 + threadIdx.x;
    if (row < it.n_rows) {
      const int64_t g = it.row_off + row;
      const float b = rhs[g];
      const float zz = it.pinv ? it.pinv[row] * b : b;
      r1[g] = b;
      r2[g] = b;
      z[g] = zz;
      x[g] = 0.f;
      w0[g] = 0.f;
      w1[g] = 0.f;
      w2[g] = 0.f;
      acc = (double)b * (double)zz;
    }
  }
  const double bs = block_sum(acc, sh);
  if (threadIdx.x == 0) partial_b[t] = bs;
}

// A(k) of every instance in modes 0 (full) and 1 (final update only)
// rows of one tile of A(k) (LANES lanes per row, kBlock / LANES rows per pass)
template <int LANES>
__device__ __forceinline__ double bm_a_rows(const nfx_minres_batch_item& it, int row0, bool first, bool final_only,
                                            float s, float ratio, float sprev, float oldeps, float delta,
                                            float denom, float phi, const float* __restrict__ z,
                                            const float* __restrict__ zprev, const float* __restrict__ r1,
                                            float* __restrict__ y, float* __restrict__ x, float* __restrict__ wnew,
                                            const float* __restrict__ wa, const float* __restrict__ wb) {
  double acc = 0;
  const int lane = threadIdx.x % LANES;
  const float* zi = z + it.row_off;
  for (int rs = threadIdx.x / LANES; rs < kBmTile; rs += kBlock / LANES) {
    const int row = row0 + rs;
    float sum = 0.f;
    if (row < it.n_rows && !final_only)
      for (int j = it.rowptr[row] + lane, j1 = it.rowptr[row + 1]; j < j1; j += 3 * LANES) {
        const bool pb = j + LANES < j1, pc = j + 2 * LANES < j1;
        const int ca = it.col[j], cb = pb ? it.col[j + LANES] : 0, cc = pc ? it.col[j + 2 * LANES] : 0;
        const float va_ = it.val32[j], vb = pb ? it.val32[j + LANES] : 0.f, vc = pc ? it.val32[j + 2 * LANES] : 0.f;
        const float za = zi[ca], zb = pb ? zi[cb] : 0.f, zc = pc ? zi[cc] : 0.f;
        sum += va_ * za + vb * zb + vc * zc;
      }
    sum = group_sum<LANES>(sum);
    if (lane == 0 && row < it.n_rows) {
      const int64_t g = it.row_off + row;
      if (!first) {
        const float v = zprev[g] * sprev;
        const float w = (v - oldeps * wa[g] - delta * wb[g]) * denom;
        wnew[g] = w;
        x[g] += phi * w;
      }
      if (!final_only) {
        const float sg = it.sign ? it.sign[row] : 1.f;
        const float yi = s * sg * sum - ratio * r1[g];
        y[g] = yi;
        acc += (double)(s * z[g]) * (double)yi;
      }
    }
  }
  return acc;
}

// A(k) of every instance in modes 0 (full) and 1 (final update only)
__global__ __launch_bounds__(kBlock) void k_bm_a(const nfx_minres_batch_item* __restrict__ items,
                                                 const int32_t* __restrict__ tiles, const int32_t* __restrict__ act,
                                                 const int32_t* __restrict__ mode,
                                                 const MinresState* __restrict__ st, int first,
                                                 const float* __restrict__ z, const float* __restrict__ zprev,
                                                 const float* __restrict__ r1, float* __restrict__ y,
                                                 float* __restrict__ x, float* __restrict__ wnew,
                                                 const float* __restrict__ wa, const float* __restrict__ wb,
                                                 double* __restrict__ partial_a) {
  __shared__ double sh[kBlock / 64];
  const int t = act[blockIdx.x], inst = tiles[2 * t], row0 = tiles[2 * t + 1];
  const int md = mode[inst];
  double acc = 0;
  if (md <= 1) {
    const nfx_minres_batch_item it = items[inst];
    const MinresState s0 = st[inst];
    const bool final_only = md == 1;
    const float s = (float)s0.s;
    const float ratio = first ? 0.f : (float)s0.ratio;
    const float sprev = s0.oldb > 0 ? (float)(1.0 / s0.oldb) : 0.f;
    const float oldeps = (float)s0.oldeps, delta = (float)s0.delta, denom = (float)s0.denom, phi = (float)s0.phi;
    // the lane count is uniform over the block (one instance per tile)
    if (it.lanes == 16)
      acc = bm_a_rows<16>(it, row0, first, final_only, s, ratio, sprev, oldeps, delta, denom, phi, z, zprev, r1, y, x,
                          wnew, wa, wb);
    else if (it.lanes == 8)
      acc = bm_a_rows<8>(it, row0, first, final_only, s, ratio, sprev, oldeps, delta, denom, phi, z, zprev, r1, y, x,
                         wnew, wa, wb);
    else
      acc = bm_a_rows<4>(it, row0, first, final_only, s, ratio, sprev, oldeps, delta, denom, phi, z, zprev, r1, y, x,
                         wnew, wa, wb);
  }
  const double bs = block_sum(acc, sh);
  if (threadIdx.x == 0) partial_a[t] = bs;
}

// B(k): y' = y - (alpha / beta) r2 (into ynew), z = P^-1 y', partials of y' . z
__global__ __launch_bounds__(kBlock) void k_bm_b(const nfx_minres_batch_item* __restrict__ items,
                                                 const int32_t* __restrict__ tiles, const int32_t* __restrict__ act,
                                                 const int32_t* __restrict__ mode,
                                                 const MinresState* __restrict__ st, const float* __restrict__ y,
                                                 const float* __restrict__ r2, float* __restrict__ ynew,
                                                 float* __restrict__ znew, double* __restrict__ partial_b) {
  __shared__ double sh[kBlock / 64];
  const int t = act[blockIdx.x], inst = tiles[2 * t], row0 = tiles[2 * t + 1];
  double acc = 0;
  if (mode[inst] == 0) {
    const nfx_minres_batch_item it = items[inst];
    const float c = (float)(st[inst].alfa / st[inst].beta);
    const int row = row0 + threadIdx.x;
    if (row < it.n_rows) {
      const int64_t g = it.row_off + row;
      const float yi = y[g] - c * r2[g];
      ynew[g] = yi;
      const float zi = it.pinv ? it.pinv[row] * yi : yi;
      znew[g] = zi;
      acc = (double)yi * (double)zi;
    }
  }
  const double bs = block_sum(acc, sh);
  if (threadIdx.x == 0) partial_b[t] = bs;
}

// Tile partials of y'.z after a preconditioner that is not the fused diagonal (the
// two-grid groups): same tiles and order as k_bm_b.
__global__ __launch_bounds__(kBlock) void k_bm_dot(const nfx_minres_batch_item* __restrict__ items,
                                                   const int32_t* __restrict__ tiles, const int32_t* __restrict__ act,
                                                   const int32_t* __restrict__ mode, const float* __restrict__ y,
                                                   const float* __restrict__ z, double* __restrict__ partial_b) {
  __shared__ double sh[kBlock / 64];
  const int t = act[blockIdx.x], inst = tiles[2 * t], row0 = tiles[2 * t + 1];
  double acc = 0;
  if (mode[inst] == 0) {
    const nfx_minres_batch_item it = items[inst];
    const int row = row0 + threadIdx.x;
    if (row < it.n_rows) {
      const int64_t g = it.row_off + row;
      acc = (double)y[g] * (double)z[g];
    }
  }
  const double bs = block_sum(acc, sh);
  if (threadIdx.x == 0) partial_b[t] = bs;
}

// One block per instance: the sum of its tile partials in tile order. what = 0: alpha of
// A(k) (and mode 1 -> 2: the final update has been applied); 1: the scalar step that A(k+1)
// reads (st_out from st_in); 2: nothing but the sum (sum_out).
__global__ __launch_bounds__(kBlock) void k_bm_reduce(int n_inst, const int32_t* __restrict__ tile_ptr,
                                                      const double* __restrict__ partial, int what, int first,
                                                      int32_t* __restrict__ mode, const MinresState* __restrict__ st_in,
                                                      MinresState* __restrict__ st_out, double* __restrict__ sum_out) {
  __shared__ double sh[kBlock / 64];
  const int inst = blockIdx.x;
  if (inst >= n_inst) return;
  double v = 0;
  for (int t = tile_ptr[inst] + threadIdx.x; t < tile_ptr[inst + 1]; t += kBlock) v += partial[t];
  v = block_sum(v, sh);
  if (threadIdx.x != 0) return;
  if (what == 2) {
    sum_out[inst] = v;
    return;
  }
  const int md = mode[inst];
  if (what == 0) {
    if (md == 0) st_out[inst].alfa = v;
    if (md == 1) mode[inst] = 2;
    return;
  }
  if (md <= 1) {
    MinresState s;
    minres_scalars(st_in[inst], s, sqrt(v > 0 ? v : 0), first);
    st_out[inst] = s;
  }
}

// FP64 residual rows of one tile (LANES lanes per row); rr / bb accumulate r.r and b.b
template <int LANES>
__device__ __forceinline__ void bm_res_rows(const nfx_minres_batch_item& it, int row0, const double* __restrict__ x,
                                            const double* __restrict__ b, double* __restrict__ r, double& rr,
                                            double& bb) {
  const int lane = threadIdx.x % LANES;
  const double* xi = x + it.row_off;
  for (int rs = threadIdx.x / LANES; rs < kBmTile; rs += kBlock / LANES) {
    const int row = row0 + rs;
    double s = 0;
    if (row < it.n_rows)
      for (int j = it.rowptr[row] + lane; j < it.rowptr[row + 1]; j += LANES) s += it.val64[j] * xi[it.col[j]];
#pragma unroll
    for (int off = LANES / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, LANES);
    if (lane == 0 && row < it.n_rows) {
      const int64_t g = it.row_off + row;
      const double ri = b[g] - s;
      r[g] = ri;
      rr += ri * ri;
      bb += b[g] * b[g];
    }
  }
}

// FP64 residual r = b - A x of every instance whose mode is 0, partials of r.r and b.b
__global__ __launch_bounds__(kBlock) void k_bm_res64(const nfx_minres_batch_item* __restrict__ items,
                                                     const int32_t* __restrict__ tiles,
                                                     const int32_t* __restrict__ act,
                                                     const int32_t* __restrict__ mode, const double* __restrict__ x,
                                                     const double* __restrict__ b, double* __restrict__ r,
                                                     double* __restrict__ prr, double* __restrict__ pbb) {
  __shared__ double sh[kBlock / 64];
  const int t = act[blockIdx.x], inst = tiles[2 * t], row0 = tiles[2 * t + 1];
  double rr = 0, bb = 0;
  if (mode[inst] == 0) {
    const nfx_minres_batch_item it = items[inst];
    if (it.lanes == 16)
      bm_res_rows<16>(it, row0, x, b, r, rr, bb);
    else if (it.lanes == 8)
      bm_res_rows<8>(it, row0, x, b, r, rr, bb);
    else
      bm_res_rows<4>(it, row0, x, b, r, rr, bb);
  }
  const double a = block_sum(rr, sh);
  const double c = block_sum(bb, sh);
  if (threadIdx.x == 0) {
    prr[t] = a;
    pbb[t] = c;
  }
}

// rhs = D r / scale (FP32) and, after the inner solve, x += scale dx (FP64), instances in
// mode 0 only
__global__ __launch_bounds__(kBlock) void k_bm_scale(const nfx_minres_batch_item* __restrict__ items,
                                                     const int32_t* __restrict__ tiles,
                                                     const int32_t* __restrict__ act,
                                                     const int32_t* __restrict__ mode,
                                                     const double* __restrict__ scale, int dir,
                                                     const double* __restrict__ r, float* __restrict__ rhs,
                                                     const float* __restrict__ dx, double* __restrict__ x) {
  const int t = act[blockIdx.x], inst = tiles[2 * t], row0 = tiles[2 * t + 1];
  if (mode[inst] != 0) return;
  const nfx_minres_batch_item it = items[inst];
  const int row = row0 + threadIdx.x;
  if (row >= it.n_rows) return;
  const int64_t g = it.row_off + row;
  if (dir == 0) {
    const double v = r[g] / scale[inst];
    rhs[g] = (float)(it.sign ? it.sign[row] * v : v);
  } else {
    x[g] += scale[inst] * (double)dx[g];
  }
}

}  // namespace nfx

struct nfx_minres_batch {
  int n_inst = 0, n_tiles = 0;
  int64_t n_total = 0;
  hipStream_t stream = nullptr;
  nfx_minres_batch_item* items = nullptr;  // device copy
  int32_t* tiles = nullptr;
  int32_t* tile_ptr = nullptr;
  int32_t* act = nullptr;    // tile ids the kernels run over
  int32_t* h_act = nullptr;  // pinned staging of act
  std::vector<int32_t> h_tile_ptr;
  int n_act = 0;
  int32_t* mode = nullptr;
  float* buf = nullptr;  // 10 vectors of n_total floats
  double* r64 = nullptr;
  double* pa = nullptr;  // n_tiles partials (alpha / r.r)
  double* pb = nullptr;  // n_tiles partials (y'.z / b.b)
  double* sums = nullptr;   // 2 n_inst
  double* scale = nullptr;  // n_inst
  MinresState* st = nullptr;  // 2 n_inst
  MinresState* h_st = nullptr;
  int32_t* h_mode = nullptr;
  double* h_sums = nullptr;
  double* h_scale = nullptr;
  std::vector<nfx_batch_tg> pc;  // two-grid groups on the pressure rows (optional)
};

static void bm_free(nfx_minres_batch* s) {
  if (!s) return;
  (void)hipFree(s->items);
  (void)hipFree(s->tiles);
  (void)hipFree(s->tile_ptr);
  (void)hipFree(s->act);
  (void)hipHostFree(s->h_act);
  (void)hipFree(s->mode);
  (void)hipFree(s->buf);
  (void)hipFree(s->r64);
  (void)hipFree(s->pa);
  (void)hipFree(s->pb);
  (void)hipFree(s->sums);
  (void)hipFree(s->scale);
  (void)hipFree(s->st);
  (void)hipHostFree(s->h_st);
  (void)hipHostFree(s->h_mode);
  (void)hipHostFree(s->h_sums);
  (void)hipHostFree(s->h_scale);
  delete s;
}

extern "C" int nfx_minres_batch_create(nfx_minres_batch** out, int n_inst, const nfx_minres_batch_item* h_items,
                                       void* stream) {
  NFX_REQUIRE(out && n_inst > 0 && h_items, "nfx_minres_batch_create: bad arguments");
  std::vector<nfx_minres_batch_item> items(h_items, h_items + n_inst);
  std::vector<int32_t> tiles, tile_ptr(1, 0);
  int64_t off = 0;
  for (int i = 0; i < n_inst; ++i) {
    auto& it = items[i];
    NFX_REQUIRE(it.n_rows > 0 && it.rowptr && it.col && it.val32 && it.val64,
                "nfx_minres_batch_create: incomplete instance");
    it.row_off = off;
    off += it.n_rows;
    // 4 lanes per row for every instance: on a batch of 32 RT1 / RT2 instances (4..16 cubes)
    // 8 and 16 lanes for the longer RT2 rows (the single-system choice) made the batched solve
    // 1.5 x slower (a tile's rows then take several passes, each ending in a one-lane update)
    it.lanes = 4;
    for (int r = 0; r < it.n_rows; r += kBmTile) {
      tiles.push_back(i);
      tiles.push_back(r);
    }
    NFX_REQUIRE(tiles.size() / 2 < (size_t)INT32_MAX, "nfx_minres_batch_create: too many tiles");
    tile_ptr.push_back((int32_t)(tiles.size() / 2));
  }
  nfx_minres_batch* s = new nfx_minres_batch();
  s->n_inst = n_inst;
  s->n_tiles = (int)(tiles.size() / 2);
  s->n_total = off;
  s->stream = static_cast<hipStream_t>(stream);
  hipError_t e = hipMalloc(&s->items, sizeof(nfx_minres_batch_item) * n_inst);
  if (e == hipSuccess) e = hipMalloc(&s->tiles, sizeof(int32_t) * tiles.size());
  if (e == hipSuccess) e = hipMalloc(&s->tile_ptr, sizeof(int32_t) * tile_ptr.size());
  if (e == hipSuccess) e = hipMalloc(&s->mode, sizeof(int32_t) * n_inst);
  if (e == hipSuccess) e = hipMalloc(&s->act, sizeof(int32_t) * s->n_tiles);
  if (e == hipSuccess) e = hipHostMalloc(&s->h_act, sizeof(int32_t) * s->n_tiles, hipHostMallocDefault);
  if (e == hipSuccess) e = hipMalloc(&s->buf, sizeof(float) * (size_t)off * 10);
  if (e == hipSuccess) e = hipMalloc(&s->r64, sizeof(double) * (size_t)off);
  if (e == hipSuccess) e = hipMalloc(&s->pa, sizeof(double) * s->n_tiles);
  if (e == hipSuccess) e = hipMalloc(&s->pb, sizeof(double) * s->n_tiles);
  if (e == hipSuccess) e = hipMalloc(&s->sums, sizeof(double) * 2 * n_inst);
  if (e == hipSuccess) e = hipMalloc(&s->scale, sizeof(double) * n_inst);
  if (e == hipSuccess) e = hipMalloc(&s->st, sizeof(MinresState) * 2 * n_inst);
  if (e == hipSuccess) e = hipHostMalloc(&s->h_st, sizeof(MinresState) * n_inst, hipHostMallocDefault);
  if (e == hipSuccess) e = hipHostMalloc(&s->h_mode, sizeof(int32_t) * n_inst, hipHostMallocDefault);
  if (e == hipSuccess) e = hipHostMalloc(&s->h_sums, sizeof(double) * 2 * n_inst, hipHostMallocDefault);
  if (e == hipSuccess) e = hipHostMalloc(&s->h_scale, sizeof(double) * n_inst, hipHostMallocDefault);
  if (e == hipSuccess)
    e = hipMemcpy(s->items, items.data(), sizeof(nfx_minres_batch_item) * n_inst, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(s->tiles, tiles.data(), sizeof(int32_t) * tiles.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(s->tile_ptr, tile_ptr.data(), sizeof(int32_t) * tile_ptr.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    set_error(std::string("nfx_minres_batch_create: ") + hipGetErrorString(e));
    bm_free(s);
    return NFX_E_HIP;
  }
  s->h_tile_ptr = tile_ptr;
  *out = s;
  return NFX_OK;
}

// act = the tiles of the instances whose host mode is < limit (uploaded; the caller's stream
// must have finished with the previous list: it is synchronised before every rebuild)
static int bm_set_active(nfx_minres_batch* s, int limit) {
  NFX_CHECK_HIP(hipStreamSynchronize(s->stream));  // no copy from h_act is pending
  int n = 0;
  for (int i = 0; i < s->n_inst; ++i)
    if (s->h_mode[i] < limit)
      for (int t = s->h_tile_ptr[i]; t < s->h_tile_ptr[i + 1]; ++t) s->h_act[n++] = t;
  s->n_act = n;
  if (n) NFX_CHECK_HIP(hipMemcpyAsync(s->act, s->h_act, sizeof(int32_t) * n, hipMemcpyHostToDevice, s->stream));
  return NFX_OK;
}

extern "C" int nfx_minres_batch_destroy(nfx_minres_batch* s) {
  bm_free(s);
  return NFX_OK;
}

extern "C" int64_t nfx_minres_batch_rows(const nfx_minres_batch* s) { return s ? s->n_total : -1; }

extern "C" int nfx_minres_batch_set_pc(nfx_minres_batch* s, int n_groups, const nfx_batch_tg* h_groups) {
  NFX_REQUIRE(s && n_groups >= 0 && (n_groups == 0 || h_groups), "nfx_minres_batch_set_pc: bad arguments");
  NFX_REQUIRE(n_groups <= 2 && (n_groups < 2 || h_groups[0].nb != h_groups[1].nb),
              "nfx_minres_batch_set_pc: at most two groups, of different nb");
  std::vector<nfx_batch_tg> g(h_groups, h_groups + n_groups);
  for (const nfx_batch_tg& d : g) {
    NFX_REQUIRE(d.nb == 1 || d.nb == 4, "nfx_minres_batch_set_pc: nb must be 1 or 4");
    NFX_REQUIRE(d.n_cells > 0 && d.n_agg > 0 && d.cell_inst && d.prow && d.work,
                "nfx_minres_batch_set_pc: incomplete group");
  }
  s->pc.swap(g);
  return NFX_OK;
}

// The two-grid groups on top of the diagonal preconditioner k_bm_init / k_bm_b applied:
// z on the groups' rows of active instances from y, then the tile partials of y'.z again.
static int bm_apply_pc(nfx_minres_batch* s, const float* y, float* z) {
  if (s->pc.empty() || !s->n_act) return NFX_OK;
  const int rc = nfx::batch_tg_apply(s->pc.data(), (int)s->pc.size(), s->mode, y, z, s->stream);
  if (rc) return rc;
  hipLaunchKernelGGL(k_bm_dot, dim3(s->n_act), dim3(kBmTile), 0, s->stream, s->items, s->tiles, s->act, s->mode, y,
                     z, s->pb);
  NFX_LAUNCH_CHECK();
  return NFX_OK;
}

// One batched FP32 MINRES on every instance whose h_mode is 0 on entry (rhs already scaled);
// iters[i] gets the instance's iteration count.

static int bm_inner(nfx_minres_batch* s, const float* rhs, float* x, const double* tol, int maxit, int* iters) {
  const int64_t N = s->n_total;
  float* V = s->buf;
  float* P[2] = {V + N * 0, V + N * 1};
  float* F = V + N * 2;
  float* Z[2] = {V + N * 3, V + N * 4};
  float* W[3] = {V + N * 5, V + N * 6, V + N * 7};
  MinresState* S[2] = {s->st, s->st + s->n_inst};
  const int m = s->n_inst;
  hipStream_t st = s->stream;
  NFX_CHECK_HIP(hipMemcpyAsync(s->mode, s->h_mode, sizeof(int32_t) * m, hipMemcpyHostToDevice, st));
  int rc = bm_set_active(s, 1);
  if (rc) return rc;
  if (s->n_act) {
    hipLaunchKernelGGL(k_bm_init, dim3(s->n_act), dim3(kBmTile), 0, st, s->items, s->tiles, s->act, s->mode, rhs,
                       P[0], P[1], Z[1], x, W[0], W[1], W[2], s->pb);
    rc = bm_apply_pc(s, P[1], Z[1]);
    if (rc) return rc;
  }
  hipLaunchKernelGGL(k_bm_reduce, dim3(m), dim3(kBlock), 0, st, m, s->tile_ptr, s->pb, 1, 1, s->mode, S[0], S[1],
                     nullptr);
  NFX_LAUNCH_CHECK();
  std::vector<int> active;
  for (int i = 0; i < m; ++i) {
    iters[i] = 0;
    if (s->h_mode[i] == 0) active.push_back(i);
  }
  int k = 1;
  bool pending_final = false;
  for (; !active.empty() || pending_final; ++k) {
    float* r1 = P[(k + 1) % 2];
    float* r2 = P[k % 2];
    float* z = Z[k % 2];
    float* zprev = Z[(k + 1) % 2];
    float* wnew = W[(k + 2) % 3];
    const float* wa = W[k % 3];
    const float* wb = W[(k + 1) % 3];
    hipLaunchKernelGGL(k_bm_a, dim3(s->n_act), dim3(kBlock), 0, st, s->items, s->tiles, s->act, s->mode, S[k & 1],
                       k == 1, z, zprev, r1, F, x, wnew, wa, wb, s->pa);
    hipLaunchKernelGGL(k_bm_reduce, dim3(m), dim3(kBlock), 0, st, m, s->tile_ptr, s->pa, 0, 0, s->mode, nullptr,
                       S[k & 1], nullptr);
    pending_final = false;
    if (active.empty()) break;  // the last final updates have been applied
    hipLaunchKernelGGL(k_bm_b, dim3(s->n_act), dim3(kBmTile), 0, st, s->items, s->tiles, s->act, s->mode, S[k & 1],
                       F, r2, r1, zprev, s->pb);
    rc = bm_apply_pc(s, r1, zprev);
    if (rc) return rc;
    hipLaunchKernelGGL(k_bm_reduce, dim3(m), dim3(kBlock), 0, st, m, s->tile_ptr, s->pb, 1, 0, s->mode, S[k & 1],
                       S[(k + 1) & 1], nullptr);
    NFX_LAUNCH_CHECK();
    if (k % kBmCheck == 0 || k >= maxit) {
      // the state A(k + 1) will read (its phibar is the estimate after iteration k)
      NFX_CHECK_HIP(
          hipMemcpyAsync(s->h_st, S[(k + 1) & 1], sizeof(MinresState) * m, hipMemcpyDeviceToHost, st));
      NFX_CHECK_HIP(hipMemcpyAsync(s->h_mode, s->mode, sizeof(int32_t) * m, hipMemcpyDeviceToHost, st));
      NFX_CHECK_HIP(hipStreamSynchronize(st));
      std::vector<int> still;
      bool changed = false;
      for (int i : active) {
        const MinresState& q = s->h_st[i];
        const bool done = q.stop || !(q.beta1 > 0) || q.phibar <= tol[i] * q.beta1 || k >= maxit;
        if (done) {
          s->h_mode[i] = 1;
          iters[i] = k;
          changed = true;
        } else {
          still.push_back(i);
        }
      }
      active.swap(still);
      if (changed) {
        NFX_CHECK_HIP(hipMemcpyAsync(s->mode, s->h_mode, sizeof(int32_t) * m, hipMemcpyHostToDevice, st));
        pending_final = true;
        // tiles of instances still iterating or with their final update pending
        rc = bm_set_active(s, 2);
        if (rc) return rc;
      }
    }
  }
  // the host mode array is rewritten next: wait for every copy from it
  NFX_CHECK_HIP(hipStreamSynchronize(st));
  return NFX_OK;
}

extern "C" int nfx_solve_minres_batch_ir(nfx_minres_batch* s, const double* d_b, double* d_x, double rtol,
                                         double inner_rtol, int max_inner, int max_outer, int* h_iters,
                                         int* h_outer, double* h_resid) {
  NFX_REQUIRE(s && d_b && d_x && h_iters && h_outer && h_resid, "nfx_solve_minres_batch_ir: null pointer");
  NFX_REQUIRE(max_inner > 0 && max_outer >= 0, "nfx_solve_minres_batch_ir: iteration limits");
  const int m = s->n_inst;
  const int64_t N = s->n_total;
  hipStream_t st = s->stream;
  float* rf = s->buf + N * 8;
  float* dx = s->buf + N * 9;
  std::vector<int> outer_active(m, 1), it(m);
  std::vector<double> tol(m, inner_rtol);
  for (int i = 0; i < m; ++i) {
    h_iters[i] = 0;
    h_outer[i] = 0;
    h_resid[i] = 0;
  }
  for (int outer = 0; outer <= max_outer; ++outer) {
    int n_act = 0;
    for (int i = 0; i < m; ++i) {
      s->h_mode[i] = outer_active[i] ? 0 : 2;
      n_act += outer_active[i];
    }
    if (!n_act) break;
    NFX_CHECK_HIP(hipMemcpyAsync(s->mode, s->h_mode, sizeof(int32_t) * m, hipMemcpyHostToDevice, st));
    int rc = bm_set_active(s, 1);
    if (rc) return rc;
    hipLaunchKernelGGL(k_bm_res64, dim3(s->n_act), dim3(kBlock), 0, st, s->items, s->tiles, s->act, s->mode, d_x, d_b,
                       s->r64, s->pa, s->pb);
    hipLaunchKernelGGL(k_bm_reduce, dim3(m), dim3(kBlock), 0, st, m, s->tile_ptr, s->pa, 2, 0, s->mode, nullptr,
                       nullptr, s->sums);
    hipLaunchKernelGGL(k_bm_reduce, dim3(m), dim3(kBlock), 0, st, m, s->tile_ptr, s->pb, 2, 0, s->mode, nullptr,
                       nullptr, s->sums + m);
    NFX_LAUNCH_CHECK();
    NFX_CHECK_HIP(hipMemcpyAsync(s->h_sums, s->sums, sizeof(double) * 2 * m, hipMemcpyDeviceToHost, st));
    NFX_CHECK_HIP(hipStreamSynchronize(st));
    n_act = 0;
    for (int i = 0; i < m; ++i) {
      if (!outer_active[i]) continue;
      const double rn = sqrt(s->h_sums[i]), bn = sqrt(s->h_sums[m + i]);
      const double rel = bn > 0 ? rn / bn : rn;
      h_resid[i] = rel;
      h_outer[i] = outer;
      s->h_scale[i] = rn;
      if (!(rel == rel) || rel <= rtol || outer == max_outer || rn == 0) {
        outer_active[i] = 0;
        s->h_mode[i] = 2;
      } else {
        s->h_mode[i] = 0;
        ++n_act;
        // the correction needs only the reduction still missing (the recurrence's
        // estimate is in the preconditioner's norm: half of it as a margin), never less
        // than inner_rtol
        tol[i] = std::min(0.5, std::max(inner_rtol, 0.5 * rtol / rel));
      }
    }
    if (!n_act) break;
    NFX_CHECK_HIP(hipMemcpyAsync(s->scale, s->h_scale, sizeof(double) * m, hipMemcpyHostToDevice, st));
    NFX_CHECK_HIP(hipMemcpyAsync(s->mode, s->h_mode, sizeof(int32_t) * m, hipMemcpyHostToDevice, st));
    rc = bm_set_active(s, 1);
    if (rc) return rc;
    hipLaunchKernelGGL(k_bm_scale, dim3(s->n_act), dim3(kBmTile), 0, st, s->items, s->tiles, s->act, s->mode, s->scale,
                       0, s->r64, rf, nullptr, nullptr);
    NFX_LAUNCH_CHECK();
    // the inner solve changes h_mode: keep the outer set for the update that follows
    std::vector<int32_t> mode_outer(s->h_mode, s->h_mode + m);
    rc = bm_inner(s, rf, dx, tol.data(), max_inner, it.data());
    if (rc) return rc;
    for (int i = 0; i < m; ++i) {
      h_iters[i] += it[i];
      s->h_mode[i] = mode_outer[i];
    }
    NFX_CHECK_HIP(hipMemcpyAsync(s->mode, s->h_mode, sizeof(int32_t) * m, hipMemcpyHostToDevice, st));
    rc = bm_set_active(s, 1);
    if (rc) return rc;
    hipLaunchKernelGGL(k_bm_scale, dim3(s->n_act), dim3(kBmTile), 0, st, s->items, s->tiles, s->act, s->mode, s->scale,
                       1, nullptr, nullptr, dx, d_x);
    NFX_LAUNCH_CHECK();
  }
  NFX_CHECK_HIP(hipStreamSynchronize(st));
  for (int i = 0; i < m; ++i)
    if (!(h_resid[i] <= rtol)) return NFX_E_NOT_CONVERGED;
  return NFX_OK;
}
