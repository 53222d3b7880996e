// Flexible GMRES(m) in FP32 inside FP64 iterative refinement.
//
// Right preconditioning with a preconditioner that may vary between iterations
// (Saad 1993): the preconditioned directions Z are stored next to the Arnoldi
// basis V, x = x0 + Z y with y the least-squares solution of the Hessenberg
// system. Everything except the stopping test stays on the device:
//   * Gram-Schmidt: block partial dot products <v_i, w> (w staged in LDS), one
//     block per coefficient sums them in a fixed order into the Hessenberg column,
//     a third kernel subtracts the projections and forms block partial norms, and
//     a one-block kernel reduces those, writes h_{j+1,j}, applies the previous
//     Givens rotations, computes the new one and updates the residual estimate
//     |g_{j+1}| (FP64). Classical Gram-Schmidt, optionally repeated once (CGS2).
//   * basis vectors are stored unnormalised with their norms on the device; the
//     norms enter the Gram-Schmidt coefficients and the preconditioner output, so
//     no normalisation pass is needed.
//   * the host reads the residual estimate every `check` iterations.
// Preconditioners: the two-grid Schur block preconditioner (nfx_twogrid_apply),
// a diagonal, or none. The outer loop evaluates r = b - A x with the FP64 matrix.
#include <algorithm>
#include <vector>

#include "nfx_common.h"
#include "nfx_residual.h"

namespace nfx {
namespace {

// 512 threads per Gram-Schmidt block (8 waves over a 1,024-row chunk): the 32^3 x 6 learned
// solve 12.9 ms against 13.45 ms with 256 threads (a 2,048-row chunk: 13.8 ms)
constexpr int kFB = 512;
constexpr int kFgResLanes = 4;  // lanes per row of the FP64 refinement residual
constexpr int kChunkDot = 1024;  // rows per block of the dot-product kernel
constexpr int kChunkUpd = 1024;  // rows per block of the update/norm kernel

__device__ __forceinline__ double fg_warp_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// Deterministic block sum (fixed shuffle tree, then waves in order); valid in thread 0.
__device__ __forceinline__ double fg_block_sum(double v) {
  __shared__ double red[kFB / 64];
  v = fg_warp_sum(v);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double r = 0;
  if (threadIdx.x == 0)
    for (int i = 0; i < kFB / 64; ++i) r += red[i];
  __syncthreads();
  return r;
}

// partial[b][i] = <V_i, w> over block b's rows (i < j1), w staged in LDS.
__global__ __launch_bounds__(kFB) void k_fg_mdot(int n, int j1, int ld, const float* __restrict__ V,
                                                 const float* __restrict__ w, double* __restrict__ partial) {
  __shared__ float sw[kChunkDot];
  const int start = blockIdx.x * kChunkDot;
  const int len = min(kChunkDot, n - start);
  for (int k = threadIdx.x; k < len; k += kFB) sw[k] = w[start + k];
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int i = wave; i < j1; i += kFB / 64) {
    const float* v = V + (int64_t)i * n + start;
    float a = 0.f;
#pragma unroll 4
    for (int k = lane; k < len; k += 64) a += v[k] * sw[k];
    const double d = fg_warp_sum((double)a);
    if (lane == 0) partial[(int64_t)blockIdx.x * ld + i] = d;
  }
}

// Block i: s = sum_b partial[b][i] (fixed order) = <u_i, w> with u_i the stored
// (unnormalised) basis vector of norm hnv[i]: h_i = s / hnv[i] goes to H[i] (acc: +=),
// the update coefficient of u_i is h32[i] = s / hnv[i]^2.
__global__ __launch_bounds__(kFB) void k_fg_hreduce(int nb, int ld, const double* __restrict__ partial,
                                                    const double* __restrict__ hnv, double* __restrict__ H,
                                                    float* __restrict__ h32, int acc) {
  const int i = blockIdx.x;
  double s = 0;
  for (int b = threadIdx.x; b < nb; b += kFB) s += partial[(int64_t)b * ld + i];
  s = fg_block_sum(s);
  if (threadIdx.x == 0) {
    const double nv = hnv[i];
    const double h = nv > 0 ? s / nv : 0.0;
    h32[i] = (float)(nv > 0 ? h / nv : 0.0);
    H[i] = acc ? H[i] + h : h;
  }
}

// w -= sum_i h32[i] V_i (i < j1); with norm != 0 also block partial sums of w.w.
__global__ __launch_bounds__(kFB) void k_fg_mupdate(int n, int j1, const float* __restrict__ V,
                                                    const float* __restrict__ h32, float* __restrict__ w,
                                                    double* __restrict__ partial, int norm) {
  __shared__ float sh[512];
  for (int i = threadIdx.x; i < j1; i += kFB) sh[i] = h32[i];
  __syncthreads();
  const int start = blockIdx.x * kChunkUpd;
  const int end = min(n, start + kChunkUpd);
  double ss = 0;
  for (int p = start + threadIdx.x; p < end; p += kFB) {
    float v = w[p];
#pragma unroll 4
    for (int i = 0; i < j1; ++i) v -= sh[i] * V[(int64_t)i * n + p];
    if (j1 > 0) w[p] = v;
    ss += (double)v * v;
  }
  if (!norm) return;
  ss = fg_block_sum(ss);
  if (threadIdx.x == 0) partial[blockIdx.x] = ss;
}

// The first update and the second dot pass of the reorthogonalised Gram-Schmidt in one
// kernel: w -= sum_i h32[i] V_i over block b's rows (the arithmetic of k_fg_mupdate), then
// partial[b][i] = <V_i, w> with the updated rows (the arithmetic of k_fg_mdot), so that the
// block's rows of V come from HBM once and are re-read from cache for the dots.
__global__ __launch_bounds__(kFB) void k_fg_mupdate_dot(int n, int j1, int ld, const float* __restrict__ V,
                                                        const float* __restrict__ h32, float* __restrict__ w,
                                                        double* __restrict__ partial) {
  static_assert(kChunkDot == kChunkUpd, "one row range per block for both phases");
  __shared__ float sh[512];
  __shared__ float sw[kChunkDot];
  for (int i = threadIdx.x; i < j1; i += kFB) sh[i] = h32[i];
  __syncthreads();
  const int start = blockIdx.x * kChunkDot;
  const int len = min(kChunkDot, n - start);
  for (int k = threadIdx.x; k < len; k += kFB) {
    float v = w[start + k];
#pragma unroll 4
    for (int i = 0; i < j1; ++i) v -= sh[i] * V[(int64_t)i * n + start + k];
    w[start + k] = v;
    sw[k] = v;
  }
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int i = wave; i < j1; i += kFB / 64) {
    const float* v = V + (int64_t)i * n + start;
    float a = 0.f;
#pragma unroll 4
    for (int k = lane; k < len; k += 64) a += v[k] * sw[k];
    const double d = fg_warp_sum((double)a);
    if (lane == 0) partial[(int64_t)blockIdx.x * ld + i] = d;
  }
}

// hn = sqrt(sum_b partial[b]) (one block). mode 1: hnv[j+1] = hn, h_{j+1,j} = hn,
// Givens on column j, residual estimate |g_{j+1}|. mode 2 (start of a cycle):
// hnv[0] = g[0] = hn.
__global__ __launch_bounds__(kFB) void k_fg_finish(int nb, const double* __restrict__ partial, int mode, int j,
                                                   int m, double* __restrict__ H, double* __restrict__ cs,
                                                   double* __restrict__ sn, double* __restrict__ g,
                                                   double* __restrict__ res, double* __restrict__ hnv) {
  // the column and the previous rotations are staged in LDS by the whole block, so that
  // the serial rotation chain carries only its running entry in a register (a chain of
  // dependent global loads and stores cost ~0.2 us per rotation)
  __shared__ double s_col[512], s_cs[512], s_sn[512];
  double s = 0;
  for (int b = threadIdx.x; b < nb; b += kFB) s += partial[b];
  s = fg_block_sum(s);
  double* col = H + (int64_t)j * (m + 1);
  if (mode != 2) {
    for (int i = threadIdx.x; i <= j; i += kFB) s_col[i] = col[i];
    for (int i = threadIdx.x; i < j; i += kFB) {
      s_cs[i] = cs[i];
      s_sn[i] = sn[i];
    }
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  const double hn = sqrt(s);
  if (mode == 2) {
    hnv[0] = hn;
    g[0] = hn;
    *res = hn;
    return;
  }
  hnv[j + 1] = hn;
  double cur = s_col[0];  // entry i of the column after rotations 0 .. i-1
  for (int i = 0; i < j; ++i) {
    const double a = cur, b = s_col[i + 1];
    col[i] = s_cs[i] * a + s_sn[i] * b;
    cur = -s_sn[i] * a + s_cs[i] * b;
  }
  const double a = cur, b = hn;
  const double den = sqrt(a * a + b * b);
  const double c = den > 0 ? a / den : 1.0, sg = den > 0 ? b / den : 0.0;
  cs[j] = c;
  sn[j] = sg;
  col[j] = c * a + sg * b;
  col[j + 1] = 0.0;
  g[j + 1] = -sg * g[j];
  g[j] = c * g[j];
  *res = fabs(g[j + 1]);
}

// y = H(0:jj,0:jj)^-1 g(0:jj), back substitution by one wave (FP64)
__global__ void k_fg_trisolve(int jj, int m, const double* __restrict__ H, const double* __restrict__ g,
                              double* __restrict__ y) {
  __shared__ double sy[512];
  const int lane = threadIdx.x;
  for (int i = jj - 1; i >= 0; --i) {
    double s = 0;
    for (int k = i + 1 + lane; k < jj; k += 64) s += H[(int64_t)k * (m + 1) + i] * sy[k];
    s = fg_warp_sum(s);
    if (lane == 0) sy[i] = (g[i] - s) / H[(int64_t)i * (m + 1) + i];
    __syncthreads();
  }
  for (int i = lane; i < jj; i += 64) y[i] = sy[i];
}

// x += sum_i y[i] Z_i
__global__ void k_fg_xupdate(int n, int jj, const float* __restrict__ Z, const double* __restrict__ y,
                             float* __restrict__ x) {
  __shared__ float sy[512];
  for (int i = threadIdx.x; i < jj; i += blockDim.x) sy[i] = (float)y[i];
  __syncthreads();
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  float v = x[p];
  for (int i = 0; i < jj; ++i) v += sy[i] * Z[(int64_t)i * n + p];
  x[p] = v;
}

// z = diag(d) r / *dscale (d may be null: identity)
__global__ void k_fg_diag(int n, const float* __restrict__ d, const float* __restrict__ r,
                          const double* __restrict__ dscale, float* __restrict__ z) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const double ds = *dscale;
  const float sc = ds > 0 ? (float)(1.0 / ds) : 0.f;
  z[p] = sc * (d ? d[p] * r[p] : r[p]);
}

// FP64 residual r = b - A x with block partials of r.r and b.b
__global__ void k_fg_to_f32(int n, const double* __restrict__ r, double scale, float* __restrict__ rf) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) rf[p] = (float)(r[p] / scale);
}

__global__ void k_fg_axpy64(int n, double scale, const float* __restrict__ d, double* __restrict__ x) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) x[p] += scale * (double)d[p];
}

__global__ void k_fg_copy(int n, const float* __restrict__ a, float* __restrict__ b) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) b[p] = a[p];
}

__global__ void k_fg_zero(int n, float* __restrict__ a) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) a[p] = 0.f;
}

}  // namespace
}  // namespace nfx

using namespace nfx;

struct nfx_fgmres {
  int n = 0, m = 0, blocks = 0, ublocks = 0;
  hipStream_t stream = nullptr;
  float *V = nullptr, *Z = nullptr, *rhs = nullptr, *x32 = nullptr, *h32 = nullptr;
  double *H = nullptr, *cs = nullptr, *sn = nullptr, *g = nullptr, *y = nullptr;
  double *partial = nullptr, *upartial = nullptr, *hnv = nullptr, *res = nullptr, *r64 = nullptr, *h_res = nullptr;
  double *partial_r = nullptr, *h_partial = nullptr;
  int rblocks = 0;
};

extern "C" int nfx_fgmres_destroy(nfx_fgmres* s) {
  if (!s) return NFX_OK;
  void* dev[] = {s->V, s->Z, s->rhs, s->x32, s->h32, s->H, s->cs, s->sn, s->g, s->y,
                 s->partial, s->upartial, s->hnv, s->res, s->r64, s->partial_r};
  for (void* p : dev)
    if (p) (void)hipFree(p);
  if (s->h_res) (void)hipHostFree(s->h_res);
  if (s->h_partial) (void)hipHostFree(s->h_partial);
  delete s;
  return NFX_OK;
}

extern "C" int nfx_fgmres_create(nfx_fgmres** out, int n_rows, int restart, void* stream) {
  NFX_REQUIRE(out && n_rows > 0, "nfx_fgmres_create: bad arguments");
  NFX_REQUIRE(restart >= 1 && restart <= 500, "nfx_fgmres_create: restart must be in [1, 500]");
  nfx_fgmres* s = new nfx_fgmres();
  s->n = n_rows;
  s->m = restart;
  s->stream = static_cast<hipStream_t>(stream);
  s->blocks = grid_for(n_rows, kChunkDot);
  s->ublocks = grid_for(n_rows, kChunkUpd);
  // residual lanes per row: the Darcy saddle points FGMRES serves have rows of at most 12
  // entries (RT1-P0), 4 lanes x 3 entries cover one in a single pass
  s->rblocks = residual_blocks(n_rows, kFgResLanes);
  const size_t n = n_rows, m = restart;
  hipError_t e = hipMalloc(&s->V, sizeof(float) * n * (m + 1));
  if (e == hipSuccess) e = hipMalloc(&s->Z, sizeof(float) * n * m);
  if (e == hipSuccess) e = hipMalloc(&s->rhs, sizeof(float) * n);
  if (e == hipSuccess) e = hipMalloc(&s->x32, sizeof(float) * n);
  if (e == hipSuccess) e = hipMalloc(&s->h32, sizeof(float) * (m + 1));
  if (e == hipSuccess) e = hipMalloc(&s->H, sizeof(double) * (m + 1) * m);
  if (e == hipSuccess) e = hipMalloc(&s->cs, sizeof(double) * m);
  if (e == hipSuccess) e = hipMalloc(&s->sn, sizeof(double) * m);
  if (e == hipSuccess) e = hipMalloc(&s->g, sizeof(double) * (m + 1));
  if (e == hipSuccess) e = hipMalloc(&s->y, sizeof(double) * m);
  if (e == hipSuccess) e = hipMalloc(&s->partial, sizeof(double) * (size_t)s->blocks * (m + 1));
  if (e == hipSuccess) e = hipMalloc(&s->upartial, sizeof(double) * (size_t)s->ublocks);
  if (e == hipSuccess) e = hipMalloc(&s->hnv, sizeof(double) * (m + 1));
  if (e == hipSuccess) e = hipMalloc(&s->res, sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&s->r64, sizeof(double) * n);
  if (e == hipSuccess) e = hipMalloc(&s->partial_r, sizeof(double) * 2 * s->rblocks);
  if (e == hipSuccess) e = hipHostMalloc(&s->h_res, sizeof(double), hipHostMallocDefault);
  if (e == hipSuccess) e = hipHostMalloc(&s->h_partial, sizeof(double) * 2 * s->rblocks, hipHostMallocDefault);
  if (e != hipSuccess) {
    set_error(std::string("nfx_fgmres_create: ") + hipGetErrorString(e));
    nfx_fgmres_destroy(s);
    return NFX_E_HIP;
  }
  *out = s;
  return NFX_OK;
}

extern "C" int nfx_spmv(int n_rows, const int32_t* d_rowptr, const int32_t* d_col, const float* d_val,
                        float alpha, const float* d_x, float beta, float* d_y, void* stream);

// z = M^-1 (u / *hn) for the stored basis vector u of norm *hn
static int fg_precond(nfx_fgmres* s, const nfx_twogrid_pc* pc, const float* pinv, const float* u, const double* hn,
                      float* z) {
  if (pc) return twogrid_apply_scaled(pc, u, z, hn, s->stream);
  hipLaunchKernelGGL(k_fg_diag, dim3(grid_for(s->n, 256)), dim3(256), 0, s->stream, s->n, pinv, u, hn, z);
  NFX_LAUNCH_CHECK();
  return NFX_OK;
}

// FP32 FGMRES on A x = rhs (x = 0 initially), ||rhs|| = 1 expected (scaled by the caller).
static int fg_inner(nfx_fgmres* s, const int32_t* rp, const int32_t* ci, const float* va,
                    const nfx_twogrid_pc* pc, const float* pinv, double tol, int max_it, int reorth,
                    int check, int* iters, double* est) {
  const int n = s->n, m = s->m, B = s->blocks, BU = s->ublocks, ld = m + 1;
  const int gn = grid_for(n, 256);
  hipStream_t st = s->stream;
  hipLaunchKernelGGL(k_fg_zero, dim3(gn), dim3(256), 0, st, n, s->x32);
  int total = 0;
  double beta0 = -1, resid = 0;
  bool done = false;
  while (!done) {
    // u_0 = rhs - A x (stored unnormalised in V_0); hnv[0] = g[0] = ||u_0||
    hipLaunchKernelGGL(k_fg_copy, dim3(gn), dim3(256), 0, st, n, s->rhs, s->V);
    if (total > 0) {
      int rc = nfx_spmv(n, rp, ci, va, -1.f, s->x32, 1.f, s->V, st);
      if (rc) return rc;
    }
    hipLaunchKernelGGL(k_fg_mupdate, dim3(BU), dim3(kFB), 0, st, n, 0, s->V, s->h32, s->V, s->upartial, 1);
    hipLaunchKernelGGL(k_fg_finish, dim3(1), dim3(kFB), 0, st, BU, s->upartial, 2, 0, m, s->H, s->cs, s->sn, s->g,
                       s->res, s->hnv);
    NFX_LAUNCH_CHECK();
    if (beta0 < 0) {
      NFX_CHECK_HIP(hipMemcpyAsync(s->h_res, s->res, sizeof(double), hipMemcpyDeviceToHost, st));
      NFX_CHECK_HIP(hipStreamSynchronize(st));
      beta0 = *s->h_res;
      if (!(beta0 > 0)) {
        resid = 0;
        break;
      }
    }
    int jj = 0;
    for (int j = 0; j < m && total < max_it; ++j) {
      const float* uj = s->V + (size_t)j * n;
      float* w = s->V + (size_t)(j + 1) * n;
      float* zj = s->Z + (size_t)j * n;
      int rc = fg_precond(s, pc, pinv, uj, s->hnv + j, zj);
      if (rc) return rc;
      rc = nfx_spmv(n, rp, ci, va, 1.f, zj, 0.f, w, st);
      if (rc) return rc;
      double* Hcol = s->H + (size_t)j * ld;
      const int passes = reorth ? 2 : 1;
      // pass p: dots (the first pass's own kernel, later ones fused into the previous
      // update), coefficients, update (fused with the next pass's dots, the last with the norm)
      hipLaunchKernelGGL(k_fg_mdot, dim3(B), dim3(kFB), 0, st, n, j + 1, ld, s->V, w, s->partial);
      for (int pass = 0; pass < passes; ++pass) {
        hipLaunchKernelGGL(k_fg_hreduce, dim3(j + 1), dim3(kFB), 0, st, B, ld, s->partial, s->hnv, Hcol, s->h32,
                           pass);
        if (pass < passes - 1)
          hipLaunchKernelGGL(k_fg_mupdate_dot, dim3(B), dim3(kFB), 0, st, n, j + 1, ld, s->V, s->h32, w,
                             s->partial);
        else
          hipLaunchKernelGGL(k_fg_mupdate, dim3(BU), dim3(kFB), 0, st, n, j + 1, s->V, s->h32, w, s->upartial, 1);
      }
      hipLaunchKernelGGL(k_fg_finish, dim3(1), dim3(kFB), 0, st, BU, s->upartial, 1, j, m, s->H, s->cs, s->sn,
                         s->g, s->res, s->hnv);
      NFX_LAUNCH_CHECK();
      ++total;
      jj = j + 1;
      if (jj % check == 0 || jj == m || total >= max_it) {
        NFX_CHECK_HIP(hipMemcpyAsync(s->h_res, s->res, sizeof(double), hipMemcpyDeviceToHost, st));
        NFX_CHECK_HIP(hipStreamSynchronize(st));
        resid = *s->h_res;
        if (!(resid == resid)) {
          set_error("fgmres: non-finite residual estimate");
          return NFX_E_NOT_CONVERGED;
        }
        if (resid <= tol * beta0) {
          done = true;
          break;
        }
      }
    }
    if (jj > 0) {
      hipLaunchKernelGGL(k_fg_trisolve, dim3(1), dim3(64), 0, st, jj, m, s->H, s->g, s->y);
      hipLaunchKernelGGL(k_fg_xupdate, dim3(gn), dim3(256), 0, st, n, jj, s->Z, s->y, s->x32);
      NFX_LAUNCH_CHECK();
    }
    if (total >= max_it) break;
  }
  *iters = total;
  *est = beta0 > 0 ? resid / beta0 : 0.0;
  return NFX_OK;
}

extern "C" int nfx_solve_fgmres_ir(nfx_fgmres* s, const int32_t* d_rowptr, const int32_t* d_col,
                                   const float* d_val32, const double* d_val64, const nfx_twogrid_pc* pc,
                                   const float* d_pinv, const double* d_b, double* d_x, double rtol,
                                   double inner_rtol, int max_inner, int max_outer, int reorth, int check_every,
                                   int* h_info, double* h_resid, double* h_hist, int hist_cap) {
  NFX_REQUIRE(s && d_rowptr && d_col && d_val32 && d_val64 && d_b && d_x, "nfx_solve_fgmres_ir: null pointer");
  NFX_REQUIRE(max_inner > 0 && max_outer >= 0 && check_every > 0, "nfx_solve_fgmres_ir: iteration limits");
  if (pc) {
    NFX_REQUIRE(pc->n_u + pc->n_cells + pc->n_tail == s->n,
                "nfx_solve_fgmres_ir: preconditioner size differs from the matrix");
    NFX_REQUIRE(pc->work && pc->Ainv && pc->nbr, "nfx_solve_fgmres_ir: incomplete two-grid preconditioner");
  }
  const int n = s->n;
  const int gn = grid_for(n, 256);
  hipStream_t st = s->stream;
  int total = 0, outer = 0;
  double rel = 0;
  for (; outer <= max_outer; ++outer) {
    NFX_CHECK_HIP(launch_residual64(kFgResLanes, s->rblocks, st, n, d_rowptr, d_col, d_val64, d_x, d_b, s->r64,
                                    s->partial_r));
    NFX_CHECK_HIP(hipMemcpyAsync(s->h_partial, s->partial_r, sizeof(double) * 2 * s->rblocks,
                                 hipMemcpyDeviceToHost, st));
    NFX_CHECK_HIP(hipStreamSynchronize(st));
    double rr = 0, bb = 0;
    for (int i = 0; i < s->rblocks; ++i) {
      rr += s->h_partial[2 * i];
      bb += s->h_partial[2 * i + 1];
    }
    const double rn = sqrt(rr), bn = sqrt(bb);
    rel = bn > 0 ? rn / bn : rn;
    if (h_hist && outer < hist_cap) h_hist[outer] = rel;
    if (!(rel == rel)) break;
    if (rel <= rtol || outer == max_outer || rn == 0) break;
    hipLaunchKernelGGL(k_fg_to_f32, dim3(gn), dim3(256), 0, st, n, s->r64, rn, s->rhs);
    NFX_LAUNCH_CHECK();
    int it = 0;
    double est = 0;
    int rc = fg_inner(s, d_rowptr, d_col, d_val32, pc, d_pinv, inner_rtol, max_inner, reorth, check_every, &it,
                      &est);
    total += it;
    if (rc) {
      if (h_info) {
        h_info[0] = outer;
        h_info[1] = total;
      }
      if (h_resid) h_resid[0] = rel;
      return rc;
    }
    hipLaunchKernelGGL(k_fg_axpy64, dim3(gn), dim3(256), 0, st, n, rn, s->x32, d_x);
    NFX_LAUNCH_CHECK();
  }
  NFX_CHECK_HIP(hipStreamSynchronize(st));
  if (h_info) {
    h_info[0] = outer;
    h_info[1] = total;
  }
  if (h_resid) h_resid[0] = rel;
  return rel <= rtol ? NFX_OK : NFX_E_NOT_CONVERGED;
}
