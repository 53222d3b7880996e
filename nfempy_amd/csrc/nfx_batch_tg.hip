// Two-grid Schur preconditioner of the pressure rows of many independent Darcy systems
// (RT1-P0: one pressure unknown per cell, RT2-P1: four), applied inside the batched
// MINRES: one set of launches per application for every instance of a group.
//
// Per cell c (cells of one instance contiguous, instance after instance):
//   S = B diag(M)^-1 B^T in blocks: S_cc (nb x nb) and one block per local face k with the
//   neighbour across it (the flux DoFs of the face are the only ones two cells share);
//   smoother W: omega S_cc^-1 per cell (block Jacobi) or, for nb = 1, the learned smoother
//   entries (diagonal + one coupling per face);
//   aggregates of cells, never spanning two instances, with the piecewise-constant
//   prolongation (every unknown of every cell of an aggregate gets the aggregate's value);
//   coarse operator P^T S P per instance, dense, inverted per instance (block diagonal over
//   the batch, so no value of one instance reaches another).
// The cycle z = x2 + W (r - S x2), x2 = x1 + P Ac^-1 P^T (r - S x1), x1 = W r is
// symmetric; with the smoother a contraction in the S-norm it is SPD, as MINRES needs.
#include "nfx_blockinv.h"
#include "nfx_common.h"

namespace nfx {
namespace {

constexpr int kTB = 256;

// S blocks of cell c from the FP64 CSR values of its instance (vals[inst]):
//   S_cc[i][j] = sum_a B(c, a, i) B(c, a, j) / M(f_a)          over the cell's flux DoFs a
//   S_ck[i][j] = sum_m B(c, a_m, i) B(d, a'_m, j) / M(f_a_m)   over the DoFs of face k
// with B(c, a, i) the value at posb[c][a][i] (pressure row i of c, flux column f_a) and
// M(f) the flux diagonal at posd[c][a].
template <int NB>
__global__ __launch_bounds__(kTB) void k_btg_schur(int n_cells, int nbf, int fd, const int32_t* __restrict__ cell_inst,
                                                   const double* const* __restrict__ vals,
                                                   const int32_t* __restrict__ posb, const int32_t* __restrict__ posd,
                                                   const int32_t* __restrict__ nbr, const int32_t* __restrict__ fdofs,
                                                   const int32_t* __restrict__ nloc, double* __restrict__ sd64,
                                                   double* __restrict__ so64, float* __restrict__ sd,
                                                   float* __restrict__ so) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n_cells) return;
  const double* v = vals[cell_inst[c]];
  double S[NB * NB];
#pragma unroll
  for (int e = 0; e < NB * NB; ++e) S[e] = 0.0;
  for (int a = 0; a < nbf; ++a) {
    const double dinv = 1.0 / v[posd[(int64_t)c * nbf + a]];
    double b[NB];
#pragma unroll
    for (int i = 0; i < NB; ++i) b[i] = v[posb[((int64_t)c * nbf + a) * NB + i]];
#pragma unroll
    for (int i = 0; i < NB; ++i)
#pragma unroll
      for (int j = 0; j < NB; ++j) S[i * NB + j] += b[i] * b[j] * dinv;
  }
#pragma unroll
  for (int e = 0; e < NB * NB; ++e) {
    sd64[(int64_t)c * NB * NB + e] = S[e];
    sd[(int64_t)c * NB * NB + e] = (float)S[e];
  }
  for (int k = 0; k < 4; ++k) {
    const int d = nbr[4 * c + k];
#pragma unroll
    for (int e = 0; e < NB * NB; ++e) S[e] = 0.0;
    if (d >= 0) {
      for (int m = 0; m < fd; ++m) {
        const int a = fdofs[k * fd + m];
        const int a2 = nloc[((int64_t)c * 4 + k) * fd + m];
        const double dinv = 1.0 / v[posd[(int64_t)c * nbf + a]];
        double bi[NB], bj[NB];
#pragma unroll
        for (int i = 0; i < NB; ++i) {
          bi[i] = v[posb[((int64_t)c * nbf + a) * NB + i]];
          bj[i] = v[posb[((int64_t)d * nbf + a2) * NB + i]];
        }
#pragma unroll
        for (int i = 0; i < NB; ++i)
#pragma unroll
          for (int j = 0; j < NB; ++j) S[i * NB + j] += bi[i] * bj[j] * dinv;
      }
    }
#pragma unroll
    for (int e = 0; e < NB * NB; ++e) {
      so64[((int64_t)c * 4 + k) * NB * NB + e] = S[e];
      so[((int64_t)c * 4 + k) * NB * NB + e] = (float)S[e];
    }
  }
}

// The same blocks with the flux mass in 3 x 3 blocks instead of its diagonal (RT2-P1: the
// three DoFs of a face, the three interior DoFs of a cell): D_g^-1 from the CSR positions
// posmf (n_cells, 4, 3, 3) of each local face's block and posmi (n_cells, 3, 3) of the
// interior block, in the cell's local DoF order (local DoFs 12..14 are the interior ones).
//   S_cc = sum_g B_g D_g^-1 B_g^T over the cell's 5 groups, S_ck = B_g(c) D_g^-1 B_g(d)^T.
__global__ __launch_bounds__(kTB) void k_btg_schur_blk(int n_cells, const int32_t* __restrict__ cell_inst,
                                                       const double* const* __restrict__ vals,
                                                       const int32_t* __restrict__ posb,
                                                       const int32_t* __restrict__ posmf,
                                                       const int32_t* __restrict__ posmi,
                                                       const int32_t* __restrict__ nbr,
                                                       const int32_t* __restrict__ fdofs,
                                                       const int32_t* __restrict__ nloc, double* __restrict__ sd64,
                                                       double* __restrict__ so64, float* __restrict__ sd,
                                                       float* __restrict__ so, int32_t* __restrict__ status) {
  constexpr int NB = 4, G = 3, NBF = 15;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n_cells) return;
  const double* v = vals[cell_inst[c]];
  auto bval = [&](int cell, int a, int i) { return v[posb[((int64_t)cell * NBF + a) * NB + i]]; };
  double S[NB * NB];
#pragma unroll
  for (int e = 0; e < NB * NB; ++e) S[e] = 0.0;
  // groups 0..3: faces, 4: interior; the face blocks are kept for the neighbour terms
  double Dinv[5][G][G];
  bool ok = true;
  for (int g = 0; g < 5; ++g) {
    const int32_t* pm = g < 4 ? posmf + ((int64_t)c * 4 + g) * G * G : posmi + (int64_t)c * G * G;
    double A[G][G];
#pragma unroll
    for (int i = 0; i < G; ++i)
#pragma unroll
      for (int j = 0; j < G; ++j) A[i][j] = v[pm[i * G + j]];
    ok = spd_block_inverse<G>(A) && ok;
#pragma unroll
    for (int i = 0; i < G; ++i)
#pragma unroll
      for (int j = 0; j < G; ++j) Dinv[g][i][j] = A[i][j];
    int la[G];
#pragma unroll
    for (int m = 0; m < G; ++m) la[m] = g < 4 ? fdofs[g * G + m] : 4 * G + m;
    double b[G][NB], t[G][NB];
#pragma unroll
    for (int m = 0; m < G; ++m)
#pragma unroll
      for (int i = 0; i < NB; ++i) b[m][i] = bval(c, la[m], i);
#pragma unroll
    for (int m = 0; m < G; ++m)
#pragma unroll
      for (int j = 0; j < NB; ++j) {
        double acc = 0.0;
#pragma unroll
        for (int q = 0; q < G; ++q) acc += A[m][q] * b[q][j];
        t[m][j] = acc;
      }
#pragma unroll
    for (int i = 0; i < NB; ++i)
#pragma unroll
      for (int j = 0; j < NB; ++j) {
        double acc = 0.0;
#pragma unroll
        for (int m = 0; m < G; ++m) acc += b[m][i] * t[m][j];
        S[i * NB + j] += acc;
      }
  }
  if (!ok) atomicOr(status, 2);
#pragma unroll
  for (int e = 0; e < NB * NB; ++e) {
    sd64[(int64_t)c * NB * NB + e] = S[e];
    sd[(int64_t)c * NB * NB + e] = (float)S[e];
  }
  for (int k = 0; k < 4; ++k) {
    const int d = nbr[4 * c + k];
#pragma unroll
    for (int e = 0; e < NB * NB; ++e) S[e] = 0.0;
    if (d >= 0) {
      double bc[G][NB], bd[G][NB];
#pragma unroll
      for (int m = 0; m < G; ++m) {
        const int a = fdofs[k * G + m];
        const int a2 = nloc[((int64_t)c * 4 + k) * G + m];
#pragma unroll
        for (int i = 0; i < NB; ++i) {
          bc[m][i] = bval(c, a, i);
          bd[m][i] = bval(d, a2, i);
        }
      }
#pragma unroll
      for (int i = 0; i < NB; ++i)
#pragma unroll
        for (int j = 0; j < NB; ++j) {
          double acc = 0.0;
#pragma unroll
          for (int m = 0; m < G; ++m)
#pragma unroll
            for (int q = 0; q < G; ++q) acc += bc[m][i] * Dinv[k][m][q] * bd[q][j];
          S[i * NB + j] = acc;
        }
    }
#pragma unroll
    for (int e = 0; e < NB * NB; ++e) {
      so64[((int64_t)c * 4 + k) * NB * NB + e] = S[e];
      so[((int64_t)c * 4 + k) * NB * NB + e] = (float)S[e];
    }
  }
}

// Flux-block preconditioner: inverse of each 3 x 3 flux block (FP64, FP32 out) from the CSR
// positions fb_pos (n_fb, 3, 3) of its instance's values.
__global__ __launch_bounds__(kTB) void k_btg_fluxinv(int n_fb, const int32_t* __restrict__ fb_inst,
                                                     const double* const* __restrict__ vals,
                                                     const int32_t* __restrict__ fb_pos, float* __restrict__ fb_inv,
                                                     int32_t* __restrict__ status) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n_fb) return;
  const double* v = vals[fb_inst[g]];
  double A[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) A[i][j] = v[fb_pos[(int64_t)g * 9 + i * 3 + j]];
  if (!spd_block_inverse<3>(A)) atomicOr(status, 4);
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) fb_inv[(int64_t)g * 9 + i * 3 + j] = (float)A[i][j];
}

// rs[c][b][i] = sum_j S_b[i][j] over the cell's own block (b = 0) and neighbour blocks
template <int NB>
__global__ __launch_bounds__(kTB) void k_btg_rowsums(int n_cells, const double* __restrict__ sd64,
                                                     const double* __restrict__ so64, float* __restrict__ rs) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n_cells) return;
  for (int b = 0; b < 5; ++b) {
    const double* S = b == 0 ? sd64 + (int64_t)c * NB * NB : so64 + ((int64_t)c * 4 + (b - 1)) * NB * NB;
#pragma unroll
    for (int i = 0; i < NB; ++i) {
      double v = 0.0;
#pragma unroll
      for (int j = 0; j < NB; ++j) v += S[i * NB + j];
      rs[((int64_t)c * 5 + b) * NB + i] = (float)v;
    }
  }
}

// W_c = omega S_cc^-1 (FP64 inverse, FP32 out); bit 0 of *status when a block is not SPD
template <int NB>
__global__ __launch_bounds__(kTB) void k_btg_smoother(int n_cells, const double* __restrict__ sd64, double omega,
                                                      float* __restrict__ wd, int32_t* __restrict__ status) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n_cells) return;
  double A[NB][NB];
#pragma unroll
  for (int i = 0; i < NB; ++i)
#pragma unroll
    for (int j = 0; j < NB; ++j) A[i][j] = sd64[(int64_t)c * NB * NB + i * NB + j];
  if (!spd_block_inverse<NB>(A)) {
    atomicOr(status, 1);
#pragma unroll
    for (int i = 0; i < NB; ++i)
#pragma unroll
      for (int j = 0; j < NB; ++j) A[i][j] = 0.0;
  }
#pragma unroll
  for (int i = 0; i < NB; ++i)
#pragma unroll
    for (int j = 0; j < NB; ++j) wd[(int64_t)c * NB * NB + i * NB + j] = (float)(omega * A[i][j]);
}

// Row a of the packed coarse operator of a's instance: sum over the cells p of a of the
// entry sums of S_pp (column a) and of S_pk (column of the neighbour's aggregate).
// One thread per aggregate, members in aggregate order: a fixed summation order.
template <int NB>
__global__ __launch_bounds__(64) void k_btg_coarse(int n_agg, const int32_t* __restrict__ agg_ptr,
                                                   const int32_t* __restrict__ agg_mem,
                                                   const int32_t* __restrict__ agg_of, const int32_t* __restrict__ nbr,
                                                   const double* __restrict__ sd64, const double* __restrict__ so64,
                                                   const int64_t* __restrict__ arow, const int32_t* __restrict__ a0,
                                                   double* __restrict__ ac) {
  const int ag = blockIdx.x * blockDim.x + threadIdx.x;
  if (ag >= n_agg) return;
  double* row = ac + arow[ag];
  const int base = a0[ag];
  for (int k = agg_ptr[ag]; k < agg_ptr[ag + 1]; ++k) {
    const int p = agg_mem[k];
    double s = 0.0;
#pragma unroll
    for (int e = 0; e < NB * NB; ++e) s += sd64[(int64_t)p * NB * NB + e];
    row[ag - base] += s;
    for (int j = 0; j < 4; ++j) {
      const int q = nbr[4 * p + j];
      if (q < 0) continue;
      double o = 0.0;
#pragma unroll
      for (int e = 0; e < NB * NB; ++e) o += so64[((int64_t)p * 4 + j) * NB * NB + e];
      row[agg_of[q] - base] += o;
    }
  }
}

template <int NB>
__device__ __forceinline__ void load_r(int c, const int32_t* __restrict__ prow, const float* __restrict__ r,
                                       float (&v)[NB]) {
#pragma unroll
  for (int i = 0; i < NB; ++i) v[i] = r[prow[(int64_t)c * NB + i]];
}

// out = S_cc x_c + sum_k S_ck x_nbr(k), x read through a functor x(q, i)
template <int NB, typename F>
__device__ __forceinline__ void s_row(int c, const int32_t* __restrict__ nbr, const float* __restrict__ sd,
                                      const float* __restrict__ so, F x, float (&out)[NB]) {
  float xc[NB];
#pragma unroll
  for (int j = 0; j < NB; ++j) xc[j] = x(c, j);
#pragma unroll
  for (int i = 0; i < NB; ++i) {
    float s = 0.f;
#pragma unroll
    for (int j = 0; j < NB; ++j) s += sd[(int64_t)c * NB * NB + i * NB + j] * xc[j];
    out[i] = s;
  }
  for (int k = 0; k < 4; ++k) {
    const int q = nbr[4 * c + k];
    if (q < 0) continue;
    float xq[NB];
#pragma unroll
    for (int j = 0; j < NB; ++j) xq[j] = x(q, j);
#pragma unroll
    for (int i = 0; i < NB; ++i) {
      float s = 0.f;
#pragma unroll
      for (int j = 0; j < NB; ++j) s += so[((int64_t)c * 4 + k) * NB * NB + i * NB + j] * xq[j];
      out[i] += s;
    }
  }
}

// out = W_cc t_c (+ sum_k W_ck t_nbr(k) when wo is given, nb = 1)
template <int NB, typename F>
__device__ __forceinline__ void w_row(int c, const int32_t* __restrict__ nbr, const float* __restrict__ wd,
                                      const float* __restrict__ wo, F t, float (&out)[NB]) {
  float tc[NB];
#pragma unroll
  for (int j = 0; j < NB; ++j) tc[j] = t(c, j);
#pragma unroll
  for (int i = 0; i < NB; ++i) {
    float s = 0.f;
#pragma unroll
    for (int j = 0; j < NB; ++j) s += wd[(int64_t)c * NB * NB + i * NB + j] * tc[j];
    out[i] = s;
  }
  if (NB == 1 && wo) {
    for (int k = 0; k < 4; ++k) {
      const int q = nbr[4 * c + k];
      if (q >= 0) out[0] += wo[4 * c + k] * t(q, 0);
    }
  }
}

// Work of one group: x1 | x2 | t2 | t1 (n_cells nb each) | cv | y (n_agg each).
struct GroupWork {
  float *x1, *x2, *t2, *t1, *cv, *y;
};

__device__ __forceinline__ GroupWork group_work(const nfx_batch_tg& pc) {
  GroupWork w;
  const int64_t cn = (int64_t)pc.n_cells * pc.nb;
  w.x1 = pc.work;
  w.x2 = w.x1 + cn;
  w.t2 = w.x2 + cn;
  w.t1 = w.t2 + cn;
  w.cv = w.t1 + cn;
  w.y = w.cv + pc.n_agg;
  return w;
}

// Stage bodies, one work item (cell, aggregate wave or flux block) each; the kernels below
// run two groups (RT1-P0 and RT2-P1) in one launch: blocks [0, split) serve the first,
// the rest the second. Items of an instance whose mode is not 0 (converged or not
// selected) return at once: every stage reads only its own instance's cells, aggregates
// and coarse block, and each application writes its work arrays before reading them.

__device__ __forceinline__ bool cell_idle(const nfx_batch_tg& pc, const int32_t* __restrict__ mode, int c) {
  return mode && mode[pc.cell_inst[c]] != 0;
}

__device__ __forceinline__ bool agg_idle(const nfx_batch_tg& pc, const int32_t* __restrict__ mode, int a) {
  return mode && mode[pc.cell_inst[pc.agg_mem[pc.agg_ptr[a]]]] != 0;
}

// x1 = W r
template <int NB>
__device__ __forceinline__ void st_pre(const nfx_batch_tg& pc, int c, const int32_t* __restrict__ mode,
                                       const float* __restrict__ r) {
  if (c >= pc.n_cells || cell_idle(pc, mode, c)) return;
  const GroupWork w = group_work(pc);
  const int32_t* prow = pc.prow;
  float o[NB];
  w_row<NB>(c, pc.nbr, pc.wd, pc.wo, [&](int q, int i) { return r[prow[(int64_t)q * NB + i]]; }, o);
#pragma unroll
  for (int i = 0; i < NB; ++i) w.x1[(int64_t)c * NB + i] = o[i];
}

// t1 = r - S x1 on the cells of aggregate `wave` (kept for the post step) and
// cv[wave] = the sum of t1 over them (one wave per aggregate, lanes over its cells)
template <int NB>
__device__ __forceinline__ void st_restrict(const nfx_batch_tg& pc, int wave, int lane,
                                            const int32_t* __restrict__ mode, const float* __restrict__ r) {
  if (wave >= pc.n_agg || agg_idle(pc, mode, wave)) return;
  const GroupWork w = group_work(pc);
  float acc = 0.f;
  for (int k = pc.agg_ptr[wave] + lane; k < pc.agg_ptr[wave + 1]; k += 64) {
    const int p = pc.agg_mem[k];
    float rp[NB], s[NB];
    load_r<NB>(p, pc.prow, r, rp);
    s_row<NB>(p, pc.nbr, pc.sd, pc.so, [&](int q, int i) { return w.x1[(int64_t)q * NB + i]; }, s);
#pragma unroll
    for (int i = 0; i < NB; ++i) {
      const float t = rp[i] - s[i];
      w.t1[(int64_t)p * NB + i] = t;
      acc += t;
    }
  }
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if (lane == 0) w.cv[wave] = acc;
}

// y[a] = sum_t ainv[arow[a] + t] cv[a0[a] + t] over the aggregates of a's instance
__device__ __forceinline__ void st_gemv(const nfx_batch_tg& pc, int wave, int lane, const int32_t* __restrict__ mode) {
  if (wave >= pc.n_agg || agg_idle(pc, mode, wave)) return;
  const GroupWork w = group_work(pc);
  const float* a = pc.ainv + pc.arow[wave];
  const float* x = w.cv + pc.a0[wave];
  const int n = pc.alen[wave];
  float acc = 0.f;
  for (int t = lane; t < n; t += 64) acc += a[t] * x[t];
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if (lane == 0) w.y[wave] = acc;
}

// x2 = x1 + P y; t2 = r - S x2 = t1 - S P y, with S P y from the row sums of the blocks
// (rs[c][0][i] for the cell's own block, rs[c][1 + k][i] for neighbour k)
template <int NB>
__device__ __forceinline__ void st_post1(const nfx_batch_tg& pc, int c, const int32_t* __restrict__ mode) {
  if (c >= pc.n_cells || cell_idle(pc, mode, c)) return;
  const GroupWork w = group_work(pc);
  const float yc = w.y[pc.agg_of[c]];
  float spy[NB];
#pragma unroll
  for (int i = 0; i < NB; ++i) spy[i] = pc.rs[((int64_t)c * 5) * NB + i] * yc;
  for (int k = 0; k < 4; ++k) {
    const int q = pc.nbr[4 * c + k];
    if (q < 0) continue;
    const float yq = w.y[pc.agg_of[q]];
#pragma unroll
    for (int i = 0; i < NB; ++i) spy[i] += pc.rs[((int64_t)c * 5 + 1 + k) * NB + i] * yq;
  }
#pragma unroll
  for (int i = 0; i < NB; ++i) {
    w.x2[(int64_t)c * NB + i] = w.x1[(int64_t)c * NB + i] + yc;
    w.t2[(int64_t)c * NB + i] = w.t1[(int64_t)c * NB + i] - spy[i];
  }
}

// z = x2 + W t2 on the rows of cells whose instance is active (mode 0, or all without mode)
template <int NB>
__device__ __forceinline__ void st_post2(const nfx_batch_tg& pc, int c, const int32_t* __restrict__ mode,
                                         float* __restrict__ z) {
  if (c >= pc.n_cells || cell_idle(pc, mode, c)) return;
  const GroupWork w = group_work(pc);
  float o[NB];
  w_row<NB>(c, pc.nbr, pc.wd, pc.wo, [&](int q, int i) { return w.t2[(int64_t)q * NB + i]; }, o);
#pragma unroll
  for (int i = 0; i < NB; ++i) z[pc.prow[(int64_t)c * NB + i]] = w.x2[(int64_t)c * NB + i] + o[i];
}

// z = D_g^-1 y on the flux rows of block g of an active instance
__device__ __forceinline__ void st_flux(const nfx_batch_tg& pc, int g, const int32_t* __restrict__ mode,
                                        const float* __restrict__ y, float* __restrict__ z) {
  if (g >= pc.n_fb) return;
  if (mode && mode[pc.fb_inst[g]] != 0) return;
  int r[3];
  float yv[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    r[i] = pc.fb_rows[(int64_t)g * 3 + i];
    yv[i] = y[r[i]];
  }
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    float s = 0.f;
#pragma unroll
    for (int j = 0; j < 3; ++j) s += pc.fb_inv[(int64_t)g * 9 + i * 3 + j] * yv[j];
    z[r[i]] = s;
  }
}

template <int NA, int NBB>
__global__ __launch_bounds__(kTB) void k_btg_pre(const nfx_batch_tg pa, const nfx_batch_tg pb, int split,
                                                 const int32_t* __restrict__ mode, const float* __restrict__ r) {
  const int b = blockIdx.x;
  if (b < split)
    st_pre<NA>(pa, b * kTB + threadIdx.x, mode, r);
  else
    st_pre<NBB>(pb, (b - split) * kTB + threadIdx.x, mode, r);
}

template <int NA, int NBB>
__global__ __launch_bounds__(kTB) void k_btg_restrict(const nfx_batch_tg pa, const nfx_batch_tg pb, int split,
                                                      const int32_t* __restrict__ mode,
                                                      const float* __restrict__ r) {
  const int b = blockIdx.x, lane = threadIdx.x & 63;
  if (b < split)
    st_restrict<NA>(pa, (b * kTB + threadIdx.x) >> 6, lane, mode, r);
  else
    st_restrict<NBB>(pb, ((b - split) * kTB + threadIdx.x) >> 6, lane, mode, r);
}

__global__ __launch_bounds__(kTB) void k_btg_gemv(const nfx_batch_tg pa, const nfx_batch_tg pb, int split,
                                                  const int32_t* __restrict__ mode) {
  const int b = blockIdx.x, lane = threadIdx.x & 63;
  if (b < split)
    st_gemv(pa, (b * kTB + threadIdx.x) >> 6, lane, mode);
  else
    st_gemv(pb, ((b - split) * kTB + threadIdx.x) >> 6, lane, mode);
}

template <int NA, int NBB>
__global__ __launch_bounds__(kTB) void k_btg_post1(const nfx_batch_tg pa, const nfx_batch_tg pb, int split,
                                                   const int32_t* __restrict__ mode) {
  const int b = blockIdx.x;
  if (b < split)
    st_post1<NA>(pa, b * kTB + threadIdx.x, mode);
  else
    st_post1<NBB>(pb, (b - split) * kTB + threadIdx.x, mode);
}

// post2 of both groups, then the flux blocks of either group (blocks [split2, grid))
template <int NA, int NBB>
__global__ __launch_bounds__(kTB) void k_btg_post2(const nfx_batch_tg pa, const nfx_batch_tg pb, int split,
                                                   int split2, int split3, const int32_t* __restrict__ mode,
                                                   const float* __restrict__ y, float* __restrict__ z) {
  const int b = blockIdx.x;
  if (b < split)
    st_post2<NA>(pa, b * kTB + threadIdx.x, mode, z);
  else if (b < split2)
    st_post2<NBB>(pb, (b - split) * kTB + threadIdx.x, mode, z);
  else if (b < split3)
    st_flux(pa, (b - split2) * kTB + threadIdx.x, mode, y, z);
  else
    st_flux(pb, (b - split3) * kTB + threadIdx.x, mode, y, z);
}

template <int NA, int NBB>
int apply_pair(const nfx_batch_tg& pa, const nfx_batch_tg* pbp, const int32_t* mode, const float* r, float* z,
               hipStream_t st) {
  // a single group runs as the pair (pa, pa) with no blocks for the second member
  const nfx_batch_tg& pb = pbp ? *pbp : pa;
  const bool two = pbp != nullptr;
  const int ca = grid_for(pa.n_cells, kTB), cb = two ? grid_for(pb.n_cells, kTB) : 0;
  const int aa = grid_for((int64_t)pa.n_agg * 64, kTB), ab = two ? grid_for((int64_t)pb.n_agg * 64, kTB) : 0;
  const int fa = grid_for(pa.n_fb, kTB), fb = two ? grid_for(pb.n_fb, kTB) : 0;
  hipLaunchKernelGGL((k_btg_pre<NA, NBB>), dim3(ca + cb), dim3(kTB), 0, st, pa, pb, ca, mode, r);
  hipLaunchKernelGGL((k_btg_restrict<NA, NBB>), dim3(aa + ab), dim3(kTB), 0, st, pa, pb, aa, mode, r);
  hipLaunchKernelGGL(k_btg_gemv, dim3(aa + ab), dim3(kTB), 0, st, pa, pb, aa, mode);
  hipLaunchKernelGGL((k_btg_post1<NA, NBB>), dim3(ca + cb), dim3(kTB), 0, st, pa, pb, ca, mode);
  hipLaunchKernelGGL((k_btg_post2<NA, NBB>), dim3(ca + cb + fa + fb), dim3(kTB), 0, st, pa, pb, ca, ca + cb,
                     ca + cb + fa, mode, r, z);
  NFX_LAUNCH_CHECK();
  return NFX_OK;
}

int check_desc(const nfx_batch_tg* pc) {
  NFX_REQUIRE(pc, "nfx_batch_tg: null descriptor");
  NFX_REQUIRE(pc->nb == 1 || pc->nb == 4, "nfx_batch_tg: nb must be 1 or 4");
  NFX_REQUIRE(pc->n_cells > 0 && pc->n_agg > 0, "nfx_batch_tg: empty group");
  NFX_REQUIRE(pc->cell_inst && pc->prow && pc->nbr && pc->sd && pc->so && pc->rs && pc->wd && pc->agg_ptr &&
                  pc->agg_mem && pc->agg_of && pc->arow && pc->a0 && pc->alen && pc->ainv && pc->work,
              "nfx_batch_tg: null pointer");
  NFX_REQUIRE(pc->nb == 1 || !pc->wo, "nfx_batch_tg: smoother couplings need nb = 1");
  NFX_REQUIRE(pc->n_fb >= 0 && (pc->n_fb == 0 || (pc->fb_inst && pc->fb_rows && pc->fb_inv)),
              "nfx_batch_tg: incomplete flux blocks");
  return NFX_OK;
}

}  // namespace

int batch_tg_apply(const nfx_batch_tg* pc, int n_groups, const int32_t* mode, const float* r, float* z,
                   hipStream_t st) {
  NFX_REQUIRE(pc && (n_groups == 1 || n_groups == 2), "nfx_batch_tg_apply: one or two groups");
  for (int g = 0; g < n_groups; ++g) {
    const int rc = check_desc(pc + g);
    if (rc) return rc;
  }
  if (n_groups == 1)
    return pc->nb == 1 ? apply_pair<1, 1>(pc[0], nullptr, mode, r, z, st) : apply_pair<4, 4>(pc[0], nullptr, mode, r, z, st);
  NFX_REQUIRE(pc[0].nb != pc[1].nb, "nfx_batch_tg_apply: two groups need different nb");
  return pc[0].nb == 1 ? apply_pair<1, 4>(pc[0], pc + 1, mode, r, z, st) : apply_pair<1, 4>(pc[1], pc, mode, r, z, st);
}

}  // namespace nfx

using namespace nfx;

extern "C" {

int nfx_batch_tg_schur(int nb, int n_cells, int nbf, int fd, const int32_t* d_cell_inst, const double* const* d_vals,
                       const int32_t* d_posb, const int32_t* d_posd, const int32_t* d_nbr, const int32_t* d_fdofs,
                       const int32_t* d_nloc, double* d_sd64, double* d_so64, float* d_sd, float* d_so,
                       void* stream) {
  NFX_REQUIRE(nb == 1 || nb == 4, "nfx_batch_tg_schur: nb must be 1 or 4");
  NFX_REQUIRE(n_cells >= 0 && nbf > 0 && fd > 0 && 4 * fd <= nbf, "nfx_batch_tg_schur: bad sizes");
  if (n_cells == 0) return NFX_OK;
  NFX_REQUIRE(d_cell_inst && d_vals && d_posb && d_posd && d_nbr && d_fdofs && d_nloc && d_sd64 && d_so64 && d_sd &&
                  d_so,
              "nfx_batch_tg_schur: null pointer");
  hipStream_t st = (hipStream_t)stream;
  if (nb == 1)
    hipLaunchKernelGGL(k_btg_schur<1>, dim3(grid_for(n_cells, kTB)), dim3(kTB), 0, st, n_cells, nbf, fd, d_cell_inst,
                       d_vals, d_posb, d_posd, d_nbr, d_fdofs, d_nloc, d_sd64, d_so64, d_sd, d_so);
  else
    hipLaunchKernelGGL(k_btg_schur<4>, dim3(grid_for(n_cells, kTB)), dim3(kTB), 0, st, n_cells, nbf, fd, d_cell_inst,
                       d_vals, d_posb, d_posd, d_nbr, d_fdofs, d_nloc, d_sd64, d_so64, d_sd, d_so);
  NFX_LAUNCH_CHECK();
  return NFX_OK;
}

int nfx_batch_tg_schur_blk(int n_cells, const int32_t* d_cell_inst, const double* const* d_vals,
                           const int32_t* d_posb, const int32_t* d_posmf, const int32_t* d_posmi,
                           const int32_t* d_nbr, const int32_t* d_fdofs, const int32_t* d_nloc, double* d_sd64,
                           double* d_so64, float* d_sd, float* d_so, int32_t* d_status, void* stream) {
  NFX_REQUIRE(n_cells >= 0, "nfx_batch_tg_schur_blk: n_cells < 0");
  if (n_cells == 0) return NFX_OK;
  NFX_REQUIRE(d_cell_inst && d_vals && d_posb && d_posmf && d_posmi && d_nbr && d_fdofs && d_nloc && d_sd64 &&
                  d_so64 && d_sd && d_so && d_status,
              "nfx_batch_tg_schur_blk: null pointer");
  hipLaunchKernelGGL(k_btg_schur_blk, dim3(grid_for(n_cells, kTB)), dim3(kTB), 0, (hipStream_t)stream, n_cells,
                     d_cell_inst, d_vals, d_posb, d_posmf, d_posmi, d_nbr, d_fdofs, d_nloc, d_sd64, d_so64, d_sd,
                     d_so, d_status);
  NFX_LAUNCH_CHECK();
  return NFX_OK;
}

int nfx_batch_tg_fluxinv(int n_fb, const int32_t* d_fb_inst, const double* const* d_vals, const int32_t* d_fb_pos,
                         float* d_fb_inv, int32_t* d_status, void* stream) {
  NFX_REQUIRE(n_fb >= 0, "nfx_batch_tg_fluxinv: n_fb < 0");
  if (n_fb == 0) return NFX_OK;
  NFX_REQUIRE(d_fb_inst && d_vals && d_fb_pos && d_fb_inv && d_status, "nfx_batch_tg_fluxinv: null pointer");
  hipLaunchKernelGGL(k_btg_fluxinv, dim3(grid_for(n_fb, kTB)), dim3(kTB), 0, (hipStream_t)stream, n_fb, d_fb_inst,
                     d_vals, d_fb_pos, d_fb_inv, d_status);
  NFX_LAUNCH_CHECK();
  return NFX_OK;
}

int nfx_batch_tg_smoother(int nb, int n_cells, const double* d_sd64, double omega, float* d_wd, int32_t* d_status,
                          void* stream) {
  NFX_REQUIRE(nb == 1 || nb == 4, "nfx_batch_tg_smoother: nb must be 1 or 4");
  NFX_REQUIRE(n_cells >= 0, "nfx_batch_tg_smoother: n_cells < 0");
  if (n_cells == 0) return NFX_OK;
  NFX_REQUIRE(d_sd64 && d_wd && d_status, "nfx_batch_tg_smoother: null pointer");
  hipStream_t st = (hipStream_t)stream;
  if (nb == 1)
    hipLaunchKernelGGL(k_btg_smoother<1>, dim3(grid_for(n_cells, kTB)), dim3(kTB), 0, st, n_cells, d_sd64, omega,
                       d_wd, d_status);
  else
    hipLaunchKernelGGL(k_btg_smoother<4>, dim3(grid_for(n_cells, kTB)), dim3(kTB), 0, st, n_cells, d_sd64, omega,
                       d_wd, d_status);
  NFX_LAUNCH_CHECK();
  return NFX_OK;
}

int nfx_batch_tg_coarse(int nb, int n_agg, const int32_t* d_agg_ptr, const int32_t* d_agg_mem,
                        const int32_t* d_agg_of, const int32_t* d_nbr, const double* d_sd64, const double* d_so64,
                        const int64_t* d_arow, const int32_t* d_a0, double* d_ac, void* stream) {
  NFX_REQUIRE(nb == 1 || nb == 4, "nfx_batch_tg_coarse: nb must be 1 or 4");
  NFX_REQUIRE(n_agg >= 0, "nfx_batch_tg_coarse: n_agg < 0");
  if (n_agg == 0) return NFX_OK;
  NFX_REQUIRE(d_agg_ptr && d_agg_mem && d_agg_of && d_nbr && d_sd64 && d_so64 && d_arow && d_a0 && d_ac,
              "nfx_batch_tg_coarse: null pointer");
  hipStream_t st = (hipStream_t)stream;
  if (nb == 1)
    hipLaunchKernelGGL(k_btg_coarse<1>, dim3(grid_for(n_agg, 64)), dim3(64), 0, st, n_agg, d_agg_ptr, d_agg_mem,
                       d_agg_of, d_nbr, d_sd64, d_so64, d_arow, d_a0, d_ac);
  else
    hipLaunchKernelGGL(k_btg_coarse<4>, dim3(grid_for(n_agg, 64)), dim3(64), 0, st, n_agg, d_agg_ptr, d_agg_mem,
                       d_agg_of, d_nbr, d_sd64, d_so64, d_arow, d_a0, d_ac);
  NFX_LAUNCH_CHECK();
  return NFX_OK;
}

int nfx_batch_tg_rowsums(int nb, int n_cells, const double* d_sd64, const double* d_so64, float* d_rs,
                         void* stream) {
  NFX_REQUIRE(nb == 1 || nb == 4, "nfx_batch_tg_rowsums: nb must be 1 or 4");
  NFX_REQUIRE(n_cells >= 0, "nfx_batch_tg_rowsums: n_cells < 0");
  if (n_cells == 0) return NFX_OK;
  NFX_REQUIRE(d_sd64 && d_so64 && d_rs, "nfx_batch_tg_rowsums: null pointer");
  hipStream_t st = (hipStream_t)stream;
  if (nb == 1)
    hipLaunchKernelGGL(k_btg_rowsums<1>, dim3(grid_for(n_cells, kTB)), dim3(kTB), 0, st, n_cells, d_sd64, d_so64, d_rs);
  else
    hipLaunchKernelGGL(k_btg_rowsums<4>, dim3(grid_for(n_cells, kTB)), dim3(kTB), 0, st, n_cells, d_sd64, d_so64, d_rs);
  NFX_LAUNCH_CHECK();
  return NFX_OK;
}

int nfx_batch_tg_apply(int n_groups, const nfx_batch_tg* pc, const int32_t* d_mode, const float* d_r, float* d_z,
                       void* stream) {
  NFX_REQUIRE(d_r && d_z, "nfx_batch_tg_apply: null vector");
  return nfx::batch_tg_apply(pc, n_groups, d_mode, d_r, d_z, (hipStream_t)stream);
}

}  // extern "C"
