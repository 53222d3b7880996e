// Dual-mixed linear elasticity, BDM1 stress rows - P0 displacement - P0 rotation,
// tetrahedra: the element operator of LEDualWeakForm.evaluate_form_vectorized
// (mode 0) and of LERieszMapWeakForm.evaluate_form_vectorized (mode 1).
//
// One wavefront assembles one element (4 elements per 256-thread block). The
// pushed-forward stress basis psi_a(x_q) = J phi_a / detJ (12 functions x 8
// points x 3 components), its divergence and the per-point weights are staged
// in LDS; lanes then evaluate the 903 entries (I <= J) of the symmetric 42 x 42
// matrix and store them through the slot map, which is element-major (the codes of
// one element are contiguous: a wave reads them with coalesced loads; an entry-major
// map measured 379 us per assembly at 16^3 x 6). Dense maps hold 1,764 codes per
// element, (I, J) and (J, I); pair maps 903, one per upper-triangle entry (an upper
// SBAIJ operator keeps at most one of the two: half the code traffic).
//
// Reference local order: stress (a, c) -> 3a + c (a < 12), displacement u_c ->
// 36 + c, rotation t_c -> 39 + c. Entries:
//   (s_ac, s_bd) = [c == d] sum_q k1 psi_a . psi_b - sum_q k2 psi_ac psi_bd
//                  (+ [c == d] sum_q w detJ div_a div_b for the Riesz map)
//       k1 = w detJ / (2 mu), k2 = w detJ (1 / 2mu) lambda / (2mu + 3 lambda)
//   (u_c, s_bd)  = [c == d] sum_q w divref_b                        (mode 0)
//   (t_c, s_bd)  = eps(c, d, k) sum_q w detJ psi_bk                  (mode 0)
//   (u_c, u_c), (t_c, t_c) = sum_q w detJ                            (mode 1)
//   r(u_c) = -sum_q w detJ f_c,  r(t_c) = -sum_q w detJ f_{3+c}      (mode 0)
#include "nfx_common.h"

namespace nfx {

constexpr int kEB = 12;        // BDM1 basis functions
constexpr int kEQ = 8;         // quadrature points (degree 3)
constexpr int kEL = 42;        // local DoFs
constexpr int kWavesPerBlock = 4;

template <int MODE, typename R>
__global__ __launch_bounds__(64 * kWavesPerBlock) void k_elasticity(
    int n_elem, const int32_t* __restrict__ cells, const R* __restrict__ coords,
    const R* __restrict__ tab, const R* __restrict__ lam, const R* __restrict__ mu, int cs_e,
    int cs_q, const R* __restrict__ f, int fs_e, int fs_q, int fs_c,
    const int32_t* __restrict__ slot, int pairs, const int32_t* __restrict__ rhs_row, R* __restrict__ vals,
    R* __restrict__ stage, R* __restrict__ rhs) {
  __shared__ R s_psi[kWavesPerBlock][kEB * kEQ * 3];
  __shared__ R s_div[kWavesPerBlock][kEB * kEQ];
  __shared__ R s_k1[kWavesPerBlock][kEQ], s_k2[kWavesPerBlock][kEQ], s_wd[kWavesPerBlock][kEQ];
  __shared__ R s_A[kWavesPerBlock][kEB * 3];
  __shared__ R s_U[kWavesPerBlock][kEB];
  const int wv = threadIdx.x >> 6;
  const int lane = threadIdx.x & 63;
  const int e = blockIdx.x * kWavesPerBlock + wv;
  const bool live = e < n_elem;
  const int ee = live ? e : 0;
  const R* tw = tab;
  const R* tphi = tab + kEQ;
  const R* tdiv = tphi + kEQ * kEB * 3;

  // geometry (every lane, same addresses)
  R X[4][3];
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const int v = cells[ee * 4 + a];
#pragma unroll
    for (int c = 0; c < 3; ++c) X[a][c] = coords[3 * v + c];
  }
  R J[3][3];
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int d = 0; d < 3; ++d) J[c][d] = X[d + 1][c] - X[0][c];
  R detJ = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
           J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
           J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
  detJ = detJ < R(0) ? -detJ : detJ;
  const R invdet = R(1) / detJ;

  for (int idx = lane; idx < kEB * kEQ; idx += 64) {
    const int q = idx / kEB, a = idx % kEB;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      R s = 0;
#pragma unroll
      for (int d = 0; d < 3; ++d) s += J[c][d] * tphi[(q * kEB + a) * 3 + d];
      s_psi[wv][(a * kEQ + q) * 3 + c] = s * invdet;
    }
    s_div[wv][a * kEQ + q] = tdiv[q * kEB + a];  // reference divergence
  }
  if (lane < kEQ) {
    const int q = lane;
    const R m = mu[(int64_t)ee * cs_e + q * cs_q];
    const R l = lam[(int64_t)ee * cs_e + q * cs_q];
    const R wd = tw[q] * detJ;
    s_wd[wv][q] = wd;
    s_k1[wv][q] = wd / (R(2) * m);
    s_k2[wv][q] = wd * (R(1) / (R(2) * m)) * (l / (R(2) * m + R(3) * l));
  }
  __syncthreads();
  if (MODE == 0) {
    if (lane < kEB * 3) {
      const int b = lane / 3, k = lane % 3;
      R s = 0;
      for (int q = 0; q < kEQ; ++q) s += s_wd[wv][q] * s_psi[wv][(b * kEQ + q) * 3 + k];
      s_A[wv][b * 3 + k] = s;
    } else if (lane < kEB * 3 + kEB) {
      const int b = lane - kEB * 3;
      R s = 0;
      for (int q = 0; q < kEQ; ++q) s += tw[q] * s_div[wv][b * kEQ + q];
      s_U[wv][b] = s;
    }
  }
  __syncthreads();
  if (!live) return;

  constexpr int kPairs = kEL * (kEL + 1) / 2;
  const int32_t* es = slot + (int64_t)e * (pairs ? kPairs : kEL * kEL);
  for (int p = lane; p < kEL * (kEL + 1) / 2; p += 64) {
    // p-th entry of the upper triangle in row-major order
    int I = 0, rem = p;
    while (rem >= kEL - I) {
      rem -= kEL - I;
      ++I;
    }
    const int Jj = I + rem;
    R v = 0;
    if (Jj < 36) {
      const int a = I / 3, c = I % 3, b = Jj / 3, d = Jj % 3;
      const R* pa = &s_psi[wv][a * kEQ * 3];
      const R* pb = &s_psi[wv][b * kEQ * 3];
      R vol = 0, dot = 0, dd = 0;
#pragma unroll
      for (int q = 0; q < kEQ; ++q) {
        vol += s_k2[wv][q] * pa[q * 3 + c] * pb[q * 3 + d];
        if (c == d) {
          dot += s_k1[wv][q] * (pa[q * 3] * pb[q * 3] + pa[q * 3 + 1] * pb[q * 3 + 1] +
                                pa[q * 3 + 2] * pb[q * 3 + 2]);
          if (MODE == 1)
            dd += s_wd[wv][q] * s_div[wv][a * kEQ + q] * s_div[wv][b * kEQ + q] * invdet * invdet;
        }
      }
      v = dot - vol + dd;
    } else if (I < 36) {
      if (MODE == 0) {
        const int b = I / 3, d = I % 3;
        if (Jj < 39) {
          v = (Jj - 36 == d) ? s_U[wv][b] : R(0);
        } else {
          const int c = Jj - 39;
          // eps(c, d, k) with k the third index
          if (c != d) {
            const int k = 3 - c - d;
            const bool even = (c == 0 && d == 1) || (c == 1 && d == 2) || (c == 2 && d == 0);
            v = even ? s_A[wv][b * 3 + k] : -s_A[wv][b * 3 + k];
          }
        }
      }
    } else {
      if (MODE == 1 && I == Jj) {
        R s = 0;
#pragma unroll
        for (int q = 0; q < kEQ; ++q) s += s_wd[wv][q];
        v = s;
      }
    }
    if (pairs) {
      put_slot<R>(es[p], v, vals, stage);
    } else {
      put_slot<R>(es[I * kEL + Jj], v, vals, stage);
      if (I != Jj) put_slot<R>(es[Jj * kEL + I], v, vals, stage);
    }
  }
  const int64_t E = n_elem;
  if (MODE == 0 && lane < 6) {
    R s = 0;
#pragma unroll
    for (int q = 0; q < kEQ; ++q)
      s -= s_wd[wv][q] * f[(int64_t)e * fs_e + q * fs_q + lane * fs_c];
    put_row<R>(rhs_row[(int64_t)lane * E + e], s, rhs);
  }
}

template <int MODE, typename R>
static int launch_elasticity(int n, const int32_t* cells, const void* coords, const void* tab,
                             const void* lam, const void* mu, int cse, int csq, const void* f,
                             int fse, int fsq, int fsc, const int32_t* slot, int pairs, const int32_t* rhs_row,
                             void* vals, void* stage, void* rhs, hipStream_t s) {
  hipLaunchKernelGGL((k_elasticity<MODE, R>), dim3(grid_for(n, kWavesPerBlock)),
                     dim3(64 * kWavesPerBlock), 0, s, n, cells, static_cast<const R*>(coords),
                     static_cast<const R*>(tab), static_cast<const R*>(lam),
                     static_cast<const R*>(mu), cse, csq, static_cast<const R*>(f), fse, fsq, fsc,
                     slot, pairs, rhs_row, static_cast<R*>(vals), static_cast<R*>(stage),
                     static_cast<R*>(rhs));
  NFX_LAUNCH_CHECK();
  return NFX_OK;
}

// ---------------------------------------------------------------------------
// Face-owner assembly: one wavefront owns the 9 CSR rows of one face (its 3 BDM1 DoFs x 3
// stress rows, consecutive in the DoF numbering and therefore one contiguous CSR slice),
// evaluates the rows of those DoFs in the element matrices of the face's (at most two)
// incident tetrahedra, sums them in element order and writes the slice with coalesced
// stores. No slot codes per element entry, no staging, no atomics. Each CSR entry of a
// stress row carries a 16-bit source code: bits 0..3 the row inside the face (3 j + c),
// bits 4..9 and 10..15 the element-local reference column (3 b + d, 36 + d for u_d,
// 39 + d for t_d) of the entry in the first / second incident element, 63 = no
// contribution (entries dropped by the value filter, explicit zero diagonals).
//
// Per incident element the wave computes only the columns the upper triangle needs: the
// blocks of the element's faces whose DoF block index is >= the owning face's, plus the
// u / t columns. The (s, s) contraction V = sum_q k2_q psi_ac(q) psi_bd(q) (9 rows x 9 nb
// columns, K = 8 quadrature points) runs on v_mfma_f32_16x16x4_f32 in FP32; FP64 builds
// evaluate the same tiles on the vector ALU. The diagonal-in-component part
// D_ab = sum_q k1_q psi_a . psi_b (and the Riesz map's div-div term) is a 3 x 12 block per
// element on the vector ALU.
//
// Cell rows (u_c, t_c of each tetrahedron) hold one entry, their diagonal: 0 in A (the
// explicit zero of the SBAIJ scatter), the mass |K| in P; their residual rows take
// -sum_q w detJ f. One thread per cell, in blocks launched after the face blocks.
//
// Table layout (R): w[8] | phi[8][12][3] | div[8][12] | Uref[12] | Phi[12][3] | DDref[12][12]
// with Uref_b = sum_q w_q divref_b(q), Phi_be = sum_q w_q phihat_be(q),
// DDref_ab = sum_q w_q divref_a(q) divref_b(q).
constexpr int kFW = 4;  // faces (waves) per block of the face kernel
constexpr int kCellBlock = 256;
constexpr int kNoSrc = 63;

typedef float f4_t __attribute__((ext_vector_type(4)));

template <typename R>
struct FaceLds {
  R psi[2][12][3][kEQ];      // psi_{b d}(q)
  R kpsi[2][2][3][3][kEQ];   // [i][0][j][c][q] = k2_q psi_{a_j c}(q); [i][1] k1_q psi_{a_j c}(q)
  R D[2][3][12];             // D_{a_j b} (+ div-div, Riesz map)
  R K[2][9][kEL];            // signed element rows of the face: s(a) s(b) K_e[(a, c), col]
};

template <typename R>
__device__ __forceinline__ void face_geometry(const int32_t* __restrict__ cells, const R* __restrict__ coords,
                                              int e, R (&J)[3][3], R& detJ) {
  R X[4][3];
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const int v = cells[e * 4 + a];
#pragma unroll
    for (int c = 0; c < 3; ++c) X[a][c] = coords[3 * v + c];
  }
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int d = 0; d < 3; ++d) J[c][d] = X[d + 1][c] - X[0][c];
  R dt = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) - J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
         J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
  detJ = dt < R(0) ? -dt : dt;
}

// Per incident element of a face: Jacobian and |det J|, the element's 4 face DoF blocks,
// orientation word (bits 3b, 3b + 1: face slot of reference basis b; bit 3b + 2: its sign),
// the reference basis functions at the owning face's slots j (4 bits each) and the list of
// the element's local faces whose block is >= the owning face's (2 bits each, nb of them).
template <typename R>
struct FaceInc {
  R J[3][3];
  R detJ, invdet;
  uint64_t orient;
  int ajpack, blpack, nb;
};

template <typename R>
__device__ __forceinline__ R face_psi(const FaceInc<R>& g, int d, R p0, R p1, R p2) {
  return (g.J[d][0] * p0 + g.J[d][1] * p1 + g.J[d][2] * p2) * g.invdet;
}

template <typename R>
__device__ __forceinline__ R basis_sign(const FaceInc<R>& g, int b) {
  return ((g.orient >> (3 * b + 2)) & 1) ? R(-1) : R(1);
}

// rows of incidence i (uniform): (s, s) blocks of the needed faces into L.K[i]
template <typename R>
__device__ __forceinline__ void face_ss_rows(FaceLds<R>& L, const FaceInc<R>& g, int i, int l) {
  const int nc = 9 * g.nb;
  const int ntile = (nc + 15) >> 4;
  const int lr = l & 15, lq = l >> 4;
#pragma unroll
  for (int t = 0; t < 3; ++t) {
    if (t >= ntile) break;
    const int n = 16 * t + lr;
    const bool cn = n < nc;
    const int bi = cn ? n / 9 : 0, w = n - 9 * bi;
    const int lf = (g.blpack >> (2 * bi)) & 3;
    const int b = 3 * lf + w / 3, d = w % 3;
    R v[4];
    if constexpr (sizeof(R) == 4) {
      f4_t acc = f4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int s = 0; s < 2; ++s) {
        const int q = 4 * s + lq;
        const float av = lr < 9 ? L.kpsi[i][0][lr / 3][lr % 3][q] : 0.f;
        const float bv = cn ? L.psi[i][b][d][q] : 0.f;
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv, acc, 0, 0, 0);
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) v[r] = acc[r];
    } else {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = 4 * lq + r;
        R s = 0;
        if (m < 9 && cn)
#pragma unroll
          for (int q = 0; q < kEQ; ++q) s += L.kpsi[i][0][m / 3][m % 3][q] * L.psi[i][b][d][q];
        v[r] = s;
      }
    }
    if (cn) {
      const R sb = basis_sign(g, b);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = 4 * lq + r;
        if (m < 9) {
          const int j = m / 3, c = m % 3, a = (g.ajpack >> (4 * j)) & 15;
          const R kv = (c == d ? L.D[i][j][b] : R(0)) - v[r];
          L.K[i][m][9 * lf + w] = basis_sign(g, a) * sb * kv;
        }
      }
    }
  }
}

template <int MODE, typename R>
__global__ __launch_bounds__(64 * kFW) void k_elasticity_faces(
    int n_faces, int n_elem, int n_s, int n_u, const int32_t* __restrict__ face_inc,
    const int32_t* __restrict__ face_lf, const int32_t* __restrict__ el_face,
    const uint64_t* __restrict__ el_orient, const int32_t* __restrict__ cells, const R* __restrict__ coords,
    const R* __restrict__ tab, const R* __restrict__ lam, const R* __restrict__ mu, int cs_e, int cs_q,
    const R* __restrict__ f, int fs_e, int fs_q, int fs_c, const int32_t* __restrict__ rowptr,
    const uint16_t* __restrict__ codes, R* __restrict__ vals, R* __restrict__ rhs) {
  const R* tw = tab;
  const R* tphi = tab + kEQ;
  const R* tdiv = tphi + kEQ * kEB * 3;
  const R* tU = tdiv + kEQ * kEB;
  const R* tPhi = tU + kEB;
  const R* tDD = tPhi + kEB * 3;
  const int n_face_blocks = (n_faces + kFW - 1) / kFW;
  if ((int)blockIdx.x >= n_face_blocks) {
    // cell rows: u_c = n_s + 3 e + c, t_c = n_s + n_u + 3 e + c
    const int e = ((int)blockIdx.x - n_face_blocks) * kCellBlock + threadIdx.x;
    if (e >= n_elem) return;
    R J[3][3], detJ;
    face_geometry<R>(cells, coords, e, J, detJ);
    R mass = 0;
#pragma unroll
    for (int q = 0; q < kEQ; ++q) mass += tw[q] * detJ;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const int ru = n_s + 3 * e + c, rt = n_s + n_u + 3 * e + c;
      vals[rowptr[ru]] = MODE == 1 ? mass : R(0);
      vals[rowptr[rt]] = MODE == 1 ? mass : R(0);
      if (MODE == 0) {
        R su = 0, st = 0;
#pragma unroll
        for (int q = 0; q < kEQ; ++q) {
          const R wd = tw[q] * detJ;
          su -= wd * f[(int64_t)e * fs_e + q * fs_q + c * fs_c];
          st -= wd * f[(int64_t)e * fs_e + q * fs_q + (3 + c) * fs_c];
        }
        rhs[ru] = su;
        rhs[rt] = st;
      }
    }
    return;
  }
  __shared__ FaceLds<R> lds_all[kFW];
  const int wv = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int p = __builtin_amdgcn_readfirstlane((int)blockIdx.x * kFW + wv);
  if (p >= n_faces) return;
  FaceLds<R>& L = lds_all[wv];
  const int e0 = face_inc[2 * p], e1r = face_inc[2 * p + 1];
  const int ni = e1r >= 0 ? 2 : 1;
  const int e1 = ni == 2 ? e1r : e0;
  const int lfw = face_lf[p];
  FaceInc<R> g0, g1;
  {
    FaceInc<R>* gs[2] = {&g0, &g1};
    const int es[2] = {e0, e1};
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      FaceInc<R>& g = *gs[i];
      const int e = es[i];
      face_geometry<R>(cells, coords, e, g.J, g.detJ);
      g.invdet = R(1) / g.detJ;
      g.orient = el_orient[e];
      const int lf = (lfw >> (2 * i)) & 3;
      g.ajpack = 0;
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const int b = 3 * lf + k;
        g.ajpack |= b << (4 * (int)((g.orient >> (3 * b)) & 3));
      }
      g.nb = 0;
      g.blpack = 0;
#pragma unroll
      for (int k = 0; k < 4; ++k)
        if (el_face[e * 4 + k] >= p) {
          g.blpack |= k << (2 * g.nb);
          ++g.nb;
        }
    }
  }
  // pushed-forward basis psi_{b d}(q) of both elements
  for (int t = l; t < 2 * kEB * kEQ; t += 64) {
    const int i = t / (kEB * kEQ), rem = t % (kEB * kEQ), b = rem / kEQ, q = rem % kEQ;
    if (i >= ni) continue;
    const R p0 = tphi[(q * kEB + b) * 3], p1 = tphi[(q * kEB + b) * 3 + 1], p2 = tphi[(q * kEB + b) * 3 + 2];
#pragma unroll
    for (int d = 0; d < 3; ++d) L.psi[i][b][d][q] = i ? face_psi(g1, d, p0, p1, p2) : face_psi(g0, d, p0, p1, p2);
  }
  // coefficient-weighted rows of the owning face's basis functions
  if (l < 2 * 3 * kEQ) {
    const int i = l / (3 * kEQ), j = (l / kEQ) % 3, q = l % kEQ;
    if (i < ni) {
      const int e = i ? e1 : e0;
      const int a = ((i ? g1.ajpack : g0.ajpack) >> (4 * j)) & 15;
      const R m = mu[(int64_t)e * cs_e + q * cs_q];
      const R lm = lam[(int64_t)e * cs_e + q * cs_q];
      const R wd = tw[q] * (i ? g1.detJ : g0.detJ);
      const R k1 = wd / (R(2) * m);
      const R k2 = wd * (R(1) / (R(2) * m)) * (lm / (R(2) * m + R(3) * lm));
      const R p0 = tphi[(q * kEB + a) * 3], p1 = tphi[(q * kEB + a) * 3 + 1], p2 = tphi[(q * kEB + a) * 3 + 2];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const R ps = i ? face_psi(g1, c, p0, p1, p2) : face_psi(g0, c, p0, p1, p2);
        L.kpsi[i][0][j][c][q] = k2 * ps;
        L.kpsi[i][1][j][c][q] = k1 * ps;
      }
    }
  }
  __builtin_amdgcn_wave_barrier();
  // D_{a_j b} = sum_q k1_q psi_{a_j} . psi_b  (+ DDref_{a_j b} / detJ for the Riesz map)
  for (int t = l; t < 2 * 3 * kEB; t += 64) {
    const int i = t / (3 * kEB), j = (t / kEB) % 3, b = t % kEB;
    if (i >= ni) continue;
    R s = 0;
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
      for (int q = 0; q < kEQ; ++q) s += L.kpsi[i][1][j][c][q] * L.psi[i][b][c][q];
    if (MODE == 1) {
      const int a = ((i ? g1.ajpack : g0.ajpack) >> (4 * j)) & 15;
      s += tDD[a * kEB + b] * (i ? g1.invdet : g0.invdet);
    }
    L.D[i][j][b] = s;
  }
  __builtin_amdgcn_wave_barrier();
  // element rows of the face: (s, s) blocks of the needed faces, then the u / t columns
  face_ss_rows<R>(L, g0, 0, l);
  if (ni == 2) face_ss_rows<R>(L, g1, 1, l);
  if (MODE == 0) {
    for (int t = l; t < 2 * 9 * 6; t += 64) {
      const int i = t / 54, m = (t / 6) % 9, col = t % 6;
      if (i >= ni) continue;
      const int j = m / 3, c = m % 3;
      const int a = ((i ? g1.ajpack : g0.ajpack) >> (4 * j)) & 15;
      const R sa = i ? basis_sign(g1, a) : basis_sign(g0, a);
      R v = 0;
      if (col < 3) {
        v = col == c ? tU[a] : R(0);
      } else {
        const int ct = col - 3;
        if (ct != c) {
          const int k = 3 - c - ct;
          const R P0 = tPhi[a * 3], P1 = tPhi[a * 3 + 1], P2 = tPhi[a * 3 + 2];
          // row k of J by selects (a runtime index into a register array would go to scratch)
          R Jk[3];
#pragma unroll
          for (int x = 0; x < 3; ++x) {
            const R u0 = k == 0 ? g0.J[0][x] : (k == 1 ? g0.J[1][x] : g0.J[2][x]);
            const R u1 = k == 0 ? g1.J[0][x] : (k == 1 ? g1.J[1][x] : g1.J[2][x]);
            Jk[x] = i ? u1 : u0;
          }
          const R Aak = Jk[0] * P0 + Jk[1] * P1 + Jk[2] * P2;
          const bool even = (ct == 0 && c == 1) || (ct == 1 && c == 2) || (ct == 2 && c == 0);
          v = even ? Aak : -Aak;
        }
      }
      L.K[i][m][36 + col] = sa * v;
    }
  }
  __builtin_amdgcn_wave_barrier();
  // the face's CSR slice: contiguous, coalesced stores
  const int base = rowptr[9 * p];
  const int cnt = rowptr[9 * p + 9] - base;
  for (int k = l; k < cnt; k += 64) {
    const int code = codes[base + k];
    const int r = code & 15, s0 = (code >> 4) & 63, s1 = (code >> 10) & 63;
    R v = 0;
    if (s0 != kNoSrc) v += L.K[0][r][s0];
    if (s1 != kNoSrc) v += L.K[1][r][s1];
    vals[base + k] = v;
  }
  if (MODE == 0 && l < 9) rhs[9 * p + l] = R(0);
}

template <int MODE, typename R>
static int launch_elasticity_faces(int n_faces, int n_elem, int n_s, int n_u, const int32_t* face_inc,
                                   const int32_t* face_lf, const int32_t* el_face, const uint64_t* el_orient,
                                   const int32_t* cells, const void* coords, const void* tab, const void* lam,
                                   const void* mu, int cse, int csq, const void* f, int fse, int fsq, int fsc,
                                   const int32_t* rowptr, const uint16_t* codes, void* vals, void* rhs,
                                   hipStream_t s) {
  const int blocks = grid_for(n_faces, kFW) + grid_for(n_elem, kCellBlock);
  hipLaunchKernelGGL((k_elasticity_faces<MODE, R>), dim3(blocks), dim3(64 * kFW), 0, s, n_faces, n_elem, n_s, n_u,
                     face_inc, face_lf, el_face, el_orient, cells, static_cast<const R*>(coords),
                     static_cast<const R*>(tab), static_cast<const R*>(lam), static_cast<const R*>(mu), cse, csq,
                     static_cast<const R*>(f), fse, fsq, fsc, rowptr, codes, static_cast<R*>(vals),
                     static_cast<R*>(rhs));
  NFX_LAUNCH_CHECK();
  return NFX_OK;
}

// gather: dst[i] = src[idx[i]] (FP32 or FP64)
template <typename R>
__global__ __launch_bounds__(256) void k_gather(int64_t n, const int32_t* __restrict__ idx,
                                                const R* __restrict__ src, R* __restrict__ dst) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) dst[i] = src[idx[i]];
}

// In-place Gauss-Jordan inverse of an SPD N x N block (FP64, no pivoting: the pivots of
// an SPD matrix are positive); returns false on a non-positive pivot. The result is
// symmetrised, (A^-1 + A^-T) / 2.
template <int N>
__device__ bool spd_block_inverse(double (&A)[N][N]) {
  for (int k = 0; k < N; ++k) {
    const double piv = A[k][k];
    if (!(piv > 0.0)) return false;
    const double ip = 1.0 / piv;
    A[k][k] = 1.0;
    for (int j = 0; j < N; ++j) A[k][j] *= ip;
    for (int i = 0; i < N; ++i) {
      if (i == k) continue;
      const double f = A[i][k];
      A[i][k] = 0.0;
      for (int j = 0; j < N; ++j) A[i][j] -= f * A[k][j];
    }
  }
  for (int i = 0; i < N; ++i)
    for (int j = i + 1; j < N; ++j) {
      const double v = 0.5 * (A[i][j] + A[j][i]);
      A[i][j] = v;
      A[j][i] = v;
    }
  return true;
}

// Face blocks of the stress mass block: rows 9b .. 9b+8 restricted to columns of the same
// block, inverted; the inverse goes to dinv (FP64) and to the preconditioner's FP32
// values (9 per row, columns 9b .. 9b+8).
__global__ __launch_bounds__(64) void k_face_block_inverse(int nb, const int32_t* __restrict__ rp,
                                                           const int32_t* __restrict__ ci,
                                                           const double* __restrict__ va,
                                                           double* __restrict__ dinv, float* __restrict__ pc_val,
                                                           int* __restrict__ status) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  double D[9][9];
  for (int i = 0; i < 9; ++i)
    for (int j = 0; j < 9; ++j) D[i][j] = 0.0;
  for (int i = 0; i < 9; ++i) {
    const int r = 9 * b + i;
    for (int k = rp[r]; k < rp[r + 1]; ++k) {
      const int c = ci[k] - 9 * b;
      if (c >= 0 && c < 9) D[i][c] = va[k];
    }
  }
  if (!spd_block_inverse<9>(D)) {
    atomicOr(status, 1);
    return;
  }
  for (int i = 0; i < 9; ++i)
    for (int j = 0; j < 9; ++j) {
      dinv[(int64_t)b * 81 + i * 9 + j] = D[i][j];
      pc_val[((int64_t)9 * b + i) * 9 + j] = (float)D[i][j];
    }
}

// Cell blocks of the Schur complement S = B D^-1 B^T on (u, t): rows of cell c are
// ns + 3c + k (u) and ns + nu + 3c + k (t), k < 3; B their stress columns, grouped by
// face block (at most 4 faces per cell). S is inverted and written as the
// preconditioner's FP32 values (6 per row, columns the cell's u then t DoFs).
__global__ __launch_bounds__(64) void k_cell_schur_inverse(int n_cells, int ns, int nu,
                                                           const int32_t* __restrict__ rp,
                                                           const int32_t* __restrict__ ci,
                                                           const double* __restrict__ va,
                                                           const double* __restrict__ dinv,
                                                           float* __restrict__ pc_val, int* __restrict__ status) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n_cells) return;
  int rows[6];
  for (int k = 0; k < 3; ++k) {
    rows[k] = ns + 3 * c + k;
    rows[3 + k] = ns + nu + 3 * c + k;
  }
  int faces[4];
  int nf = 0;
  bool over = false;
  for (int a = 0; a < 6; ++a)
    for (int k = rp[rows[a]]; k < rp[rows[a] + 1]; ++k) {
      const int col = ci[k];
      if (col >= ns) continue;
      const int f = col / 9;
      bool seen = false;
      for (int t = 0; t < nf; ++t) seen = seen || faces[t] == f;
      if (seen) continue;
      if (nf < 4)
        faces[nf++] = f;
      else
        over = true;
    }
  if (over) {
    atomicOr(status, 2);
    return;
  }
  double S[6][6];
  for (int a = 0; a < 6; ++a)
    for (int b = 0; b < 6; ++b) S[a][b] = 0.0;
  for (int t = 0; t < nf; ++t) {
    const int f = faces[t];
    double Bf[6][9];
    for (int a = 0; a < 6; ++a)
      for (int j = 0; j < 9; ++j) Bf[a][j] = 0.0;
    for (int a = 0; a < 6; ++a)
      for (int k = rp[rows[a]]; k < rp[rows[a] + 1]; ++k) {
        const int col = ci[k];
        if (col < ns && col / 9 == f) Bf[a][col - 9 * f] = va[k];
      }
    const double* Di = dinv + (int64_t)f * 81;
    for (int a = 0; a < 6; ++a) {
      double T[9];
      for (int j = 0; j < 9; ++j) {
        double sacc = 0.0;
        for (int i = 0; i < 9; ++i) sacc += Bf[a][i] * Di[i * 9 + j];
        T[j] = sacc;
      }
      for (int b = 0; b < 6; ++b) {
        double sacc = 0.0;
        for (int j = 0; j < 9; ++j) sacc += T[j] * Bf[b][j];
        S[a][b] += sacc;
      }
    }
  }
  if (!spd_block_inverse<6>(S)) {
    atomicOr(status, 4);
    return;
  }
  for (int a = 0; a < 6; ++a) {
    const int64_t r = rows[a] - ns;  // preconditioner row among the (u, t) rows
    for (int b = 0; b < 6; ++b) pc_val[(int64_t)ns * 9 + r * 6 + b] = (float)S[a][b];
  }
}

}  // namespace nfx

using namespace nfx;

extern "C" int nfx_elasticity_mass_schur_pc(int n, int n_s, int n_u, const int32_t* d_rowptr, const int32_t* d_col,
                                            const double* d_vals64, double* d_dinv, float* d_pc_val,
                                            int* d_status, void* stream) {
  NFX_REQUIRE(n_s >= 0 && n_u >= 0 && n_s % 9 == 0 && n_u % 3 == 0 && n == n_s + 2 * n_u,
              "nfx_elasticity_mass_schur_pc: sizes");
  if (n == 0) return NFX_OK;
  NFX_REQUIRE(d_rowptr && d_col && d_vals64 && d_dinv && d_pc_val && d_status,
              "nfx_elasticity_mass_schur_pc: null pointer");
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int nb = n_s / 9, n_cells = n_u / 3;
  if (nb)
    hipLaunchKernelGGL(k_face_block_inverse, dim3(grid_for(nb, 64)), dim3(64), 0, st, nb, d_rowptr, d_col, d_vals64,
                       d_dinv, d_pc_val, d_status);
  if (n_cells)
    hipLaunchKernelGGL(k_cell_schur_inverse, dim3(grid_for(n_cells, 64)), dim3(64), 0, st, n_cells, n_s, n_u,
                       d_rowptr, d_col, d_vals64, d_dinv, d_pc_val, d_status);
  NFX_LAUNCH_CHECK();
  return NFX_OK;
}

extern "C" int nfx_assemble_elasticity(int mode, int n_elem, const int32_t* d_cells,
                                       const void* d_coords, const void* d_tables,
                                       const void* d_lam, const void* d_mu, int coef_stride_e,
                                       int coef_stride_q, const void* d_f, int f_stride_e,
                                       int f_stride_q, int f_stride_c, const int32_t* d_slot, int slot_pairs,
                                       const int32_t* d_rhs_row, int fp64, void* d_vals,
                                       void* d_stage, void* d_rhs, void* stream) {
  NFX_REQUIRE(mode == 0 || mode == 1, "nfx_assemble_elasticity: mode");
  NFX_REQUIRE(n_elem >= 0, "nfx_assemble_elasticity: n_elem < 0");
  if (n_elem == 0) return NFX_OK;
  NFX_REQUIRE(d_cells && d_coords && d_tables && d_lam && d_mu && d_slot && d_vals,
              "nfx_assemble_elasticity: null pointer");
  NFX_REQUIRE(mode == 1 || (d_f && d_rhs_row && d_rhs), "nfx_assemble_elasticity: residual args");
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (mode == 0)
    return fp64 ? launch_elasticity<0, double>(n_elem, d_cells, d_coords, d_tables, d_lam, d_mu,
                                               coef_stride_e, coef_stride_q, d_f, f_stride_e,
                                               f_stride_q, f_stride_c, d_slot, slot_pairs, d_rhs_row, d_vals,
                                               d_stage, d_rhs, s)
                : launch_elasticity<0, float>(n_elem, d_cells, d_coords, d_tables, d_lam, d_mu,
                                              coef_stride_e, coef_stride_q, d_f, f_stride_e,
                                              f_stride_q, f_stride_c, d_slot, slot_pairs, d_rhs_row, d_vals,
                                              d_stage, d_rhs, s);
  return fp64 ? launch_elasticity<1, double>(n_elem, d_cells, d_coords, d_tables, d_lam, d_mu,
                                             coef_stride_e, coef_stride_q, nullptr, 0, 0, 0,
                                             d_slot, slot_pairs, nullptr, d_vals, d_stage, nullptr, s)
              : launch_elasticity<1, float>(n_elem, d_cells, d_coords, d_tables, d_lam, d_mu,
                                            coef_stride_e, coef_stride_q, nullptr, 0, 0, 0,
                                            d_slot, slot_pairs, nullptr, d_vals, d_stage, nullptr, s);
}

extern "C" int nfx_assemble_elasticity_faces(int mode, int n_faces, int n_elem, int n_s, int n_u,
                                             const int32_t* d_face_inc, const int32_t* d_face_lf,
                                             const int32_t* d_el_face, const uint64_t* d_el_orient,
                                             const int32_t* d_cells, const void* d_coords, const void* d_tables,
                                             const void* d_lam, const void* d_mu, int coef_stride_e,
                                             int coef_stride_q, const void* d_f, int f_stride_e, int f_stride_q,
                                             int f_stride_c, const int32_t* d_rowptr, const uint16_t* d_codes,
                                             int fp64, void* d_vals, void* d_rhs, void* stream) {
  NFX_REQUIRE(mode == 0 || mode == 1, "nfx_assemble_elasticity_faces: mode");
  NFX_REQUIRE(n_faces >= 0 && n_elem >= 0 && n_s == 9 * n_faces && n_u == 3 * n_elem,
              "nfx_assemble_elasticity_faces: sizes (n_s = 9 n_faces, n_u = 3 n_elem)");
  if (n_faces == 0 && n_elem == 0) return NFX_OK;
  NFX_REQUIRE(d_face_inc && d_face_lf && d_el_face && d_el_orient && d_cells && d_coords && d_tables && d_lam &&
                  d_mu && d_rowptr && d_codes && d_vals,
              "nfx_assemble_elasticity_faces: null pointer");
  NFX_REQUIRE(mode == 1 || (d_f && d_rhs), "nfx_assemble_elasticity_faces: residual args");
  hipStream_t s = static_cast<hipStream_t>(stream);
#define NFX_EF(M, T)                                                                                            \
  launch_elasticity_faces<M, T>(n_faces, n_elem, n_s, n_u, d_face_inc, d_face_lf, d_el_face, d_el_orient, d_cells, \
                                d_coords, d_tables, d_lam, d_mu, coef_stride_e, coef_stride_q, d_f, f_stride_e,   \
                                f_stride_q, f_stride_c, d_rowptr, d_codes, d_vals, d_rhs, s)
  if (mode == 0) return fp64 ? NFX_EF(0, double) : NFX_EF(0, float);
  return fp64 ? NFX_EF(1, double) : NFX_EF(1, float);
#undef NFX_EF
}

extern "C" int nfx_gather(int fp64, int64_t n, const int32_t* d_idx, const void* d_src, void* d_dst,
                          void* stream) {
  NFX_REQUIRE(n >= 0, "nfx_gather: n < 0");
  if (n == 0) return NFX_OK;
  NFX_REQUIRE(d_idx && d_src && d_dst, "nfx_gather: null pointer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int blocks = static_cast<int>((n + 255) / 256);
  if (fp64)
    hipLaunchKernelGGL(k_gather<double>, dim3(blocks), dim3(256), 0, s, n, d_idx,
                       static_cast<const double*>(d_src), static_cast<double*>(d_dst));
  else
    hipLaunchKernelGGL(k_gather<float>, dim3(blocks), dim3(256), 0, s, n, d_idx,
                       static_cast<const float*>(d_src), static_cast<float*>(d_dst));
  NFX_LAUNCH_CHECK();
  return NFX_OK;
}
