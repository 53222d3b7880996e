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
#include "nfx_blockinv.h"
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
// Face-owner assembly: one wavefront owns the CSR rows of one facet (a face of a tetrahedron
// or an edge of a triangle): its BDM1 DoFs x stress rows, consecutive in the DoF numbering
// and therefore one contiguous CSR slice. It evaluates the rows of those DoFs in the element
// matrices of the facet's (at most two) incident elements, sums them in element order and
// writes the slice with coalesced stores. No slot codes per element entry, no staging, no
// atomics. Each CSR entry of a stress row carries a 16-bit source code: bits 0..3 the row
// inside the facet block (j * NC + c), bits 4..9 and 10..15 the element-local reference
// column (b * NC + d for stress, NSS + d for u_d, NSS + NC + d for t_d) of the entry in the
// first / second incident element, 63 = no contribution (entries dropped by the value
// filter, explicit zero diagonals).
//
// Per incident element the wave computes only the columns the upper triangle needs: the
// blocks of the element's facets whose DoF block index is >= the owning facet's, plus the
// u / t columns. The (s, s) contraction V = sum_q k2_q psi_ac(q) psi_bd(q) (rows x needed
// columns, K = quadrature points) runs on v_mfma_f32_16x16x4_f32 in FP32; FP64 builds
// evaluate the same tiles on the vector ALU. The diagonal-in-component part
// D_ab = sum_q k1_q psi_a . psi_b (and the Riesz map's div-div term) is a small block per
// element on the vector ALU.
//
// Cell rows (u_c, t_c of each element) hold one entry, their diagonal: 0 in A (the explicit
// zero of the SBAIJ scatter), the mass |K| in P; their residual rows take -sum_q w detJ f.
// One thread per cell, in blocks launched after the facet blocks.
//
// Table layout (R): w[NQ] | phi[NQ][NB][DIM] | div[NQ][NB] | Uref[NB] | Phi[NB][DIM] |
// DDref[NB][NB] with Uref_b = sum_q w_q divref_b(q), Phi_be = sum_q w_q phihat_be(q),
// DDref_ab = sum_q w_q divref_a(q) divref_b(q).
constexpr int kFW = 4;  // facets (waves) per block of the face kernel
constexpr int kCellBlock = 256;
constexpr int kNoSrc = 63;

// BDM1 - P0 - P0 sizes: NB basis functions, NQ quadrature points (degree 3), NC stress rows
// (= displacement components), NT rotation components, NBF basis functions per facet, NFC
// facets per element, NL local DoFs.
template <int DIM>
struct ElSizes;
template <>
struct ElSizes<3> {
  static constexpr int NB = 12, NQ = 8, NC = 3, NT = 3, NBF = 3, NFC = 4;
};
template <>
struct ElSizes<2> {
  static constexpr int NB = 6, NQ = 4, NC = 2, NT = 1, NBF = 2, NFC = 3;
};

typedef float f4_t __attribute__((ext_vector_type(4)));

template <int DIM, typename R>
struct FaceLds {
  using S = ElSizes<DIM>;
  static constexpr int ROWS = S::NBF * S::NC, NL = S::NB * S::NC + S::NC + S::NT;
  // psi staging passes of 64 lanes over (b, q): lanes past NB rows store to the junk slot
  static constexpr int NBP = (S::NB * S::NQ + 63) / 64 * 64 / S::NQ;
  // a K row holds NL columns and one junk slot (column NL) that masked lanes store to
  static constexpr int NLP = NL + 1, DJ = S::NBF * S::NB;
  // psi, kpsi and D serve one incidence at a time (the second reuses them after the first's
  // rows are in K)
  // s(b) psi_{b d}(q); rows padded by two (a stride of 26 / 10 words): the reads of the tiles
  // (rows b and b + 1 in adjacent lanes, q and q + 1 in the other half) hit distinct banks
  struct PsiRow {
    R v[DIM][S::NQ];
    R pad[2];
  } psi[S::NB];
  R kpsi[2][S::NBF][DIM][S::NQ];  // [0][j][c][q] = k2_q s(a_j) psi_{a_j c}(q); [1] with k1_q
  R D[DJ + 1];                    // s(a) s(b) D_{a_j b} at j NB + b (+ div-div, Riesz map); DJ junk
  R K[2][ROWS][NLP];              // signed element rows of the facet: s(a) s(b) K_e[(a, c), col]
};

// the element-independent table, staged in LDS once per block
template <int DIM>
struct TabSizes {
  using S = ElSizes<DIM>;
  static constexpr int W = 0, PHI = S::NQ, DIV = PHI + S::NQ * S::NB * DIM, U = DIV + S::NQ * S::NB,
                       PH = U + S::NB, DD = PH + S::NB * DIM, N = DD + S::NB * S::NB;
};

// Per-element geometry record (computed once per mesh): J / |det J| row-major, then |det J|.
template <int DIM, typename R>
__device__ __forceinline__ void face_geometry(const R* __restrict__ geom, int e, R (&Jn)[DIM][DIM], R& detJ) {
  const R* gp = geom + (int64_t)e * (DIM * DIM + 1);
#pragma unroll
  for (int c = 0; c < DIM; ++c)
#pragma unroll
    for (int d = 0; d < DIM; ++d) Jn[c][d] = gp[c * DIM + d];
  detJ = gp[DIM * DIM];
}

// Per incident element of a facet: element, Jacobian and |det J|, the facet slots (2 bits per
// reference basis function) and orientation signs (1 bit per function) of the element, the
// reference basis functions at the owning facet's slots j (4 bits each) and the list of the
// element's local facets whose block is >= the owning facet's (2 bits each, nb of them).
template <int DIM, typename R>
struct FaceInc {
  int e;
  R J[DIM][DIM];  // J / |det J|
  R detJ;
  uint32_t slots, signs;
  int ajpack, blpack, nb;
  R lam_q, mu_q;  // coefficients at quadrature point q = lane mod NQ
};

template <int DIM, typename R>
__device__ __forceinline__ R basis_sign(const FaceInc<DIM, R>& g, int b) {
  return ((g.signs >> b) & 1) ? R(-1) : R(1);
}

// Every global load of an incidence issued together (one latency wait per facet instead of a
// chain per phase): geometry record, orientation word, facet blocks, the lane's coefficients.
template <int DIM, typename R>
__device__ __forceinline__ FaceInc<DIM, R> face_inc_data(const R* __restrict__ geom,
                                                          const uint64_t* __restrict__ el_orient,
                                                          const int32_t* __restrict__ el_face,
                                                          const R* __restrict__ lam, const R* __restrict__ mu,
                                                          int cs_e, int cs_q, int e, int lf, int p, int l) {
  using S = ElSizes<DIM>;
  FaceInc<DIM, R> g;
  g.e = e;
  const int q = l & (S::NQ - 1);
  g.mu_q = mu[(int64_t)e * cs_e + q * cs_q];
  g.lam_q = lam[(int64_t)e * cs_e + q * cs_q];
  const uint64_t o = el_orient[e];
  int fb[S::NFC];
#pragma unroll
  for (int k = 0; k < S::NFC; ++k) fb[k] = el_face[e * S::NFC + k];
  face_geometry<DIM, R>(geom, e, g.J, g.detJ);
  g.slots = (uint32_t)o;
  g.signs = (uint32_t)(o >> 32);
  g.ajpack = 0;
#pragma unroll
  for (int k = 0; k < S::NBF; ++k) {
    const int b = S::NBF * lf + k;
    g.ajpack |= b << (4 * (int)((g.slots >> (2 * b)) & 3));
  }
  g.nb = 0;
  g.blpack = 0;
#pragma unroll
  for (int k = 0; k < S::NFC; ++k)
    if (fb[k] >= p) {
      g.blpack |= k << (2 * g.nb);
      ++g.nb;
    }
  return g;
}

// All element rows of the owning facet in incidence I (uniform): psi, coefficient-weighted
// rows, the D block, the u / t columns and the (s, s) blocks of the needed facets into L.K[I].
// The matrix-core tiles use power-of-two lane layouts, so no lane index needs a division:
// accumulator row 4 j + c (facet slot j < NBF, component c < NC), column 4 w + d (basis w of
// one needed facet, component d); one 16 x 16 tile per needed facet. Lanes without an entry
// to store write to the junk slots (D[DJ], column NL of K row 0) instead of branching.
template <int DIM, int MODE, int I, typename R>
__device__ __forceinline__ void face_incidence(FaceLds<DIM, R>& L, const R* __restrict__ T,
                                               const FaceInc<DIM, R>& g, int l) {
  using S = ElSizes<DIM>;
  using TS = TabSizes<DIM>;
  using LD = FaceLds<DIM, R>;
  constexpr int NB = S::NB, NQ = S::NQ, NC = S::NC, NT = S::NT, NBF = S::NBF, NFC = S::NFC;
  constexpr int ROWS = LD::ROWS, NSS = NB * NC, NLP = LD::NLP, NL = LD::NL;
  constexpr int LQ = NQ == 8 ? 3 : 2;  // log2 NQ
  static_assert(NBF <= 4 && NC <= 4 && NB < LD::NBP, "lane layout");
  R* const Kf = &L.K[I][0][0];
  constexpr int JUNK = NL;  // column NL of row 0: written by masked lanes, never read
  __builtin_amdgcn_wave_barrier();  // a previous incidence's reads of psi / kpsi / D are done
  // psi_{b d}(q) = (J phihat_b(q))_d / detJ, lane <-> (b, q); lanes past the NB rows read
  // the table at b = 0 and store to the junk slot
#pragma unroll
  for (int pass = 0; pass < LD::NBP * NQ / 64; ++pass) {
    const int idx = l + 64 * pass;
    const bool in = idx < NB * NQ;
    const int b = in ? idx >> LQ : 0, q = idx & (NQ - 1);
    R ph[DIM];
#pragma unroll
    for (int x = 0; x < DIM; ++x) ph[x] = T[TS::PHI + (q * NB + b) * DIM + x];
    // the orientation sign of b folded in: every product below is then signed
    const R sc = basis_sign(g, b);
    R* const dst = in ? &L.psi[b].v[0][q] : Kf + JUNK;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      R v = 0;
#pragma unroll
      for (int x = 0; x < DIM; ++x) v += g.J[d][x] * ph[x];
      dst[in ? d * NQ : 0] = v * sc;
    }
  }
  __builtin_amdgcn_wave_barrier();
  // k2_q psi_{a_j c}(q) and k1_q psi_{a_j c}(q), lane <-> (j, q)
  if (l < NBF * NQ) {
    const int j = l >> LQ, q = l & (NQ - 1);
    const int a = (g.ajpack >> (4 * j)) & 15;
    const R m = g.mu_q, lm = g.lam_q;
    const R wd = T[TS::W + q] * g.detJ;
    R k1, k2;
    if constexpr (sizeof(R) == 4) {
      // FP32 rows: hardware reciprocals (1 ulp), well inside the FP32 value tolerance
      const float r2m = __builtin_amdgcn_rcpf(2.f * m);
      k1 = wd * r2m;
      k2 = k1 * (lm * __builtin_amdgcn_rcpf(2.f * m + float(DIM) * lm));
    } else {
      k1 = wd / (R(2) * m);
      k2 = wd * (R(1) / (R(2) * m)) * (lm / (R(2) * m + R(DIM) * lm));
    }
#pragma unroll
    for (int c = 0; c < DIM; ++c) {
      const R ps = L.psi[a].v[c][q];
      L.kpsi[0][j][c][q] = k2 * ps;
      L.kpsi[1][j][c][q] = k1 * ps;
    }
  }
  __builtin_amdgcn_wave_barrier();
  const int lr = l & 15, lq = l >> 4;
  // D_{a_j b} = sum_{c, q} k1_q psi_{a_j c}(q) psi_{b c}(q) (+ DDref / detJ for the Riesz map):
  // an NBF x NB block, K = NC NQ; accumulator row j = r (lq = 0), column b = lr
  {
    R dv[4];
    if constexpr (sizeof(R) == 4) {
      f4_t acc = f4_t{0.f, 0.f, 0.f, 0.f};
      // rows j >= NBF and columns b >= NB of the tile read clamped entries: they only reach
      // accumulator rows / columns that are never stored
      const int jr = lr < NBF ? lr : 0, br = lr < NB ? lr : 0;
#pragma unroll
      for (int s = 0; s < NC * NQ / 4; ++s) {
        const int k = 4 * s + lq, c = k >> LQ, q = k & (NQ - 1);
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(L.kpsi[1][jr][c][q], L.psi[br].v[c][q], acc, 0, 0, 0);
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) dv[r] = acc[r];
    } else {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        R sacc = 0;
        if (r < NBF)
#pragma unroll
          for (int c = 0; c < NC; ++c)
#pragma unroll
            for (int q = 0; q < NQ; ++q) sacc += L.kpsi[1][r < NBF ? r : 0][c][q] * L.psi[lr < NB ? lr : 0].v[c][q];
        dv[r] = sacc;
      }
    }
    const bool dl = lq == 0 && lr < NB;
    const int lrc = lr < NB ? lr : 0;
#pragma unroll
    for (int r = 0; r < NBF; ++r) {
      R v = dv[r];
      if (MODE == 1) {
        const int a = (g.ajpack >> (4 * r)) & 15;
        v += basis_sign(g, a) * basis_sign(g, lrc) * T[TS::DD + a * NB + lrc] / g.detJ;
      }
      L.D[dl ? r * NB + lr : LD::DJ] = v;
    }
  }
  // u / t columns of the facet rows: lane 16 j + 4 c + x holds row (j, c), columns u_x and
  // t_x (x < NC, resp. x < NT)
  if (MODE == 0) {
    const int j = l >> 4, c = (l >> 2) & 3, x = l & 3;
    const bool ok = j < NBF && c < NC;
    const int jj = ok ? j : 0, cc = ok ? c : 0;
    const int a = (g.ajpack >> (4 * jj)) & 15;
    const R sa = basis_sign(g, a);
    const int rowo = (NC * jj + cc) * NLP;
    // displacement column u_x: sum_q w divref_a on the diagonal component
    Kf[ok && x < NC ? rowo + NSS + x : NL] = sa * (x == cc ? T[TS::U + a] : R(0));
    // rotation column t_x; in 2-D the single rotation is the one about z
    const int ct = DIM == 3 ? x : 2;
    const int k = 3 - cc - ct;  // < DIM when ct != c
    // A_{a k} = sum_y J_ky Phi_ay for every row k, then a select (no runtime index into J)
    R Aak = 0;
#pragma unroll
    for (int y = 0; y < DIM; ++y) {
      R ar = 0;
#pragma unroll
      for (int z = 0; z < DIM; ++z) ar += g.J[y][z] * T[TS::PH + a * DIM + z];
      Aak = k == y ? ar * g.detJ : Aak;  // J = (J / |det J|) |det J|
    }
    const bool even = (ct == 0 && cc == 1) || (ct == 1 && cc == 2) || (ct == 2 && cc == 0);
    const R vt = ct == cc ? R(0) : (even ? Aak : -Aak);
    Kf[ok && x < NT ? rowo + NSS + NC + x : NL] = sa * vt;
  }
  __builtin_amdgcn_wave_barrier();
  // (s, s) blocks: V = sum_q k2_q psi_ac psi_bd on the matrix cores, one tile per needed facet
  // (rows 4 j + c, columns 4 w + d), then K = delta_cd D - V with the orientation signs
  const int aj = lr >> 2, ac = lr & 3;   // A operand row (j, c)
  const int wj = lr >> 2, dd = lr & 3;   // accumulator column (w, d)
  const bool arow = aj < NBF && ac < NC;
  const bool ccol = wj < NBF && dd < DIM;
  const int ajc = arow ? aj : 0, acc_ = arow ? ac : 0;
  const int wjc = ccol ? wj : 0, ddc = ccol ? dd : 0;
  const bool vrow = lq < NBF;  // accumulator rows 4 lq + r, r < NC
  // row offsets of the three stored rows (junk when the lane holds no entry)
  int rowo[NC];
#pragma unroll
  for (int r = 0; r < NC; ++r) rowo[r] = (vrow && ccol) ? (NC * lq + r) * NLP : JUNK;
  // the A operand is the same for every tile (rows (j, c) past NBF x NC read a clamped entry:
  // their accumulator rows are never stored)
  R a_op[NQ / 4];
  if constexpr (sizeof(R) == 4) {
#pragma unroll
    for (int s = 0; s < NQ / 4; ++s) a_op[s] = L.kpsi[0][ajc][acc_][4 * s + lq];
  }
#pragma unroll
  for (int t = 0; t < NFC; ++t) {
    if (t >= g.nb) break;
    const int lf = (g.blpack >> (2 * t)) & 3;  // uniform
    const int b = NBF * lf + wjc;
    R v[4];
    if constexpr (sizeof(R) == 4) {
      f4_t accv = f4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int s = 0; s < NQ / 4; ++s) {
        const float bv = L.psi[b].v[ddc][4 * s + lq];
        accv = __builtin_amdgcn_mfma_f32_16x16x4f32(a_op[s], bv, accv, 0, 0, 0);
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) v[r] = accv[r];
    } else {
#pragma unroll
      for (int r = 0; r < NC; ++r) {
        R sacc = 0;
#pragma unroll
        for (int q = 0; q < NQ; ++q) sacc += L.kpsi[0][vrow ? lq : 0][r][q] * L.psi[b].v[ddc][q];
        v[r] = sacc;
      }
    }
    // K column of (b, d) inside the needed-facet blocks: NBF NC lf + NC w + d
    const int colk = NBF * NC * lf + NC * wjc + ddc;
    const R dlv = L.D[(vrow ? lq : 0) * NB + b];
#pragma unroll
    for (int r = 0; r < NC; ++r) {
      const int ro = rowo[r];
      Kf[ro == JUNK ? JUNK : ro + colk] = (r == ddc ? dlv : R(0)) - v[r];
    }
  }
}

template <int DIM, int MODE, typename R>
__global__ __launch_bounds__(64 * kFW, sizeof(R) == 4 ? 7 : 3) void k_elasticity_faces(
    int n_faces, int n_elem, const int32_t* __restrict__ face_inc, const int32_t* __restrict__ face_lf,
    const int32_t* __restrict__ el_face, const uint64_t* __restrict__ el_orient, const R* __restrict__ geom,
    const R* __restrict__ tab, const R* __restrict__ lam, const R* __restrict__ mu,
    int cs_e, int cs_q, const R* __restrict__ f, int fs_e, int fs_q, int fs_c, const int32_t* __restrict__ rowptr,
    const uint16_t* __restrict__ codes, R* __restrict__ vals, R* __restrict__ rhs) {
  using S = ElSizes<DIM>;
  using TS = TabSizes<DIM>;
  constexpr int NQ = S::NQ, NC = S::NC, NT = S::NT, NBF = S::NBF, NFC = S::NFC;
  constexpr int ROWS = FaceLds<DIM, R>::ROWS;
  const int n_s = ROWS * n_faces, n_u = NC * n_elem;
  const int n_face_blocks = (n_faces + kFW - 1) / kFW;
  if ((int)blockIdx.x >= n_face_blocks) {
    // cell rows: u_c = n_s + NC e + c, t_c = n_s + n_u + NT e + c
    const int e = ((int)blockIdx.x - n_face_blocks) * kCellBlock + threadIdx.x;
    if (e >= n_elem) return;
    R J[DIM][DIM], detJ;
    face_geometry<DIM, R>(geom, e, J, detJ);
    R mass = 0;
#pragma unroll
    for (int q = 0; q < NQ; ++q) mass += tab[TS::W + q] * detJ;
#pragma unroll
    for (int c = 0; c < NC + NT; ++c) {
      const int row = c < NC ? n_s + NC * e + c : n_s + n_u + NT * e + (c - NC);
      vals[rowptr[row]] = MODE == 1 ? mass : R(0);
      if (MODE == 0) {
        R sr = 0;
#pragma unroll
        for (int q = 0; q < NQ; ++q) sr -= tab[TS::W + q] * detJ * f[(int64_t)e * fs_e + q * fs_q + c * fs_c];
        rhs[row] = sr;
      }
    }
    return;
  }
  __shared__ R T[TS::N];
  __shared__ FaceLds<DIM, R> lds_all[kFW];
  for (int i = threadIdx.x; i < TS::N; i += 64 * kFW) T[i] = tab[i];
  __syncthreads();
  const int wv = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int p = __builtin_amdgcn_readfirstlane((int)blockIdx.x * kFW + wv);
  if (p >= n_faces) return;
  FaceLds<DIM, R>& L = lds_all[wv];
  const int e0 = face_inc[2 * p], e1r = face_inc[2 * p + 1];
  const int ni = e1r >= 0 ? 2 : 1;
  const int lfw = face_lf[p];
  // the facet's CSR slice; its first source codes are loaded while the rows are computed
  const int base = rowptr[ROWS * p];
  const int cnt = rowptr[ROWS * p + ROWS] - base;
  constexpr int kPre = 6;
  int pre[kPre];
#pragma unroll
  for (int u = 0; u < kPre; ++u) pre[u] = l + 64 * u < cnt ? codes[base + l + 64 * u] : 0;
  // both incidences' loads before any arithmetic (the second reuses the first element on a
  // boundary facet; its rows are then not used)
  const FaceInc<DIM, R> g0 = face_inc_data<DIM, R>(geom, el_orient, el_face, lam, mu, cs_e, cs_q, e0, lfw & 3, p, l);
  const FaceInc<DIM, R> g1 = face_inc_data<DIM, R>(geom, el_orient, el_face, lam, mu, cs_e, cs_q,
                                                   ni == 2 ? e1r : e0, (lfw >> 2) & 3, p, l);
  face_incidence<DIM, MODE, 0, R>(L, T, g0, l);
  if (ni == 2) face_incidence<DIM, MODE, 1, R>(L, T, g1, l);
  __builtin_amdgcn_wave_barrier();
  // the facet's CSR slice: contiguous, coalesced stores; a missing source (code 63) reads a
  // clamped slot and is selected away (0 + K0 + K1 in that order, a missing term as 0)
  constexpr int NLP = FaceLds<DIM, R>::NLP, NL = FaceLds<DIM, R>::NL;
  const R* const K0 = &L.K[0][0][0];
  const R* const K1 = &L.K[1][0][0];
  auto entry = [&](int code) -> R {
    const int r = code & 15, s0 = (code >> 4) & 63, s1 = (code >> 10) & 63;
    const R a0 = K0[r * NLP + (s0 < NL ? s0 : 0)];
    const R a1 = K1[r * NLP + (s1 < NL ? s1 : 0)];
    return (R(0) + (s0 < NL ? a0 : R(0))) + (s1 < NL ? a1 : R(0));
  };
#pragma unroll
  for (int u = 0; u < kPre; ++u) {
    const int k = l + 64 * u;
    if (k < cnt) vals[base + k] = entry(pre[u]);
  }
  for (int k = l + 64 * kPre; k < cnt; k += 64) vals[base + k] = entry(codes[base + k]);
  if (MODE == 0 && l < ROWS) rhs[ROWS * p + l] = R(0);
}

template <int DIM, int MODE, typename R>
static int launch_elasticity_faces(int n_faces, int n_elem, const int32_t* face_inc, const int32_t* face_lf,
                                   const int32_t* el_face, const uint64_t* el_orient, const void* geom,
                                   const void* tab, const void* lam, const void* mu, int cse,
                                   int csq, const void* f, int fse, int fsq, int fsc, const int32_t* rowptr,
                                   const uint16_t* codes, void* vals, void* rhs, hipStream_t s) {
  const int blocks = grid_for(n_faces, kFW) + grid_for(n_elem, kCellBlock);
  hipLaunchKernelGGL((k_elasticity_faces<DIM, MODE, R>), dim3(blocks), dim3(64 * kFW), 0, s, n_faces, n_elem,
                     face_inc, face_lf, el_face, el_orient, static_cast<const R*>(geom), static_cast<const R*>(tab), static_cast<const R*>(lam), static_cast<const R*>(mu), cse, csq,
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

extern "C" int nfx_assemble_elasticity_faces(int dim, int mode, int n_faces, int n_elem,
                                             const int32_t* d_face_inc, const int32_t* d_face_lf,
                                             const int32_t* d_el_face, const uint64_t* d_el_orient,
                                             const void* d_geom, const void* d_tables,
                                             const void* d_lam, const void* d_mu, int coef_stride_e,
                                             int coef_stride_q, const void* d_f, int f_stride_e, int f_stride_q,
                                             int f_stride_c, const int32_t* d_rowptr, const uint16_t* d_codes,
                                             int fp64, void* d_vals, void* d_rhs, void* stream) {
  NFX_REQUIRE(dim == 2 || dim == 3, "nfx_assemble_elasticity_faces: dim must be 2 or 3");
  NFX_REQUIRE(mode == 0 || mode == 1, "nfx_assemble_elasticity_faces: mode");
  NFX_REQUIRE(n_faces >= 0 && n_elem >= 0, "nfx_assemble_elasticity_faces: sizes");
  if (n_faces == 0 && n_elem == 0) return NFX_OK;
  NFX_REQUIRE(d_face_inc && d_face_lf && d_el_face && d_el_orient && d_geom && d_tables && d_lam &&
                  d_mu && d_rowptr && d_codes && d_vals,
              "nfx_assemble_elasticity_faces: null pointer");
  NFX_REQUIRE(mode == 1 || (d_f && d_rhs), "nfx_assemble_elasticity_faces: residual args");
  hipStream_t s = static_cast<hipStream_t>(stream);
#define NFX_EF(D, M, T)                                                                                         \
  launch_elasticity_faces<D, M, T>(n_faces, n_elem, d_face_inc, d_face_lf, d_el_face, d_el_orient, d_geom,      \
                                   d_tables, d_lam, d_mu, coef_stride_e, coef_stride_q, d_f, f_stride_e,          \
                                   f_stride_q, f_stride_c, d_rowptr, d_codes, d_vals, d_rhs, s)
  if (dim == 3) {
    if (mode == 0) return fp64 ? NFX_EF(3, 0, double) : NFX_EF(3, 0, float);
    return fp64 ? NFX_EF(3, 1, double) : NFX_EF(3, 1, float);
  }
  if (mode == 0) return fp64 ? NFX_EF(2, 0, double) : NFX_EF(2, 0, float);
  return fp64 ? NFX_EF(2, 1, double) : NFX_EF(2, 1, float);
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
