// Element assembly kernels for the mixed Darcy form (RT_k - P_{k-1}, k = 1, 2,
// triangles and tetrahedra), its Dirichlet facet term, the deterministic
// staged-contribution reduction and small fill / geometry kernels.
//
// Data layout (HBM):
//   cells   int32 (E, DIM+1)        vertex ids, element-major
//   coords  float (V, 3)            vertex coordinates
//   slot    int32 (NL*NL, E)        entry-major slot codes (see nfx_common.h),
//                                   reference-basis order a*NL + b; the facet
//                                   orientation (signed permutation) of the
//                                   reference's permute_and_transform is folded
//                                   into these codes on the host
//   rhsrow  int32 (NP, E)           pressure residual rows
//   tables  R: w[NQ] | phi[NQ][NB][DIM] | div[NQ][NB] | psi[NQ][NP]
// One thread assembles one element; the element matrix never leaves registers.
#include "nfx_common.h"

namespace nfx {

template <int DIM, typename R>
struct Geom {
  R J[3][DIM];
  R G[DIM][DIM];
  R detJ;
};

// Geometry from already loaded vertex coordinates (same arithmetic as load_geometry).
template <int DIM, typename R>
__device__ __forceinline__ void geometry_from(const R (&X)[DIM + 1][3], Geom<DIM, R>& g) {
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int d = 0; d < DIM; ++d) g.J[c][d] = X[d + 1][c] - X[0][c];
#pragma unroll
  for (int i = 0; i < DIM; ++i)
#pragma unroll
    for (int j = 0; j < DIM; ++j) {
      R s = 0;
#pragma unroll
      for (int c = 0; c < 3; ++c) s += g.J[c][i] * g.J[c][j];
      g.G[i][j] = s;
    }
  if constexpr (DIM == 3) {
    const R d = g.J[0][0] * (g.J[1][1] * g.J[2][2] - g.J[1][2] * g.J[2][1]) -
                g.J[0][1] * (g.J[1][0] * g.J[2][2] - g.J[1][2] * g.J[2][0]) +
                g.J[0][2] * (g.J[1][0] * g.J[2][1] - g.J[1][1] * g.J[2][0]);
    g.detJ = d < R(0) ? -d : d;
  } else {
    const R cx = g.J[1][0] * g.J[2][1] - g.J[2][0] * g.J[1][1];
    const R cy = g.J[2][0] * g.J[0][1] - g.J[0][0] * g.J[2][1];
    const R cz = g.J[0][0] * g.J[1][1] - g.J[1][0] * g.J[0][1];
    g.detJ = sqrt(cx * cx + cy * cy + cz * cz);
  }
}

// Per-element metric of the row kernels: gd[s] = G[d][d2] / detJ over the symmetric
// pairs d <= d2 (s in row-major upper order) and detJ. The geometry kernel stores the
// same values (geo_stride floats per element: gd, detJ, zero padding), so a row
// kernel reading them computes what one deriving them from the vertices computes.
template <int DIM>
constexpr int geo_stride() {
  return DIM == 3 ? 8 : 4;
}

template <int DIM, typename R>
__device__ __forceinline__ void scaled_metric(const Geom<DIM, R>& g, R (&gd)[DIM * (DIM + 1) / 2], R& dj) {
  const R inv = R(1) / g.detJ;
  int s = 0;
#pragma unroll
  for (int d = 0; d < DIM; ++d)
#pragma unroll
    for (int d2 = d; d2 < DIM; ++d2) gd[s++] = g.G[d][d2] * inv;
  dj = g.detJ;
}

template <int DIM, typename R>
__device__ __forceinline__ void load_geometry(const int32_t* __restrict__ cells,
                                              const R* __restrict__ coords,
                                              int e, Geom<DIM, R>& g,
                                              R (&X)[DIM + 1][3]) {
#pragma unroll
  for (int a = 0; a <= DIM; ++a) {
    const int v = cells[e * (DIM + 1) + a];
#pragma unroll
    for (int c = 0; c < 3; ++c) X[a][c] = coords[3 * v + c];
  }
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int d = 0; d < DIM; ++d) g.J[c][d] = X[d + 1][c] - X[0][c];
#pragma unroll
  for (int i = 0; i < DIM; ++i)
#pragma unroll
    for (int j = 0; j < DIM; ++j) {
      R s = 0;
#pragma unroll
      for (int c = 0; c < 3; ++c) s += g.J[c][i] * g.J[c][j];
      g.G[i][j] = s;
    }
  if constexpr (DIM == 3) {
    const R d = g.J[0][0] * (g.J[1][1] * g.J[2][2] - g.J[1][2] * g.J[2][1]) -
                g.J[0][1] * (g.J[1][0] * g.J[2][2] - g.J[1][2] * g.J[2][0]) +
                g.J[0][2] * (g.J[1][0] * g.J[2][1] - g.J[1][1] * g.J[2][0]);
    g.detJ = d < R(0) ? -d : d;
  } else {
    // |J0 x J1|: the QR determinant of a 3x2 Jacobian (src/geometry/mapping.py:23-36)
    const R cx = g.J[1][0] * g.J[2][1] - g.J[2][0] * g.J[1][1];
    const R cy = g.J[2][0] * g.J[0][1] - g.J[0][0] * g.J[2][1];
    const R cz = g.J[0][0] * g.J[1][1] - g.J[1][0] * g.J[0][1];
    g.detJ = sqrt(cx * cx + cy * cy + cz * cz);
  }
}

// Mixed Darcy element (LaplaceDualWeakForm.evaluate_form):
//   M[a][b] = sum_q w kinv phi_a^T G phi_b / detJ
//   B[k][b] = sum_q w div_b psi_k                 (detJ cancels)
//   local (a, NB+k) = -B[k][a], (NB+k, b) = +B[k][b], (p, p) = 0
//   rp[k]   = -sum_q w detJ f psi_k
template <int DIM, int NB, int NP, int NQ, typename R>
__global__ __launch_bounds__(256) void k_darcy(
    int n_elem, const int32_t* __restrict__ cells, const R* __restrict__ coords,
    const R* __restrict__ tab, const R* __restrict__ kinv, int ks_e, int ks_q,
    const R* __restrict__ f, int fs_e, int fs_q, const int32_t* __restrict__ slot,
    const int32_t* __restrict__ rhs_row, R* __restrict__ vals, R* __restrict__ stage,
    R* __restrict__ rhs) {
  constexpr int NL = NB + NP;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n_elem) return;
  const R* __restrict__ tw = tab;
  const R* __restrict__ tphi = tab + NQ;
  const R* __restrict__ tdiv = tphi + NQ * NB * DIM;
  const R* __restrict__ tpsi = tdiv + NQ * NB;

  Geom<DIM, R> g;
  R X[DIM + 1][3];
  load_geometry<DIM, R>(cells, coords, e, g, X);
  const R invdet = R(1) / g.detJ;

  R M[NB][NB];
  R B[NP][NB];
  R rp[NP];
#pragma unroll
  for (int a = 0; a < NB; ++a)
#pragma unroll
    for (int b = 0; b < NB; ++b) M[a][b] = 0;
#pragma unroll
  for (int k = 0; k < NP; ++k) {
    rp[k] = 0;
#pragma unroll
    for (int b = 0; b < NB; ++b) B[k][b] = 0;
  }

  for (int q = 0; q < NQ; ++q) {
    const R w = tw[q];
    const R ki = kinv[(int64_t)e * ks_e + q * ks_q];
    const R fq = f[(int64_t)e * fs_e + q * fs_q];
    const R cm = w * ki * invdet;
    R gphi[NB][DIM];
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
      for (int d = 0; d < DIM; ++d) {
        R s = 0;
#pragma unroll
        for (int d2 = 0; d2 < DIM; ++d2) s += g.G[d][d2] * tphi[(q * NB + b) * DIM + d2];
        gphi[b][d] = s * cm;
      }
#pragma unroll
    for (int a = 0; a < NB; ++a)
#pragma unroll
      for (int b = a; b < NB; ++b) {
        R s = 0;
#pragma unroll
        for (int d = 0; d < DIM; ++d) s += tphi[(q * NB + a) * DIM + d] * gphi[b][d];
        M[a][b] += s;
      }
#pragma unroll
    for (int k = 0; k < NP; ++k) {
      const R psi = tpsi[q * NP + k];
      const R wp = w * psi;
#pragma unroll
      for (int b = 0; b < NB; ++b) B[k][b] += wp * tdiv[q * NB + b];
      rp[k] -= wp * g.detJ * fq;
    }
  }

  const int64_t E = n_elem;
#pragma unroll
  for (int a = 0; a < NB; ++a) {
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const R v = a <= b ? M[a][b] : M[b][a];
      put_slot<R>(slot[(int64_t)(a * NL + b) * E + e], v, vals, stage);
    }
#pragma unroll
    for (int k = 0; k < NP; ++k) {
      put_slot<R>(slot[(int64_t)(a * NL + NB + k) * E + e], -B[k][a], vals, stage);
      put_slot<R>(slot[(int64_t)((NB + k) * NL + a) * E + e], B[k][a], vals, stage);
    }
  }
#pragma unroll
  for (int k = 0; k < NP; ++k) {
#pragma unroll
    for (int l = 0; l < NP; ++l)
      put_slot<R>(slot[(int64_t)((NB + k) * NL + NB + l) * E + e], R(0), vals, stage);
    put_row<R>(rhs_row[(int64_t)k * E + e], rp[k], rhs);
  }
}

// Dirichlet facet term of the mixed forms: for facet DoF a and data component c,
//   r[a, c] = scale * sum_q w |detJ_f| g_c(x_q) psi_a(x_q) . n
// (LaplaceDualWeakFormBCDirichlet: g = p_D, NC = 1, scale = +1;
//  LEDualWeakFormBCDirichlet: g = u_D, NC = 3, scale = -1).
// trace: (ncombo, NQF, FD, DIM) reference flux basis at the facet points mapped
// into the neighbour's reference cell, per (local facet, facet vertex order);
// g: (F, NQF, NC); rows: (F, FD * NC) row codes, entry a * NC + c.
// Weights and normal traces of a facet: wq[q] = w_q |det J_f|, s[q][a] = psi_a(x_q) . n
// (reference flux basis of facet DoF a traced from the neighbour with vertices X and
// geometry g, pushed forward, n the unit normal pointing out of the neighbour).
// fn: the facet's DIM vertex ids in stored order; cmb: its combo index into trace.
template <int DIM, int FD, int NQF, typename R>
__device__ __forceinline__ void facet_trace(const int32_t* __restrict__ fn, int cmb, const R (&X)[DIM + 1][3],
                                            const Geom<DIM, R>& g, const R* __restrict__ coords,
                                            const R* __restrict__ trace, const R* __restrict__ fw,
                                            R (&wq)[NQF], R (&s)[NQF][FD]) {
  // facet geometry in stored node order
  R Y[DIM][3];
#pragma unroll
  for (int a = 0; a < DIM; ++a) {
    const int v = fn[a];
#pragma unroll
    for (int c = 0; c < 3; ++c) Y[a][c] = coords[3 * v + c];
  }
  R xc0[3], xc1[3], out[3], n[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    R s0 = 0, s1 = 0;
#pragma unroll
    for (int a = 0; a <= DIM; ++a) s0 += X[a][c];
#pragma unroll
    for (int a = 0; a < DIM; ++a) s1 += Y[a][c];
    xc0[c] = s0 / R(DIM + 1);
    xc1[c] = s1 / R(DIM);
    out[c] = xc1[c] - xc0[c];
  }
  R detf;
  if constexpr (DIM == 2) {
    R v[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) v[c] = Y[1][c] - Y[0][c];
    detf = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    // n = v x (v x out) / |...| rearranged: in-plane normal of the edge
    R wv[3] = {v[1] * out[2] - v[2] * out[1], v[2] * out[0] - v[0] * out[2], v[0] * out[1] - v[1] * out[0]};
    n[0] = v[1] * wv[2] - v[2] * wv[1];
    n[1] = v[2] * wv[0] - v[0] * wv[2];
    n[2] = v[0] * wv[1] - v[1] * wv[0];
  } else {
    R a1[3], a2[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      a1[c] = Y[1][c] - Y[0][c];
      a2[c] = Y[2][c] - Y[0][c];
    }
    const R cx = a1[1] * a2[2] - a1[2] * a2[1];
    const R cy = a1[2] * a2[0] - a1[0] * a2[2];
    const R cz = a1[0] * a2[1] - a1[1] * a2[0];
    detf = sqrt(cx * cx + cy * cy + cz * cz);
    n[0] = cx;
    n[1] = cy;
    n[2] = cz;
  }
  {
    const R nn = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    const R sgn = (n[0] * out[0] + n[1] * out[1] + n[2] * out[2]) < R(0) ? R(-1) : R(1);
#pragma unroll
    for (int c = 0; c < 3; ++c) n[c] = sgn * n[c] / nn;
  }
  // J^T n / detJ (pushing the flux basis forward and dotting with n)
  R jtn[DIM];
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    R s = 0;
#pragma unroll
    for (int c = 0; c < 3; ++c) s += g.J[c][d] * n[c];
    jtn[d] = s / g.detJ;
  }
  const R* __restrict__ tr = trace + (int64_t)cmb * NQF * FD * DIM;
#pragma unroll
  for (int q = 0; q < NQF; ++q) {
    wq[q] = fw[q] * detf;
#pragma unroll
    for (int a = 0; a < FD; ++a) {
      R sv = 0;
#pragma unroll
      for (int d = 0; d < DIM; ++d) sv += tr[(q * FD + a) * DIM + d] * jtn[d];
      s[q][a] = sv;
    }
  }
}

// val[a][c] of facet f from the neighbour's vertices X and geometry g.
template <int DIM, int FD, int NQF, int NC, typename R>
__device__ __forceinline__ void facet_term(int f, const R (&X)[DIM + 1][3], const Geom<DIM, R>& g,
                                           const int32_t* __restrict__ facet_nodes,
                                           const int32_t* __restrict__ combo, const R* __restrict__ coords,
                                           const R* __restrict__ trace, const R* __restrict__ fw,
                                           const R* __restrict__ g_data, R (&val)[FD][NC]) {
  R wq[NQF], s[NQF][FD];
  facet_trace<DIM, FD, NQF, R>(facet_nodes + (int64_t)f * DIM, combo[f], X, g, coords, trace, fw, wq, s);
#pragma unroll
  for (int a = 0; a < FD; ++a)
#pragma unroll
    for (int c = 0; c < NC; ++c) val[a][c] = 0;
#pragma unroll
  for (int q = 0; q < NQF; ++q)
#pragma unroll
    for (int a = 0; a < FD; ++a)
#pragma unroll
      for (int c = 0; c < NC; ++c) val[a][c] += wq[q] * g_data[((int64_t)f * NQF + q) * NC + c] * s[q][a];
}

template <int DIM, int FD, int NQF, int NC, typename R>
__global__ __launch_bounds__(256) void k_facet_bc(
    int n_facets, const int32_t* __restrict__ facet_nodes, const int32_t* __restrict__ neigh,
    const int32_t* __restrict__ combo, const int32_t* __restrict__ cells,
    const R* __restrict__ coords, const R* __restrict__ trace, const R* __restrict__ fw,
    const R* __restrict__ g_data, const int32_t* __restrict__ rows, R scale, R* __restrict__ rhs) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= n_facets) return;
  const int e = neigh[f];
  Geom<DIM, R> g;
  R X[DIM + 1][3];
  load_geometry<DIM, R>(cells, coords, e, g, X);
  R val[FD][NC];
  facet_term<DIM, FD, NQF, NC, R>(f, X, g, facet_nodes, combo, coords, trace, fw, g_data, val);
#pragma unroll
  for (int a = 0; a < FD; ++a)
#pragma unroll
    for (int c = 0; c < NC; ++c)
      put_row<R>(rows[(int64_t)f * FD * NC + a * NC + c], scale * val[a][c], rhs);
}

// Row-owner ("gather") assembly of the mixed Darcy system.
//
// One thread owns one CSR row. Row r lists its incident (element, reference DoF)
// pairs in element order (at most two: a facet DoF is shared by the two cells of
// the facet, a cell DoF by one). For each incidence the thread evaluates row a of
// that element's matrix in the reference basis, applies the orientation signs and
// accumulates into the row's slice of an LDS buffer at the precomputed position of
// every column (first incidence then second: the reference's addv order). The
// block then stores its contiguous range of CSR values with coalesced writes.
//   inc[2][n_rows]  packed incidence e << 5 | a, -1 = none
//   pos             column position of reference DoF b inside the row, per (row, incidence):
//                   NL <= 6: one word, bits 4b..4b+3 the position of b (15 = none) and bit
//                   4 NL + b the sign of entry (a, b); otherwise NLP bytes (255 = none),
//                   NLP = NL, or NL rounded up to whole words when NL > 8
//   signs[E]        bit b set: reference DoF b of element e carries a minus sign;
//                   entry (a, b) enters with the sign (-1)^(bit a xor bit b)
// Element row a < NB: M[a][b] (b < NB) and -B[k][a] (b = NB + k);
// row a = NB + k: B[k][b] (b < NB), 0 (b >= NB), residual -sum_q w detJ f psi_k.
constexpr int kRowsPerBlock = 128;

// Bytes per incidence in the column-position array: the lowest-order elements (nl <= 6,
// rows of at most 12 entries) pack the nl positions as 4-bit fields of one 32-bit word and
// the nl entry signs of the incidence's element row above them (one word load per
// incidence, no gather of the element's sign bits); the larger elements keep one byte per
// position, padded to whole words so that a row reads them with word loads.
constexpr bool row_pos_packed(int nl) { return nl <= 6; }
// nl <= 5: bits 25..31 of an incidence's word are free; the two words of a row carry the
// row's CSR offset inside its 128-row tile (low 7 bits in word 0, high 7 in word 1), read
// by the single-system kernel instead of rowptr[r].
constexpr bool row_off_packed(int nl) { return nl <= 5; }
constexpr int row_pos_stride(int nl) { return nl <= 6 ? 4 : (nl > 8 ? (nl + 3) / 4 * 4 : nl); }

// Table layout: w[NQ] | phi[NQ][NB][DIM] | div[NQ][NB] | psi[NQ][NP] | Rs[NB][NB][NS] |
// Bc[NB][NP] | ck[NP]  (NS = DIM (DIM + 1) / 2 symmetric pairs of G = J^T J).
template <int DIM, int NB, int NP, int NQ>
struct DarcyTab {
  static constexpr int NS = DIM * (DIM + 1) / 2;
  static constexpr int NT = NQ * (1 + NB * DIM + NB + NP) + NB * NB * NS + NB * NP + NP;
};

// Facet DoFs and facet quadrature points of the pressure Dirichlet term, by element.
template <int DIM, int NB>
struct DarcyFacetShape;
template <>
struct DarcyFacetShape<2, 3> {
  static constexpr int FD = 1, NQF = 2;
};
template <>
struct DarcyFacetShape<2, 8> {
  static constexpr int FD = 2, NQF = 3;
};
template <>
struct DarcyFacetShape<3, 4> {
  static constexpr int FD = 1, NQF = 4;
};
template <>
struct DarcyFacetShape<3, 15> {
  static constexpr int FD = 3, NQF = 9;
};

// Pressure Dirichlet term folded into the row kernel. A row owned by a Dirichlet
// facet carries second incidence -2 - i, i = f * FD + a indexing the facet row codes
// (facet f, facet DoF a); its thread evaluates the facet term (facet_term, the same
// arithmetic as k_facet_bc) and stores it as the row's residual, which the volume
// term leaves at zero for flux rows. pD == nullptr: no Dirichlet data, every row
// stores its volume residual.
template <typename R>
struct DarcyRowBc {
  int n_codes;  // F * FD
  const int32_t* facet_nodes;
  const int32_t* combo;
  const R* trace;
  const R* fw;
  const R* pD;
  const int32_t* rows;
};

// One tile of TILE * RPT consecutive rows [r0, r0 + TILE * RPT) of one system: thread t
// owns rows r0 + h * TILE + t, h < RPT. st / sv are LDS buffers of at least
// DarcyTab::NT and TILE * RPT * 2 * NL values. Called by all threads of the block;
// contains block barriers. (Reading Rs / Bc / ck from global memory instead of LDS,
// with one barrier per tile, measured 6 % slower.)
// With per-element coefficients (ks_q == 0 and fs_q == 0) the quadrature sums are
// the precomputed reference tensors: flux row a of element e is
//   M_ab = kinv_e / detJ sum_s G_s Rs[a][b][s],  B part -Bc[a][k];
// pressure row k: Bc[b][k] and residual -detJ f_e ck[k]. Otherwise the sums run
// over the quadrature points. The loads of every incidence of all RPT rows are
// issued before any row is computed (RPT > 1: more loads in flight per wave). Rows
// owned by a Dirichlet facet take the facet term (DarcyRowBc) after the volume rows
// are written to LDS.
template <int DIM, typename R, bool GEO>
struct RowLoad {
  static constexpr int NS = DIM * (DIM + 1) / 2;
  int code[2], e[2], a[2];
  uint32_t sg[2];
  // GEO: the element metric comes precomputed (one load level less than cells ->
  // coordinates); otherwise the vertex coordinates
  R X[GEO ? 1 : 2][DIM + 1][3];
  R gd[2][NS], dj[2];
  R kv[2], fv[2];
  uint32_t pw[2];  // packed column positions (row_pos_packed)
};

// RS_ONLY: st holds only Rs | Bc | ck (the per-element path; tiles with per-quadrature
// coefficients are not processed). n_in_tile: rows of this tile (at most TILE * RPT).
template <int DIM, int NB, int NP, int NQ, int TILE, typename R, bool GEO = false, int RPT = 1,
          bool RS_ONLY = false, bool OFF_PACKED = false>
__device__ __forceinline__ void darcy_row_tile(
    int r0, int n_rows, const int32_t* __restrict__ inc, const uint8_t* __restrict__ pos,
    const int32_t* __restrict__ rowptr, const uint32_t* __restrict__ signs,
    const int32_t* __restrict__ cells, const R* __restrict__ coords, const R* __restrict__ tab,
    const R* __restrict__ kinv, int ks_e, int ks_q, const R* __restrict__ f, int fs_e, int fs_q,
    R* __restrict__ vals, R* __restrict__ rhs, R* st, R* sv, const DarcyRowBc<R>& bc,
    bool load_table = true, const R* __restrict__ geo = nullptr, int n_in_tile = TILE * RPT) {
  constexpr int NL = NB + NP;
  constexpr int NS = DarcyTab<DIM, NB, NP, NQ>::NS;
  constexpr int NT = DarcyTab<DIM, NB, NP, NQ>::NT;
  constexpr int NT_RS = NQ * (1 + NB * DIM + NB + NP);  // start of Rs | Bc | ck
  const int rl = min(r0 + n_in_tile, n_rows);
  const bool per_element = (ks_q == 0 && fs_q == 0);
  if (RS_ONLY && !per_element) return;
  // the reference-tensor path reads only Rs | Bc | ck; a caller that already holds the
  // table in st (same element, same path) passes load_table = false
  if (load_table) {
    if constexpr (RS_ONLY) {
      for (int i = threadIdx.x; i < NT - NT_RS; i += TILE) st[i] = tab[NT_RS + i];
    } else {
      for (int i = (per_element ? NT_RS : 0) + threadIdx.x; i < NT; i += TILE) st[i] = tab[i];
    }
  }
  const int base = rowptr[r0];
  const int span = rowptr[rl] - base;
  for (int i = threadIdx.x; i < span; i += TILE) sv[i] = R(0);
  __syncthreads();
  const R* tw = st;
  const R* tphi = st + NQ;
  const R* tdiv = tphi + NQ * NB * DIM;
  const R* tpsi = tdiv + NQ * NB;
  const R* tRs = RS_ONLY ? st : tpsi + NQ * NP;
  const R* tBc = tRs + NB * NB * NS;
  const R* tck = tBc + NB * NP;
  RowLoad<DIM, R, GEO> in[RPT];
  // load phase: incidences, then signs / metric (or vertices) / coefficients
#pragma unroll
  for (int h = 0; h < RPT; ++h) {
    const int r = r0 + h * TILE + threadIdx.x;
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int c = r < rl ? inc[(int64_t)j * n_rows + r] : -1;
      in[h].code[j] = c;
      in[h].e[j] = c < 0 ? 0 : c >> 5;
      in[h].a[j] = c & 31;
      if constexpr (row_pos_packed(NL))
        in[h].pw[j] = (OFF_PACKED ? r < rl : c >= 0) ? reinterpret_cast<const uint32_t*>(pos)[(int64_t)r * 2 + j] : 0u;
    }
  }
#pragma unroll
  for (int h = 0; h < RPT; ++h) {
    RowLoad<DIM, R, GEO>& L = in[h];
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      L.sg[j] = 0;
      L.kv[j] = L.fv[j] = R(0);
      if (L.code[j] >= 0) {
        if constexpr (row_pos_packed(NL)) {
          L.sg[j] = L.pw[j] >> (4 * NL);
        } else {
          const uint32_t s0 = signs[L.e[j]];
          L.sg[j] = s0 ^ (((s0 >> L.a[j]) & 1u) ? 0xffffffffu : 0u);
        }
        if constexpr (GEO && DIM == 3 && sizeof(R) == 4) {
          // two 16-byte loads (the per-element stride is 32 bytes, the array 256-byte aligned)
          const float4* gp = reinterpret_cast<const float4*>(geo + (int64_t)L.e[j] * geo_stride<DIM>());
          const float4 g0 = gp[0], g1 = gp[1];
          L.gd[j][0] = g0.x;
          L.gd[j][1] = g0.y;
          L.gd[j][2] = g0.z;
          L.gd[j][3] = g0.w;
          L.gd[j][4] = g1.x;
          L.gd[j][5] = g1.y;
          L.dj[j] = g1.z;
        } else if constexpr (GEO) {
          const R* gp = geo + (int64_t)L.e[j] * geo_stride<DIM>();
#pragma unroll
          for (int t = 0; t < NS; ++t) L.gd[j][t] = gp[t];
          L.dj[j] = gp[NS];
        } else {
          int v[DIM + 1];
#pragma unroll
          for (int c = 0; c <= DIM; ++c) v[c] = cells[L.e[j] * (DIM + 1) + c];
#pragma unroll
          for (int c = 0; c <= DIM; ++c)
#pragma unroll
            for (int d = 0; d < 3; ++d) L.X[j][c][d] = coords[3 * v[c] + d];
        }
        if (per_element) {
          L.kv[j] = kinv[(int64_t)L.e[j] * ks_e];
          L.fv[j] = f[(int64_t)L.e[j] * fs_e];
        }
      }
    }
  }
  // compute phase
#pragma unroll
  for (int h = 0; h < RPT; ++h) {
    const int r = r0 + h * TILE + threadIdx.x;
    if (r >= rl) continue;
    RowLoad<DIM, R, GEO>& L = in[h];
    int off;
    if constexpr (OFF_PACKED) {
      static_assert(row_off_packed(NL) && TILE == 128 && RPT == 1, "packed offsets: 128-row tiles");
      off = (int)((L.pw[0] >> 25) | ((L.pw[1] >> 25) << 7));
    } else {
      off = rowptr[r] - base;
    }
    R res = 0;
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      if (L.code[j] < 0) continue;
      if constexpr (!GEO) {
        Geom<DIM, R> g;
        geometry_from<DIM, R>(L.X[j], g);
        scaled_metric<DIM, R>(g, L.gd[j], L.dj[j]);
      }
      const int aj = L.a[j], ej = L.e[j];
      R row[NL];
#pragma unroll
      for (int b = 0; b < NL; ++b) row[b] = 0;
      if (RS_ONLY || per_element) {
        if (aj < NB) {
          R gs[NS];
#pragma unroll
          for (int t = 0; t < NS; ++t) gs[t] = L.gd[j][t] * L.kv[j];
          const R* rs = tRs + aj * NB * NS;
#pragma unroll
          for (int b = 0; b < NB; ++b) {
            R m = 0;
#pragma unroll
            for (int t = 0; t < NS; ++t) m += gs[t] * rs[b * NS + t];
            row[b] = m;
          }
#pragma unroll
          for (int k = 0; k < NP; ++k) row[NB + k] = -tBc[aj * NP + k];
        } else {
          const int k = aj - NB;
#pragma unroll
          for (int b = 0; b < NB; ++b) row[b] = tBc[b * NP + k];
          res -= L.dj[j] * L.fv[j] * tck[k];
        }
      } else if (!RS_ONLY && aj < NB) {
        // symmetric G / detJ from the packed upper triangle
        R Gd[DIM][DIM];
        {
          int t = 0;
#pragma unroll
          for (int d = 0; d < DIM; ++d)
#pragma unroll
            for (int d2 = d; d2 < DIM; ++d2) {
              Gd[d][d2] = L.gd[j][t];
              Gd[d2][d] = L.gd[j][t];
              ++t;
            }
        }
        for (int q = 0; q < NQ; ++q) {
          const R w = tw[q];
          const R cm = w * kinv[(int64_t)ej * ks_e + q * ks_q];
          R ga[DIM];
#pragma unroll
          for (int d = 0; d < DIM; ++d) {
            R s = 0;
#pragma unroll
            for (int d2 = 0; d2 < DIM; ++d2) s += Gd[d][d2] * tphi[(q * NB + aj) * DIM + d2];
            ga[d] = s * cm;
          }
#pragma unroll
          for (int b = 0; b < NB; ++b) {
            R s = 0;
#pragma unroll
            for (int d = 0; d < DIM; ++d) s += ga[d] * tphi[(q * NB + b) * DIM + d];
            row[b] += s;
          }
          const R wd = w * tdiv[q * NB + aj];
#pragma unroll
          for (int k = 0; k < NP; ++k) row[NB + k] -= wd * tpsi[q * NP + k];
        }
      } else if (!RS_ONLY) {
        const int k = aj - NB;
        for (int q = 0; q < NQ; ++q) {
          const R wp = tw[q] * tpsi[q * NP + k];
#pragma unroll
          for (int b = 0; b < NB; ++b) row[b] += wp * tdiv[q * NB + b];
          res -= wp * L.dj[j] * f[(int64_t)ej * fs_e + q * fs_q];
        }
      }
      if constexpr (row_pos_packed(NL)) {
#pragma unroll
        for (int b = 0; b < NL; ++b) {
          const int p = (L.pw[j] >> (4 * b)) & 15;
          if (p == 15) continue;
          sv[off + p] += ((L.sg[j] >> b) & 1u) ? -row[b] : row[b];
        }
      } else if constexpr (row_pos_stride(NL) % 4 == 0) {
        // padded to whole words: NLP / 4 word loads instead of NL byte loads
        constexpr int NW = row_pos_stride(NL) / 4;
        const uint32_t* pw = reinterpret_cast<const uint32_t*>(pos) + ((int64_t)r * 2 + j) * NW;
        uint32_t wd[NW];
#pragma unroll
        for (int w = 0; w < NW; ++w) wd[w] = pw[w];
#pragma unroll
        for (int b = 0; b < NL; ++b) {
          const int p = (wd[b >> 2] >> (8 * (b & 3))) & 0xff;
          if (p == 255) continue;
          sv[off + p] += ((L.sg[j] >> b) & 1u) ? -row[b] : row[b];
        }
      } else {
        const uint8_t* pr = pos + ((int64_t)r * 2 + j) * NL;
#pragma unroll
        for (int b = 0; b < NL; ++b) {
          const int p = pr[b];
          if (p == 255) continue;
          sv[off + p] += ((L.sg[j] >> b) & 1u) ? -row[b] : row[b];
        }
      }
    }
    if (bc.pD != nullptr && L.code[1] <= -2 && -L.code[1] - 2 < bc.n_codes) {
      constexpr int FD = DarcyFacetShape<DIM, NB>::FD, NQF = DarcyFacetShape<DIM, NB>::NQF;
      const int i = -L.code[1] - 2, fct = i / FD, a = i - fct * FD;
      Geom<DIM, R> g;
      R X0[DIM + 1][3];
      load_geometry<DIM, R>(cells, coords, L.e[0], g, X0);
      R val[FD][1];
      facet_term<DIM, FD, NQF, 1, R>(fct, X0, g, bc.facet_nodes, bc.combo, coords, bc.trace, bc.fw, bc.pD,
                                     val);
      R v = val[0][0];
#pragma unroll
      for (int t = 1; t < FD; ++t) v = a == t ? val[t][0] : v;
      res = (bc.rows[i] & kSlotNeg) ? -v : v;
    }
    rhs[r] = res;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < span; i += TILE) vals[base + i] = sv[i];
  __syncthreads();
}

// Rows per thread and waves per SIMD of the row kernel. The FP32 lowest-order kernel
// runs one row per thread at the hardware maximum of 8 waves (64 VGPRs). At 32^3 x 6
// (595,968 rows = 9,312 waves, more than the 8,192 wave slots) two rows per thread would
// fit one round at 5 waves per SIMD (87 VGPRs), but measured 18.7 us against 14.9 us:
// the occupancy hides the load latency better than the second row's loads do. The
// larger and FP64 elements keep the compiler's register choice.
template <int NB, typename R>
constexpr int row_kernel_rpt() {
  return 1;
}
template <int NB, typename R>
constexpr int row_kernel_min_waves() {
  return (NB <= 4 && sizeof(R) == 4) ? 8 : 1;
}

template <int DIM, int NB, int NP, int NQ, typename R, bool GEO>
__global__ __launch_bounds__(kRowsPerBlock)
__attribute__((amdgpu_waves_per_eu(row_kernel_min_waves<NB, R>(), 8))) void k_darcy_rows(
    int n_rows, const int32_t* __restrict__ inc, const uint8_t* __restrict__ pos,
    const int32_t* __restrict__ rowptr, const uint32_t* __restrict__ signs,
    const int32_t* __restrict__ cells, const R* __restrict__ coords, const R* __restrict__ tab,
    const R* __restrict__ kinv, int ks_e, int ks_q, const R* __restrict__ f, int fs_e, int fs_q,
    R* __restrict__ vals, R* __restrict__ rhs, DarcyRowBc<R> bc, const R* __restrict__ geo,
    const int32_t* __restrict__ tile_map, int n_map) {
  constexpr int NL = NB + NP;
  constexpr int NT = DarcyTab<DIM, NB, NP, NQ>::NT;
  constexpr int RPT = row_kernel_rpt<NB, R>();
  __shared__ R st[NT];
  __shared__ R sv[kRowsPerBlock * RPT * 2 * NL];
  // Slot i of the launch holds tile tile_map[i] (tile i without a map; entries outside
  // [0, n_tiles) are empty slots). Block b takes slots b, b + grid, ... (one slot per
  // block for the lowest-order elements; the larger elements run a persistent grid that
  // loads the reference table into LDS once). Workgroups are dispatched to the 8 XCDs
  // round-robin, so a map whose slots congruent to x mod 8 hold the tiles of one element
  // range keeps each element's data in one XCD's L2.
  const int n_tiles = (n_rows + kRowsPerBlock * RPT - 1) / (kRowsPerBlock * RPT);
  if constexpr (NB <= 4) {
    const int t = tile_map ? tile_map[blockIdx.x] : (int)blockIdx.x;
    if (t < 0 || t >= n_tiles) return;
    darcy_row_tile<DIM, NB, NP, NQ, kRowsPerBlock, R, GEO, RPT, false, row_off_packed(NB + NP) && RPT == 1>(
        t * kRowsPerBlock * RPT, n_rows, inc, pos, rowptr, signs, cells, coords, tab, kinv, ks_e, ks_q, f, fs_e,
        fs_q, vals, rhs, st, sv, bc, true, geo);
  } else {
    const int n_slots = tile_map ? n_map : n_tiles;
    bool first = true;
    for (int i = blockIdx.x; i < n_slots; i += gridDim.x) {
      const int t = tile_map ? tile_map[i] : i;
      if (t < 0 || t >= n_tiles) continue;
      darcy_row_tile<DIM, NB, NP, NQ, kRowsPerBlock, R, GEO, RPT>(t * kRowsPerBlock * RPT, n_rows, inc, pos,
                                                                  rowptr, signs, cells, coords, tab, kinv, ks_e,
                                                                  ks_q, f, fs_e, fs_q, vals, rhs, st, sv, bc,
                                                                  first, geo);
      first = false;
    }
  }
}

// ---------------------------------------------------------------------------
// Batched persistent assembly: many independent 3-D RT_k - P_{k-1} systems.
// The arrays are device memory: the pointers carry the global address space, so that
// the kernel reading them from the descriptor uses global loads and stores (through a
// generic pointer the compiler emits flat ones, which also occupy the LDS counter).
#define NFX_GLOBAL __attribute__((address_space(1)))
struct DarcyBatchItem {
  int k;
  int n_rows;
  int ks_e, ks_q, fs_e, fs_q;
  const NFX_GLOBAL int32_t* inc;
  const NFX_GLOBAL uint8_t* pos;
  const NFX_GLOBAL int32_t* rowptr;
  const NFX_GLOBAL uint32_t* signs;
  const NFX_GLOBAL int32_t* cells;
  const NFX_GLOBAL float* coords;
  const NFX_GLOBAL float* tab;
  const NFX_GLOBAL float* kinv;
  const NFX_GLOBAL float* f;
  NFX_GLOBAL float* vals;
  NFX_GLOBAL float* rhs;
  const NFX_GLOBAL float* geo;  // per-element metric (nfx_element_geometry), required
};
#undef NFX_GLOBAL

// tiles: (item, first row) pairs, heaviest first; block b takes tiles b, b + G, ...
// KD = 0: tiles of both degrees; KD = 1 or 2: every tile is of that degree (a tile of
// the other degree is skipped), so the kernel carries only that element's registers
// and LDS (RT1: 8 waves per SIMD, as the single-system kernel). KD = 3: RT2 tiles with
// per-element coefficients, LDS holding only Rs | Bc | ck and a staging slice of
// kRt2SpanPerRow entries per tile row; the tile's row count rides in bits 16.. of its item
// field (0 = TILE) and the host cuts tiles so that each CSR span fits, which brings a
// 256-row block to 39.4 KB of LDS: 4 blocks, 4 waves per SIMD.
constexpr int kRt2SpanPerRow = 33;
template <int KD>
constexpr int batched_min_waves() {
  return KD == 1 ? 8 : (KD == 3 ? 4 : 1);
}

template <int TILE, int KD>
__global__ __launch_bounds__(TILE)
__attribute__((amdgpu_waves_per_eu(batched_min_waves<KD>(), 8))) void k_darcy_rows_batched(
    const DarcyBatchItem* __restrict__ items, const int2* __restrict__ tiles, int n_tiles) {
  constexpr int NT2 = DarcyTab<3, 15, 4, 27>::NT, NT2_RS = 27 * (1 + 15 * 3 + 15 + 4);
  constexpr int NTS = KD == 1 ? DarcyTab<3, 4, 1, 8>::NT : (KD == 3 ? NT2 - NT2_RS : NT2);
  constexpr int NSV = KD == 1 ? TILE * 2 * 5 : (KD == 3 ? TILE * kRt2SpanPerRow : TILE * 2 * 19);
  __shared__ float st[NTS];
  __shared__ float sv[NSV];
  const DarcyRowBc<float> no_bc{0, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  // The table stays in LDS across consecutive tiles with the same table pointer and
  // path (the host gives all instances of one degree one table; tiles come grouped by
  // cost, hence by degree).
  const float* held_tab = nullptr;
  int held_path = -1;
  for (int t = blockIdx.x; t < n_tiles; t += gridDim.x) {
    const int2 tl = tiles[t];
    const int item = KD == 3 ? (tl.x & 0xFFFF) : tl.x;
    const int tile_n = KD == 3 && (tl.x >> 16) > 0 ? min(tl.x >> 16, TILE) : TILE;
    const DarcyBatchItem& it = items[item];
    // explicit casts from the global-typed fields: the compiler then knows every access
    // below is a global one
    const int32_t* inc = (const int32_t*)it.inc;
    const uint8_t* pos = (const uint8_t*)it.pos;
    const int32_t* rowptr = (const int32_t*)it.rowptr;
    const uint32_t* signs = (const uint32_t*)it.signs;
    const int32_t* cells = (const int32_t*)it.cells;
    const float* coords = (const float*)it.coords;
    const float* tab = (const float*)it.tab;
    const float* kinv = (const float*)it.kinv;
    const float* f = (const float*)it.f;
    float* vals = (float*)it.vals;
    float* rhs = (float*)it.rhs;
    const float* geo = (const float*)it.geo;
    const int path = (it.ks_q == 0 && it.fs_q == 0) ? 1 : 0;
    const bool load = tab != held_tab || path != held_path;
    held_tab = tab;
    held_path = path;
    if constexpr (KD != 2) {
      if (it.k == 1)
        darcy_row_tile<3, 4, 1, 8, TILE, float, true>(tl.y, it.n_rows, inc, pos, rowptr, signs, cells, coords,
                                                      tab, kinv, it.ks_e, it.ks_q, f, it.fs_e, it.fs_q, vals, rhs,
                                                      st, sv, no_bc, load, geo);
    }
    if constexpr (KD == 0 || KD == 2) {
      if (it.k == 2)
        darcy_row_tile<3, 15, 4, 27, TILE, float, true>(tl.y, it.n_rows, inc, pos, rowptr, signs, cells,
                                                        coords, tab, kinv, it.ks_e, it.ks_q, f, it.fs_e, it.fs_q,
                                                        vals, rhs, st, sv, no_bc, load, geo);
    }
    if constexpr (KD == 3) {
      // a tile whose CSR span exceeds the staging slice is not processed (the host's
      // tiling rules it out)
      const int rl = min(tl.y + tile_n, it.n_rows);
      if (it.k != 2 || path != 1 || rowptr[rl] - rowptr[tl.y] > NSV) {
        held_tab = nullptr;  // nothing loaded: the next tile reloads the table
        continue;
      }
      darcy_row_tile<3, 15, 4, 27, TILE, float, true, 1, true>(
            tl.y, it.n_rows, inc, pos, rowptr, signs, cells, coords, tab, kinv, it.ks_e, it.ks_q, f, it.fs_e,
            it.fs_q, vals, rhs, st, sv, no_bc, load, geo, tile_n);
    }
  }
}

template <int DIM, int NB, int NP, int NQ, typename R>
static int launch_darcy_rows(int n_rows, const int32_t* inc, const uint8_t* pos,
                             const int32_t* rowptr, const uint32_t* signs, const int32_t* cells,
                             const void* coords, const void* tab, const void* kinv, int kse,
                             int ksq, const void* f, int fse, int fsq, void* vals, void* rhs,
                             hipStream_t s, int n_bc = 0, const int32_t* facet = nullptr,
                             const int32_t* combo = nullptr, const void* trace = nullptr,
                             const void* fw = nullptr, const void* pD = nullptr,
                             const int32_t* bc_rows = nullptr, const void* geo = nullptr,
                             const int32_t* tile_map = nullptr, int n_map = 0) {
  const DarcyRowBc<R> bc{n_bc, facet, combo, static_cast<const R*>(trace), static_cast<const R*>(fw),
                         static_cast<const R*>(pD), bc_rows};
  // one slot per block for the lowest-order elements; a persistent grid (8 blocks per
  // CU of 256, a multiple of the 8 XCDs) for the larger tables
  const int n_slots = tile_map ? n_map : grid_for(n_rows, kRowsPerBlock * row_kernel_rpt<NB, R>());
  const int grid = NB <= 4 ? n_slots : std::min(n_slots, 2048);
  if (grid == 0) return NFX_OK;
  if (geo)
    hipLaunchKernelGGL((k_darcy_rows<DIM, NB, NP, NQ, R, true>), dim3(grid), dim3(kRowsPerBlock), 0, s, n_rows, inc,
                       pos, rowptr, signs, cells, static_cast<const R*>(coords), static_cast<const R*>(tab),
                       static_cast<const R*>(kinv), kse, ksq, static_cast<const R*>(f), fse, fsq,
                       static_cast<R*>(vals), static_cast<R*>(rhs), bc, static_cast<const R*>(geo), tile_map,
                       n_map);
  else
    hipLaunchKernelGGL((k_darcy_rows<DIM, NB, NP, NQ, R, false>), dim3(grid), dim3(kRowsPerBlock), 0, s, n_rows,
                       inc, pos, rowptr, signs, cells, static_cast<const R*>(coords), static_cast<const R*>(tab),
                       static_cast<const R*>(kinv), kse, ksq, static_cast<const R*>(f), fse, fsq,
                       static_cast<R*>(vals), static_cast<R*>(rhs), bc, static_cast<const R*>(nullptr), tile_map,
                       n_map);
  NFX_LAUNCH_CHECK();
  return NFX_OK;
}

// vals[target[m]] = sum of stage[segptr[m] .. segptr[m+1]) in order.
template <typename R>
__global__ __launch_bounds__(256) void k_stage_reduce(int n_seg, const int32_t* __restrict__ segptr,
                                                      const int32_t* __restrict__ target,
                                                      const R* __restrict__ stage,
                                                      R* __restrict__ vals) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= n_seg) return;
  const int b = segptr[m], en = segptr[m + 1];
  R s = stage[b];
  for (int r = b + 1; r < en; ++r) s += stage[r];
  vals[target[m]] = s;
}

template <typename R>
__global__ __launch_bounds__(256) void k_fill(int64_t n, R v, R* __restrict__ x) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t j = i; j < n; j += stride) x[j] = v;
}

__global__ __launch_bounds__(256) void k_qpoints(int dim, int n_elem, int nq,
                                                 const int32_t* __restrict__ cells,
                                                 const double* __restrict__ coords,
                                                 const double* __restrict__ bary,
                                                 double* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)n_elem * nq) return;
  const int e = (int)(t / nq), q = (int)(t % nq);
  double x[3] = {0, 0, 0};
  for (int a = 0; a <= dim; ++a) {
    const int v = cells[(int64_t)e * (dim + 1) + a];
    const double s = bary[q * (dim + 1) + a];
    for (int c = 0; c < 3; ++c) x[c] += s * coords[3 * v + c];
  }
  for (int c = 0; c < 3; ++c) out[t * 3 + c] = x[c];
}

template <int DIM, typename R>
__global__ __launch_bounds__(256) void k_element_geometry(int n_elem, const int32_t* __restrict__ cells,
                                                          const R* __restrict__ coords, R* __restrict__ geo) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n_elem) return;
  R X[DIM + 1][3];
#pragma unroll
  for (int a = 0; a <= DIM; ++a) {
    const int v = cells[(int64_t)e * (DIM + 1) + a];
#pragma unroll
    for (int c = 0; c < 3; ++c) X[a][c] = coords[3 * (int64_t)v + c];
  }
  Geom<DIM, R> g;
  geometry_from<DIM, R>(X, g);  // the function the vertex path of the row kernels uses
  constexpr int NS = DIM * (DIM + 1) / 2;
  R gd[NS], dj;
  scaled_metric<DIM, R>(g, gd, dj);
  R* o = geo + (int64_t)e * geo_stride<DIM>();
#pragma unroll
  for (int t = 0; t < NS; ++t) o[t] = gd[t];
  o[NS] = dj;
#pragma unroll
  for (int t = NS + 1; t < geo_stride<DIM>(); ++t) o[t] = R(0);
}

template <int DIM, int NB, int NP, int NQ, typename R>
static int launch_darcy(int n, const int32_t* cells, const void* coords, const void* tab,
                        const void* kinv, int kse, int ksq, const void* f, int fse, int fsq,
                        const int32_t* slot, const int32_t* rhs_row, void* vals, void* stage,
                        void* rhs, hipStream_t s) {
  hipLaunchKernelGGL((k_darcy<DIM, NB, NP, NQ, R>), dim3(grid_for(n, 256)), dim3(256), 0, s, n,
                     cells, static_cast<const R*>(coords), static_cast<const R*>(tab),
                     static_cast<const R*>(kinv), kse, ksq,
                     static_cast<const R*>(f), fse, fsq, slot,
                     rhs_row, static_cast<R*>(vals), static_cast<R*>(stage), static_cast<R*>(rhs));
  NFX_LAUNCH_CHECK();
  return NFX_OK;
}

template <int DIM, int FD, int NQF, typename R, int NC = 1>
static int launch_darcy_bc(int n, const int32_t* facet, const int32_t* neigh, const int32_t* combo,
                           const int32_t* cells, const void* coords, const void* trace,
                           const void* fw, const void* pD, const int32_t* rows, void* rhs,
                           hipStream_t s, double scale = 1.0) {
  hipLaunchKernelGGL((k_facet_bc<DIM, FD, NQF, NC, R>), dim3(grid_for(n, 256)), dim3(256), 0, s, n,
                     facet, neigh, combo, cells, static_cast<const R*>(coords),
                     static_cast<const R*>(trace), static_cast<const R*>(fw),
                     static_cast<const R*>(pD), rows, static_cast<R>(scale), static_cast<R*>(rhs));
  NFX_LAUNCH_CHECK();
  return NFX_OK;
}

// ---------------------------------------------------------------------------
// Mixed-dimensional coupling (CouplingWeakForm.evaluate_form,
// examples/md_elliptic_scalar/CouplingWeakForm.py:11-151) for RT1 facet DoFs: fracture
// triangle i lies on matrix face F with the element of side s (0 = positive, 1 = negative
// side of the fracture plane) and that element's facet DoF; with the normal trace
// s_q = phi . n_s (n_s out of that element) and 1/alpha = delta / (2 kappa_n):
//   (u_s, u_s) += (1/alpha) sum_q w |det J_F| s_q^2
//   (u_s, p) and (p, u_s) += sum_q w |det J_F| s_q          (P0 fracture pressure)
// each signed by the DoF's orientation (row code bit 30). pos (m, 6): CSR positions of
// (u0,u0), (u0,p), (p,u0), (u1,u1), (u1,p), (p,u1); every position belongs to one
// triangle (the split gives each side its own DoF), so += has no race. The (p, u_s)
// entries are multiplied by psign: +1 is the reference form (the fracture mass row adds
// the matrix outflow with the sign of the fracture divergence), -1 the conservative
// exchange (the outflow enters the fracture as a source).
template <typename R>
__global__ __launch_bounds__(256) void k_md_coupling(
    int m, const int32_t* __restrict__ side_elem, const int32_t* __restrict__ side_combo,
    const int32_t* __restrict__ side_code, const int32_t* __restrict__ facet_nodes,
    const int32_t* __restrict__ cells, const R* __restrict__ coords, const R* __restrict__ trace,
    const R* __restrict__ fw, const R* __restrict__ inv_alpha, const int32_t* __restrict__ pos, R psign,
    R* __restrict__ vals) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const R ia = inv_alpha[i];
#pragma unroll
  for (int sd = 0; sd < 2; ++sd) {
    Geom<3, R> g;
    R X[4][3];
    load_geometry<3, R>(cells, coords, side_elem[2 * i + sd], g, X);
    R wq[4], sq[4][1];
    facet_trace<3, 1, 4, R>(facet_nodes + 3 * (int64_t)i, side_combo[2 * i + sd], X, g, coords, trace, fw, wq,
                            sq);
    R s1 = 0, s2 = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      s1 += wq[q] * sq[q][0];
      s2 += wq[q] * sq[q][0] * sq[q][0];
    }
    const R sg = (side_code[2 * i + sd] & kSlotNeg) ? R(-1) : R(1);
    const int32_t* p = pos + 6 * (int64_t)i + 3 * sd;
    vals[p[0]] += ia * s2;
    vals[p[1]] += sg * s1;
    vals[p[2]] += psign * sg * s1;
  }
}

// dst[map[i]] = src[i]
template <typename R>
__global__ __launch_bounds__(256) void k_md_scatter(int n, const R* __restrict__ src, const int32_t* __restrict__ map,
                                                    R* __restrict__ dst) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[map[i]] = src[i];
}

// vals[pos[i]] = v[i] (mode 0) or += v[i] (mode 1); positions distinct
template <typename R>
__global__ __launch_bounds__(256) void k_md_update(int n, const int32_t* __restrict__ pos, const double* __restrict__ v,
                                                   int mode, R* __restrict__ vals) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const R x = (R)v[i];
  vals[pos[i]] = mode ? vals[pos[i]] + x : x;
}

// Backward-Euler right-hand side: b[i] = b0[i] on the flux rows (i < n_u) and
// b0[i] + mass[i - n_u] x[i] on the pressure rows (one thread per row).
__global__ __launch_bounds__(256) void k_md_rhs(int n, int n_u, const double* __restrict__ b0,
                                                const double* __restrict__ mass, const double* __restrict__ x,
                                                double* __restrict__ b) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  b[i] = i < n_u ? b0[i] : b0[i] + mass[i - n_u] * x[i];
}

// Block preconditioner diagonal of a coupled saddle point whose coupling blocks are not
// antisymmetric: flux rows 1/|a_kk|, pressure rows 1/(|a_kk| + sum_j |a_kj a_jk| / |a_jj|)
// over flux columns j (the absolute Schur diagonal).
__global__ __launch_bounds__(256) void k_md_block_diag(int n, int nu, const int32_t* __restrict__ rp,
                                                       const int32_t* __restrict__ ci,
                                                       const double* __restrict__ va, float* __restrict__ pinv) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  double akk = 0;
  for (int jj = rp[k]; jj < rp[k + 1]; ++jj)
    if (ci[jj] == k) akk = va[jj];
  if (k < nu) {
    pinv[k] = akk != 0 ? (float)(1.0 / fabs(akk)) : 1.f;
    return;
  }
  double s = fabs(akk);
  for (int jj = rp[k]; jj < rp[k + 1]; ++jj) {
    const int j = ci[jj];
    if (j >= nu) continue;
    double ajk = 0, ajj = 0;
    for (int t = rp[j]; t < rp[j + 1]; ++t) {
      if (ci[t] == k) ajk = va[t];
      if (ci[t] == j) ajj = va[t];
    }
    if (ajj != 0) s += fabs(va[jj] * ajk) / fabs(ajj);
  }
  pinv[k] = s > 0 ? (float)(1.0 / s) : 1.f;
}

}  // namespace nfx

using namespace nfx;

extern "C" int nfx_assemble_darcy(int dim, int k, int n_elem, const int32_t* d_cells,
                                  const void* d_coords, const void* d_tables, const void* d_kinv,
                                  int kinv_stride_e, int kinv_stride_q, const void* d_f,
                                  int f_stride_e, int f_stride_q, const int32_t* d_slot,
                                  const int32_t* d_rhs_row, int fp64, void* d_vals,
                                  void* d_stage, void* d_rhs, void* stream) {
  NFX_REQUIRE(n_elem >= 0, "nfx_assemble_darcy: n_elem < 0");
  if (n_elem == 0) return NFX_OK;
  NFX_REQUIRE(d_cells && d_coords && d_tables && d_kinv && d_f && d_slot && d_rhs_row && d_vals &&
                  d_rhs,
              "nfx_assemble_darcy: null pointer");
  hipStream_t s = static_cast<hipStream_t>(stream);
#define NFX_DARCY(DIM, NB, NP, NQ)                                                                \
  return fp64 ? launch_darcy<DIM, NB, NP, NQ, double>(n_elem, d_cells, d_coords, d_tables, d_kinv, \
                                                      kinv_stride_e, kinv_stride_q, d_f,          \
                                                      f_stride_e, f_stride_q, d_slot, d_rhs_row,  \
                                                      d_vals, d_stage, d_rhs, s)                  \
              : launch_darcy<DIM, NB, NP, NQ, float>(n_elem, d_cells, d_coords, d_tables, d_kinv,  \
                                                     kinv_stride_e, kinv_stride_q, d_f,           \
                                                     f_stride_e, f_stride_q, d_slot, d_rhs_row,   \
                                                     d_vals, d_stage, d_rhs, s)
  if (dim == 2 && k == 1) NFX_DARCY(2, 3, 1, 4);
  if (dim == 2 && k == 2) NFX_DARCY(2, 8, 3, 9);
  if (dim == 3 && k == 1) NFX_DARCY(3, 4, 1, 8);
  if (dim == 3 && k == 2) NFX_DARCY(3, 15, 4, 27);
#undef NFX_DARCY
  set_error("nfx_assemble_darcy: unsupported (dim, k)");
  return NFX_E_UNSUPPORTED;
}

extern "C" int nfx_assemble_darcy_bc(int dim, int k, int n_facets, const int32_t* d_facet,
                                     const int32_t* d_neigh, const int32_t* d_combo,
                                     const int32_t* d_cells, const void* d_coords,
                                     const void* d_trace, const void* d_fw, int nqf,
                                     const void* d_pD, const int32_t* d_rows, int fp64,
                                     void* d_rhs, void* stream) {
  NFX_REQUIRE(n_facets >= 0, "nfx_assemble_darcy_bc: n_facets < 0");
  if (n_facets == 0) return NFX_OK;
  NFX_REQUIRE(d_facet && d_neigh && d_combo && d_cells && d_coords && d_trace && d_fw && d_pD &&
                  d_rows && d_rhs,
              "nfx_assemble_darcy_bc: null pointer");
  hipStream_t s = static_cast<hipStream_t>(stream);
#define NFX_DBC(DIM, FD, NQF)                                                                     \
  return fp64 ? launch_darcy_bc<DIM, FD, NQF, double>(n_facets, d_facet, d_neigh, d_combo,         \
                                                      d_cells, d_coords, d_trace, d_fw, d_pD,      \
                                                      d_rows, d_rhs, s)                            \
              : launch_darcy_bc<DIM, FD, NQF, float>(n_facets, d_facet, d_neigh, d_combo, d_cells, \
                                                     d_coords, d_trace, d_fw, d_pD, d_rows, d_rhs, \
                                                     s)
  if (dim == 2 && k == 1 && nqf == 2) NFX_DBC(2, 1, 2);
  if (dim == 2 && k == 2 && nqf == 3) NFX_DBC(2, 2, 3);
  if (dim == 3 && k == 1 && nqf == 4) NFX_DBC(3, 1, 4);
  if (dim == 3 && k == 2 && nqf == 9) NFX_DBC(3, 3, 9);
#undef NFX_DBC
  set_error("nfx_assemble_darcy_bc: unsupported (dim, k, nqf)");
  return NFX_E_UNSUPPORTED;
}

extern "C" int nfx_stage_reduce(int n_seg, const int32_t* d_segptr, const int32_t* d_target,
                                int fp64, const void* d_stage, void* d_vals, void* stream) {
  NFX_REQUIRE(n_seg >= 0, "nfx_stage_reduce: n_seg < 0");
  if (n_seg == 0) return NFX_OK;
  NFX_REQUIRE(d_segptr && d_target && d_stage && d_vals, "nfx_stage_reduce: null pointer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (fp64)
    hipLaunchKernelGGL(k_stage_reduce<double>, dim3(grid_for(n_seg, 256)), dim3(256), 0, s, n_seg,
                       d_segptr, d_target, static_cast<const double*>(d_stage),
                       static_cast<double*>(d_vals));
  else
    hipLaunchKernelGGL(k_stage_reduce<float>, dim3(grid_for(n_seg, 256)), dim3(256), 0, s, n_seg,
                       d_segptr, d_target, static_cast<const float*>(d_stage),
                       static_cast<float*>(d_vals));
  NFX_LAUNCH_CHECK();
  return NFX_OK;
}

extern "C" int nfx_fill(int fp64, int64_t n, double value, void* d_x, void* stream) {
  NFX_REQUIRE(n >= 0, "nfx_fill: n < 0");
  if (n == 0) return NFX_OK;
  NFX_REQUIRE(d_x, "nfx_fill: null pointer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int64_t blocks64 = (n + 255) / 256;
  const int blocks = static_cast<int>(blocks64 < 65536 ? blocks64 : 65536);
  if (fp64)
    hipLaunchKernelGGL(k_fill<double>, dim3(blocks), dim3(256), 0, s, n, value,
                       static_cast<double*>(d_x));
  else
    hipLaunchKernelGGL(k_fill<float>, dim3(blocks), dim3(256), 0, s, n, static_cast<float>(value),
                       static_cast<float*>(d_x));
  NFX_LAUNCH_CHECK();
  return NFX_OK;
}

extern "C" int nfx_quadrature_points(int dim, int n_elem, int nq, const int32_t* d_cells,
                                     const double* d_coords64, const double* d_bary,
                                     double* d_out, void* stream) {
  NFX_REQUIRE(dim == 1 || dim == 2 || dim == 3, "nfx_quadrature_points: dim");
  NFX_REQUIRE(n_elem >= 0 && nq > 0, "nfx_quadrature_points: sizes");
  if (n_elem == 0) return NFX_OK;
  NFX_REQUIRE(d_cells && d_coords64 && d_bary && d_out, "nfx_quadrature_points: null pointer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(k_qpoints, dim3(grid_for((int64_t)n_elem * nq, 256)), dim3(256), 0, s, dim,
                     n_elem, nq, d_cells, d_coords64, d_bary, d_out);
  NFX_LAUNCH_CHECK();
  return NFX_OK;
}

extern "C" int nfx_assemble_darcy_rows(int dim, int k, int n_rows, const int32_t* d_inc,
                                       const uint8_t* d_pos, const int32_t* d_rowptr,
                                       const uint32_t* d_signs, const int32_t* d_cells,
                                       const void* d_coords, const void* d_tables,
                                       const void* d_kinv, int kinv_stride_e, int kinv_stride_q,
                                       const void* d_f, int f_stride_e, int f_stride_q, const void* d_geo,
                                       const int32_t* d_tile_map, int n_tile_map, int fp64, void* d_vals,
                                       void* d_rhs, void* stream) {
  NFX_REQUIRE(n_rows >= 0, "nfx_assemble_darcy_rows: n_rows < 0");
  NFX_REQUIRE(d_tile_map == nullptr || n_tile_map >= 0, "nfx_assemble_darcy_rows: n_tile_map < 0");
  NFX_REQUIRE(reinterpret_cast<uintptr_t>(d_geo) % 16 == 0, "nfx_assemble_darcy_rows: d_geo not 16-byte aligned");
  if (n_rows == 0) return NFX_OK;
  NFX_REQUIRE(d_inc && d_pos && d_rowptr && d_signs && d_cells && d_coords && d_tables && d_kinv &&
                  d_f && d_vals && d_rhs,
              "nfx_assemble_darcy_rows: null pointer");
  hipStream_t s = static_cast<hipStream_t>(stream);
#define NFX_DROWS(DIM, NB, NP, NQ)                                                                 \
  return fp64 ? launch_darcy_rows<DIM, NB, NP, NQ, double>(                                        \
                    n_rows, d_inc, d_pos, d_rowptr, d_signs, d_cells, d_coords, d_tables, d_kinv,  \
                    kinv_stride_e, kinv_stride_q, d_f, f_stride_e, f_stride_q, d_vals, d_rhs, s,   \
                    0, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, d_geo, d_tile_map,    \
                    n_tile_map)                                                                    \
              : launch_darcy_rows<DIM, NB, NP, NQ, float>(                                         \
                    n_rows, d_inc, d_pos, d_rowptr, d_signs, d_cells, d_coords, d_tables, d_kinv,  \
                    kinv_stride_e, kinv_stride_q, d_f, f_stride_e, f_stride_q, d_vals, d_rhs, s,   \
                    0, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, d_geo, d_tile_map,    \
                    n_tile_map)
  if (dim == 2 && k == 1) NFX_DROWS(2, 3, 1, 4);
  if (dim == 2 && k == 2) NFX_DROWS(2, 8, 3, 9);
  if (dim == 3 && k == 1) NFX_DROWS(3, 4, 1, 8);
  if (dim == 3 && k == 2) NFX_DROWS(3, 15, 4, 27);
#undef NFX_DROWS
  set_error("nfx_assemble_darcy_rows: unsupported (dim, k)");
  return NFX_E_UNSUPPORTED;
}

extern "C" int nfx_assemble_darcy_rows_bc(
    int dim, int k, int n_rows, const int32_t* d_inc, const uint8_t* d_pos, const int32_t* d_rowptr,
    const uint32_t* d_signs, const int32_t* d_cells, const void* d_coords, const void* d_tables,
    const void* d_kinv, int kinv_stride_e, int kinv_stride_q, const void* d_f, int f_stride_e,
    int f_stride_q, int n_facets, const int32_t* d_facet, const int32_t* d_combo, const void* d_trace,
    const void* d_fw, int nqf, const void* d_pD, const int32_t* d_rows, const void* d_geo,
    const int32_t* d_tile_map, int n_tile_map, int fp64, void* d_vals, void* d_rhs, void* stream) {
  NFX_REQUIRE(n_rows >= 0 && n_facets >= 0, "nfx_assemble_darcy_rows_bc: negative size");
  NFX_REQUIRE(d_tile_map == nullptr || n_tile_map >= 0, "nfx_assemble_darcy_rows_bc: n_tile_map < 0");
  if (n_rows == 0) return NFX_OK;
  NFX_REQUIRE(d_inc && d_pos && d_rowptr && d_signs && d_cells && d_coords && d_tables && d_kinv &&
                  d_f && d_vals && d_rhs,
              "nfx_assemble_darcy_rows_bc: null pointer");
  NFX_REQUIRE(n_facets == 0 || (d_facet && d_combo && d_trace && d_fw && d_pD && d_rows),
              "nfx_assemble_darcy_rows_bc: null facet pointer");
  NFX_REQUIRE(reinterpret_cast<uintptr_t>(d_geo) % 16 == 0, "nfx_assemble_darcy_rows_bc: d_geo not 16-byte aligned");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const void* pD = n_facets ? d_pD : nullptr;
#define NFX_DROWSBC(DIM, NB, NP, NQ)                                                                  \
  do {                                                                                                \
    constexpr int FD = DarcyFacetShape<DIM, NB>::FD;                                                  \
    constexpr int NQF = DarcyFacetShape<DIM, NB>::NQF;                                                \
    NFX_REQUIRE(n_facets == 0 || nqf == NQF,                                                          \
                "nfx_assemble_darcy_rows_bc: facet quadrature does not match the element");           \
    return fp64 ? launch_darcy_rows<DIM, NB, NP, NQ, double>(                                         \
                      n_rows, d_inc, d_pos, d_rowptr, d_signs, d_cells, d_coords, d_tables, d_kinv,   \
                      kinv_stride_e, kinv_stride_q, d_f, f_stride_e, f_stride_q, d_vals, d_rhs, s,     \
                      n_facets * FD, d_facet, d_combo, d_trace, d_fw, pD, d_rows, d_geo, d_tile_map,   \
                      n_tile_map)                                                                     \
                : launch_darcy_rows<DIM, NB, NP, NQ, float>(                                          \
                      n_rows, d_inc, d_pos, d_rowptr, d_signs, d_cells, d_coords, d_tables, d_kinv,   \
                      kinv_stride_e, kinv_stride_q, d_f, f_stride_e, f_stride_q, d_vals, d_rhs, s,     \
                      n_facets * FD, d_facet, d_combo, d_trace, d_fw, pD, d_rows, d_geo, d_tile_map,   \
                      n_tile_map);                                                                    \
  } while (0)
  if (dim == 2 && k == 1) NFX_DROWSBC(2, 3, 1, 4);
  if (dim == 2 && k == 2) NFX_DROWSBC(2, 8, 3, 9);
  if (dim == 3 && k == 1) NFX_DROWSBC(3, 4, 1, 8);
  if (dim == 3 && k == 2) NFX_DROWSBC(3, 15, 4, 27);
#undef NFX_DROWSBC
  set_error("nfx_assemble_darcy_rows_bc: unsupported (dim, k)");
  return NFX_E_UNSUPPORTED;
}

extern "C" int nfx_element_geometry(int dim, int n_elem, const int32_t* d_cells, const void* d_coords, int fp64,
                                    void* d_geo, void* stream) {
  NFX_REQUIRE((dim == 2 || dim == 3) && n_elem >= 0, "nfx_element_geometry: sizes");
  if (n_elem == 0) return NFX_OK;
  NFX_REQUIRE(d_cells && d_coords && d_geo, "nfx_element_geometry: null pointer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const dim3 grid(grid_for(n_elem, 256));
  if (dim == 3) {
    if (fp64)
      hipLaunchKernelGGL((k_element_geometry<3, double>), grid, dim3(256), 0, s, n_elem, d_cells,
                         static_cast<const double*>(d_coords), static_cast<double*>(d_geo));
    else
      hipLaunchKernelGGL((k_element_geometry<3, float>), grid, dim3(256), 0, s, n_elem, d_cells,
                         static_cast<const float*>(d_coords), static_cast<float*>(d_geo));
  } else {
    if (fp64)
      hipLaunchKernelGGL((k_element_geometry<2, double>), grid, dim3(256), 0, s, n_elem, d_cells,
                         static_cast<const double*>(d_coords), static_cast<double*>(d_geo));
    else
      hipLaunchKernelGGL((k_element_geometry<2, float>), grid, dim3(256), 0, s, n_elem, d_cells,
                         static_cast<const float*>(d_coords), static_cast<float*>(d_geo));
  }
  NFX_LAUNCH_CHECK();
  return NFX_OK;
}

extern "C" int nfx_assemble_elasticity_bc(int dim, int n_facets, const int32_t* d_facet,
                                          const int32_t* d_neigh, const int32_t* d_combo,
                                          const int32_t* d_cells, const void* d_coords,
                                          const void* d_trace, const void* d_fw, int nqf,
                                          const void* d_uD, const int32_t* d_rows, int fp64,
                                          void* d_rhs, void* stream) {
  NFX_REQUIRE(dim == 2 || dim == 3, "nfx_assemble_elasticity_bc: dim must be 2 or 3");
  NFX_REQUIRE(n_facets >= 0, "nfx_assemble_elasticity_bc: n_facets < 0");
  if (n_facets == 0) return NFX_OK;
  NFX_REQUIRE(nqf == (dim == 3 ? 4 : 2), "nfx_assemble_elasticity_bc: facet rule must have 4 (dim 3) / 2 (dim 2) points");
  NFX_REQUIRE(d_facet && d_neigh && d_combo && d_cells && d_coords && d_trace && d_fw && d_uD &&
                  d_rows && d_rhs,
              "nfx_assemble_elasticity_bc: null pointer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (dim == 2)
    return fp64 ? launch_darcy_bc<2, 2, 2, double, 2>(n_facets, d_facet, d_neigh, d_combo, d_cells, d_coords,
                                                      d_trace, d_fw, d_uD, d_rows, d_rhs, s, -1.0)
                : launch_darcy_bc<2, 2, 2, float, 2>(n_facets, d_facet, d_neigh, d_combo, d_cells, d_coords,
                                                     d_trace, d_fw, d_uD, d_rows, d_rhs, s, -1.0);
  return fp64 ? launch_darcy_bc<3, 3, 4, double, 3>(n_facets, d_facet, d_neigh, d_combo, d_cells,
                                                    d_coords, d_trace, d_fw, d_uD, d_rows, d_rhs,
                                                    s, -1.0)
              : launch_darcy_bc<3, 3, 4, float, 3>(n_facets, d_facet, d_neigh, d_combo, d_cells,
                                                   d_coords, d_trace, d_fw, d_uD, d_rows, d_rhs, s,
                                                   -1.0);
}

template <int KD>
static int launch_batched(const DarcyBatchItem* items, const int2* tiles, int n_tiles, int tile_rows, int grid,
                          hipStream_t s) {
  if (tile_rows == 64)
    hipLaunchKernelGGL((k_darcy_rows_batched<64, KD>), dim3(grid), dim3(64), 0, s, items, tiles, n_tiles);
  else if (tile_rows == 128)
    hipLaunchKernelGGL((k_darcy_rows_batched<128, KD>), dim3(grid), dim3(128), 0, s, items, tiles, n_tiles);
  else if (tile_rows == 256)
    hipLaunchKernelGGL((k_darcy_rows_batched<256, KD>), dim3(grid), dim3(256), 0, s, items, tiles, n_tiles);
  else {
    set_error("nfx_assemble_darcy_batched: tile_rows must be 64, 128 or 256");
    return NFX_E_INVALID;
  }
  NFX_LAUNCH_CHECK();
  return NFX_OK;
}

extern "C" int nfx_assemble_darcy_batched(int n_items, const void* d_items, int n_tiles,
                                          const int32_t* d_tiles, int tile_rows, int grid, int degree,
                                          void* stream) {
  NFX_REQUIRE(n_items >= 0 && n_tiles >= 0 && grid > 0, "nfx_assemble_darcy_batched: sizes");
  NFX_REQUIRE(degree >= 0 && degree <= 3, "nfx_assemble_darcy_batched: degree must be 0, 1, 2 or 3");
  if (n_tiles == 0) return NFX_OK;
  NFX_REQUIRE(d_items && d_tiles, "nfx_assemble_darcy_batched: null pointer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const DarcyBatchItem* items = static_cast<const DarcyBatchItem*>(d_items);
  const int2* tiles = reinterpret_cast<const int2*>(d_tiles);
  if (degree == 1) return launch_batched<1>(items, tiles, n_tiles, tile_rows, grid, s);
  if (degree == 2) return launch_batched<2>(items, tiles, n_tiles, tile_rows, grid, s);
  if (degree == 3) return launch_batched<3>(items, tiles, n_tiles, tile_rows, grid, s);
  return launch_batched<0>(items, tiles, n_tiles, tile_rows, grid, s);
}

extern "C" int nfx_darcy_batch_item_size(void) { return (int)sizeof(DarcyBatchItem); }

extern "C" int nfx_md_coupling(int m, const int32_t* d_side_elem, const int32_t* d_side_combo,
                               const int32_t* d_side_code, const int32_t* d_facet_nodes, const int32_t* d_cells,
                               const void* d_coords, const void* d_trace, const void* d_fw,
                               const void* d_inv_alpha, const int32_t* d_pos, int exchange_sign, int fp64,
                               void* d_vals, void* stream) {
  NFX_REQUIRE(m >= 0 && (exchange_sign == 1 || exchange_sign == -1), "nfx_md_coupling: bad arguments");
  if (m == 0) return NFX_OK;
  NFX_REQUIRE(d_side_elem && d_side_combo && d_side_code && d_facet_nodes && d_cells && d_coords && d_trace &&
                  d_fw && d_inv_alpha && d_pos && d_vals,
              "nfx_md_coupling: null pointer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (fp64)
    hipLaunchKernelGGL(k_md_coupling<double>, dim3(grid_for(m, 256)), dim3(256), 0, s, m, d_side_elem, d_side_combo,
                       d_side_code, d_facet_nodes, d_cells, static_cast<const double*>(d_coords),
                       static_cast<const double*>(d_trace), static_cast<const double*>(d_fw),
                       static_cast<const double*>(d_inv_alpha), d_pos, (double)exchange_sign,
                       static_cast<double*>(d_vals));
  else
    hipLaunchKernelGGL(k_md_coupling<float>, dim3(grid_for(m, 256)), dim3(256), 0, s, m, d_side_elem, d_side_combo,
                       d_side_code, d_facet_nodes, d_cells, static_cast<const float*>(d_coords),
                       static_cast<const float*>(d_trace), static_cast<const float*>(d_fw),
                       static_cast<const float*>(d_inv_alpha), d_pos, (float)exchange_sign,
                       static_cast<float*>(d_vals));
  NFX_LAUNCH_CHECK();
  return NFX_OK;
}

extern "C" int nfx_md_scatter(int n, const void* d_src, const int32_t* d_map, int fp64, void* d_dst, void* stream) {
  NFX_REQUIRE(n >= 0, "nfx_md_scatter: n < 0");
  if (n == 0) return NFX_OK;
  NFX_REQUIRE(d_src && d_map && d_dst, "nfx_md_scatter: null pointer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (fp64)
    hipLaunchKernelGGL(k_md_scatter<double>, dim3(grid_for(n, 256)), dim3(256), 0, s, n,
                       static_cast<const double*>(d_src), d_map, static_cast<double*>(d_dst));
  else
    hipLaunchKernelGGL(k_md_scatter<float>, dim3(grid_for(n, 256)), dim3(256), 0, s, n,
                       static_cast<const float*>(d_src), d_map, static_cast<float*>(d_dst));
  NFX_LAUNCH_CHECK();
  return NFX_OK;
}

extern "C" int nfx_md_update(int n, const int32_t* d_pos, const double* d_v, int mode, int fp64, void* d_vals,
                             void* stream) {
  NFX_REQUIRE(n >= 0 && (mode == 0 || mode == 1), "nfx_md_update: bad arguments");
  if (n == 0) return NFX_OK;
  NFX_REQUIRE(d_pos && d_v && d_vals, "nfx_md_update: null pointer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (fp64)
    hipLaunchKernelGGL(k_md_update<double>, dim3(grid_for(n, 256)), dim3(256), 0, s, n, d_pos, d_v, mode,
                       static_cast<double*>(d_vals));
  else
    hipLaunchKernelGGL(k_md_update<float>, dim3(grid_for(n, 256)), dim3(256), 0, s, n, d_pos, d_v, mode,
                       static_cast<float*>(d_vals));
  NFX_LAUNCH_CHECK();
  return NFX_OK;
}

extern "C" int nfx_md_rhs(int n, int n_u, const double* d_b0, const double* d_mass, const double* d_x,
                          double* d_b, void* stream) {
  NFX_REQUIRE(n >= 0 && n_u >= 0 && n_u <= n, "nfx_md_rhs: sizes");
  if (n == 0) return NFX_OK;
  NFX_REQUIRE(d_b0 && d_x && d_b && (n_u == n || d_mass), "nfx_md_rhs: null pointer");
  hipLaunchKernelGGL(k_md_rhs, dim3(grid_for(n, 256)), dim3(256), 0, static_cast<hipStream_t>(stream), n, n_u, d_b0,
                     d_mass, d_x, d_b);
  NFX_LAUNCH_CHECK();
  return NFX_OK;
}

extern "C" int nfx_md_block_diag(int n, int n_u, const int32_t* d_rowptr, const int32_t* d_col, const double* d_vals,
                                 float* d_pinv, void* stream) {
  NFX_REQUIRE(n >= 0 && n_u >= 0 && n_u <= n, "nfx_md_block_diag: sizes");
  if (n == 0) return NFX_OK;
  NFX_REQUIRE(d_rowptr && d_col && d_vals && d_pinv, "nfx_md_block_diag: null pointer");
  hipLaunchKernelGGL(k_md_block_diag, dim3(grid_for(n, 256)), dim3(256), 0, static_cast<hipStream_t>(stream), n, n_u,
                     d_rowptr, d_col, d_vals, d_pinv);
  NFX_LAUNCH_CHECK();
  return NFX_OK;
}
