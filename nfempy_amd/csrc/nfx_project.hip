// L2 projection: element mass matrix and right-hand side of one field, staged into CSR.
//
// Restates L2ProjectorWeakForm.evaluate_form (src/weak_forms/l2_projector_weak_form.py:8-93)
// as used by l2_projector (src/postprocess/projectors.py:8-91), per element e and
// component c of the field (DoF (a, c) at element-local index off + a * nc + c):
//   vector (RT / BDM rows): phi_a = J phi_ref_a / detJ (Piola), the reference keeps its
//     first dim components (tab[..., 0:dim]):
//       M_ab = sum_q w_q detJ sum_{d<dim} phi_a,d phi_b,d,   r_a = sum_q w_q detJ f_c(x_q) . phi_a
//   scalar (Lagrange):  M_ab = sum_q w_q detJ psi_a psi_b,   r_a = sum_q w_q detJ f_c(x_q) psi_a
// (the reference's r_el carries a minus sign and the driver solves with b = -rg; r here
// is that b). The element matrix is the same for every component and has no entries
// between components. Every matrix contribution passes the reference's scatter filter,
// which drops a local value v with |v - 1e-16| <= 1e-8 + 1e-5 * 1e-16 (numpy.isclose
// defaults; projectors.py:41); a dropped value contributes 0 here.
// Arithmetic in FP64; values stored as R.
#include "nfx_common.h"

namespace nfx {

constexpr int kPB = 256;

template <typename R>
__device__ __forceinline__ void put_filtered(int code, double v, R* __restrict__ vals, R* __restrict__ stage) {
  if (code < 0) return;
  const double sv = (code & kSlotNeg) ? -v : v;
  if (fabs(sv - 1e-16) <= 1e-8 + 1e-5 * 1e-16) v = 0.0;
  put_slot<R>(code, (R)v, vals, stage);
}

template <int VECTOR, typename R>
__global__ __launch_bounds__(kPB) void k_projector(int dim, int nb, int nq, int nc, int nl, int off, int n_elem,
                                                   const int32_t* __restrict__ cells,
                                                   const double* __restrict__ coords,
                                                   const double* __restrict__ tab, const double* __restrict__ f,
                                                   const int32_t* __restrict__ slot,
                                                   const int32_t* __restrict__ rhs_slot, R* __restrict__ vals,
                                                   R* __restrict__ stage, R* __restrict__ rhs,
                                                   R* __restrict__ rhs_stage) {
  extern __shared__ double st[];
  const int ntab = nq + nq * nb * (VECTOR ? dim : 1);
  for (int k = threadIdx.x; k < ntab; k += kPB) st[k] = tab[k];
  __syncthreads();
  const int e = blockIdx.x * kPB + threadIdx.x;
  if (e >= n_elem) return;
  const double* w = st;
  const double* phi = st + nq;
  double X[4][3];
  for (int a = 0; a <= dim; ++a) {
    const int v = cells[(int64_t)e * (dim + 1) + a];
    for (int c = 0; c < 3; ++c) X[a][c] = coords[3 * (int64_t)v + c];
  }
  double J[3][3];
  for (int c = 0; c < 3; ++c)
    for (int d = 0; d < 3; ++d) J[c][d] = d < dim ? X[d + 1][c] - X[0][c] : 0.0;
  double detJ;
  if (dim == 3) {
    detJ = fabs(J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) - J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]));
  } else {
    const double cx = J[1][0] * J[2][1] - J[2][0] * J[1][1];
    const double cy = J[2][0] * J[0][1] - J[0][0] * J[2][1];
    const double cz = J[0][0] * J[1][1] - J[1][0] * J[0][1];
    detJ = sqrt(cx * cx + cy * cy + cz * cz);
  }
  const int64_t E = n_elem;
  // mapped basis component d of DoF a at point q (times detJ: J phi_ref)
  auto jphi = [&](int q, int a, int d) {
    double s = 0;
    for (int k = 0; k < dim; ++k) s += J[d][k] * phi[(q * nb + a) * dim + k];
    return s;
  };
  for (int a = 0; a < nb; ++a) {
    for (int b = a; b < nb; ++b) {
      double m = 0;
      for (int q = 0; q < nq; ++q) {
        if (VECTOR) {
          double s = 0;
          for (int d = 0; d < dim; ++d) s += jphi(q, a, d) * jphi(q, b, d);
          m += w[q] * s / detJ;
        } else {
          m += w[q] * detJ * phi[q * nb + a] * phi[q * nb + b];
        }
      }
      for (int c = 0; c < nc; ++c) {
        const int i = off + a * nc + c, j = off + b * nc + c;
        put_filtered<R>(slot[((int64_t)i * nl + j) * E + e], m, vals, stage);
        if (a != b) put_filtered<R>(slot[((int64_t)j * nl + i) * E + e], m, vals, stage);
      }
    }
    for (int c = 0; c < nc; ++c) {
      double r = 0;
      for (int q = 0; q < nq; ++q) {
        if (VECTOR) {
          const double* fq = f + (((int64_t)e * nq + q) * nc + c) * dim;
          double s = 0;
          for (int d = 0; d < dim; ++d) s += fq[d] * jphi(q, a, d);
          r += w[q] * s;
        } else {
          r += w[q] * detJ * f[((int64_t)e * nq + q) * nc + c] * phi[q * nb + a];
        }
      }
      const int code = rhs_slot[(int64_t)(off + a * nc + c) * E + e];
      if (code >= 0) put_slot<R>(code, (R)r, rhs, rhs_stage);
    }
  }
}

template <int VECTOR, typename R>
void launch_projector(int dim, int nb, int nq, int nc, int nl, int off, int n_elem, const int32_t* cells,
                      const double* coords, const double* tab, const double* f, const int32_t* slot,
                      const int32_t* rhs_slot, void* vals, void* stage, void* rhs, void* rhs_stage, hipStream_t st) {
  const size_t lds = sizeof(double) * (nq + (size_t)nq * nb * (VECTOR ? dim : 1));
  hipLaunchKernelGGL((k_projector<VECTOR, R>), dim3(grid_for(n_elem, kPB)), dim3(kPB), lds, st, dim, nb, nq, nc, nl,
                     off, n_elem, cells, coords, tab, f, slot, rhs_slot, static_cast<R*>(vals),
                     static_cast<R*>(stage), static_cast<R*>(rhs), static_cast<R*>(rhs_stage));
}

}  // namespace nfx

using namespace nfx;

extern "C" int nfx_assemble_projector(int kind, int dim, int nb, int nq, int n_comp, int nl, int off, int n_elem,
                                      const int32_t* d_cells, const double* d_coords64,
                                      const double* d_tables64, const double* d_f, const int32_t* d_slot,
                                      const int32_t* d_rhs_slot, int fp64, void* d_vals, void* d_stage,
                                      void* d_rhs, void* d_rhs_stage, void* stream) {
  NFX_REQUIRE(kind == 0 || kind == 1, "nfx_assemble_projector: kind");
  NFX_REQUIRE((dim == 2 || dim == 3) && nb > 0 && nq > 0 && n_comp > 0 && n_elem >= 0 && off >= 0,
              "nfx_assemble_projector: sizes");
  NFX_REQUIRE(off + nb * n_comp <= nl, "nfx_assemble_projector: field block outside the element's DoFs");
  NFX_REQUIRE((size_t)nq * (1 + (size_t)nb * (kind == 0 ? dim : 1)) * sizeof(double) <= 48 * 1024,
              "nfx_assemble_projector: tables exceed 48 KiB");
  if (n_elem == 0) return NFX_OK;
  NFX_REQUIRE(d_cells && d_coords64 && d_tables64 && d_f && d_slot && d_rhs_slot && d_vals && d_rhs,
              "nfx_assemble_projector: null pointer");
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (kind == 0) {
    if (fp64)
      launch_projector<1, double>(dim, nb, nq, n_comp, nl, off, n_elem, d_cells, d_coords64, d_tables64, d_f, d_slot,
                                  d_rhs_slot, d_vals, d_stage, d_rhs, d_rhs_stage, st);
    else
      launch_projector<1, float>(dim, nb, nq, n_comp, nl, off, n_elem, d_cells, d_coords64, d_tables64, d_f, d_slot,
                                 d_rhs_slot, d_vals, d_stage, d_rhs, d_rhs_stage, st);
  } else {
    if (fp64)
      launch_projector<0, double>(dim, nb, nq, n_comp, nl, off, n_elem, d_cells, d_coords64, d_tables64, d_f, d_slot,
                                  d_rhs_slot, d_vals, d_stage, d_rhs, d_rhs_stage, st);
    else
      launch_projector<0, float>(dim, nb, nq, n_comp, nl, off, n_elem, d_cells, d_coords64, d_tables64, d_f, d_slot,
                                 d_rhs_slot, d_vals, d_stage, d_rhs, d_rhs_stage, st);
  }
  NFX_LAUNCH_CHECK();
  return NFX_OK;
}
