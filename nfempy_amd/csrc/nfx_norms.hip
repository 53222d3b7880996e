// FP64 error norms: per-element squared L2 / div errors of a discrete field
// against exact values at the element quadrature points, reduced
// deterministically (block partials summed on the host in block order).
//
// Vector (H(div)) field with n_comp rows:
//   f_h[c][d](x_q) = sum_a alpha[row(a, c)] * sign * (J phi_a(q))_d / detJ
//   err_e = sum_q w_q detJ sum_{c,d} (f_e - f_h)^2
// Scalar (Lagrange) field: f_h[c] = sum_k alpha[row(k, c)] psi_k(q).
// div: f_h[c] = sum_a alpha[row(a, c)] * sign * div_a(q) / detJ.
// tables64: w[NQ] | phi[NQ][NB][DIM] | div[NQ][NB]   (vector)
//           w[NQ] | psi[NQ][NB]                      (scalar)
#include <vector>

#include "nfx_common.h"

namespace nfx {

constexpr int kNB = 256;

__device__ __forceinline__ double nrm_warp_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

__device__ double nrm_block_sum(double v, double* sh) {
  v = nrm_warp_sum(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double r = 0;
  if (threadIdx.x == 0)
    for (int i = 0; i < (int)(blockDim.x + 63) / 64; ++i) r += sh[i];
  __syncthreads();
  return r;
}

// kind: 0 = vector L2, 1 = scalar L2, 2 = vector div
__global__ __launch_bounds__(kNB) void k_field_error(
    int kind, int dim, int nb, int nq, int nc, int n_elem, const int32_t* __restrict__ cells,
    const double* __restrict__ coords, const double* __restrict__ tab,
    const int32_t* __restrict__ rows, const double* __restrict__ alpha,
    const double* __restrict__ exact, double* __restrict__ partial) {
  __shared__ double sh[kNB / 64];
  double acc = 0;
  for (int e = blockIdx.x * kNB + threadIdx.x; e < n_elem; e += gridDim.x * kNB) {
    double X[4][3];
    for (int a = 0; a <= dim; ++a) {
      const int v = cells[(int64_t)e * (dim + 1) + a];
      for (int c = 0; c < 3; ++c) X[a][c] = coords[3 * v + c];
    }
    double J[3][3];
    for (int c = 0; c < 3; ++c)
      for (int d = 0; d < dim; ++d) J[c][d] = X[d + 1][c] - X[0][c];
    double detJ;
    if (dim == 3) {
      const double dd = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                        J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                        J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
      detJ = fabs(dd);
    } else {
      const double cx = J[1][0] * J[2][1] - J[2][0] * J[1][1];
      const double cy = J[2][0] * J[0][1] - J[0][0] * J[2][1];
      const double cz = J[0][0] * J[1][1] - J[1][0] * J[0][1];
      detJ = sqrt(cx * cx + cy * cy + cz * cz);
    }
    const double* w = tab;
    for (int q = 0; q < nq; ++q) {
      double err = 0;
      for (int c = 0; c < nc; ++c) {
        if (kind == 1) {
          const double* psi = tab + nq;
          double fh = 0;
          for (int k = 0; k < nb; ++k) {
            const int code = rows[((int64_t)e * nb + k) * nc + c];
            fh += alpha[code & kSlotIndexMask] * psi[q * nb + k];
          }
          const double d = exact[((int64_t)e * nq + q) * nc + c] - fh;
          err += d * d;
        } else if (kind == 0) {
          const double* phi = tab + nq;
          double ref[3] = {0, 0, 0};
          for (int a = 0; a < nb; ++a) {
            const int code = rows[((int64_t)e * nb + a) * nc + c];
            double al = alpha[code & kSlotIndexMask];
            if (code & kSlotNeg) al = -al;
            for (int d = 0; d < dim; ++d) ref[d] += al * phi[(q * nb + a) * dim + d];
          }
          for (int d = 0; d < dim; ++d) {
            double fh = 0;
            for (int d2 = 0; d2 < dim; ++d2) fh += J[d][d2] * ref[d2];
            fh /= detJ;
            const double diff = exact[(((int64_t)e * nq + q) * nc + c) * dim + d] - fh;
            err += diff * diff;
          }
        } else {
          const double* dv = tab + nq + nq * nb * dim;
          double fh = 0;
          for (int a = 0; a < nb; ++a) {
            const int code = rows[((int64_t)e * nb + a) * nc + c];
            double al = alpha[code & kSlotIndexMask];
            if (code & kSlotNeg) al = -al;
            fh += al * dv[q * nb + a];
          }
          fh /= detJ;
          const double d = exact[((int64_t)e * nq + q) * nc + c] - fh;
          err += d * d;
        }
      }
      acc += w[q] * detJ * err;
    }
  }
  const double bs = nrm_block_sum(acc, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = bs;
}

// Field values at np reference points of every element (the vertex and centroid values
// of the VTK output): kind 0 vector, out[e][p][c][d] = (J phi_h)_d / detJ for d < dim;
// kind 1 scalar, out[e][p][c] = sum_k alpha psi_k(p). tab: phi[np][nb][dim] or psi[np][nb].
__global__ __launch_bounds__(kNB) void k_field_values(int kind, int dim, int nb, int np, int nc, int n_elem,
                                                      const int32_t* __restrict__ cells,
                                                      const double* __restrict__ coords,
                                                      const double* __restrict__ tab,
                                                      const int32_t* __restrict__ rows,
                                                      const double* __restrict__ alpha, double* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * kNB + threadIdx.x;
  if (t >= (int64_t)n_elem * np) return;
  const int e = (int)(t / np), p = (int)(t % np);
  if (kind == 1) {
    for (int c = 0; c < nc; ++c) {
      double fh = 0;
      for (int k = 0; k < nb; ++k) {
        const int code = rows[((int64_t)e * nb + k) * nc + c];
        fh += alpha[code & kSlotIndexMask] * tab[p * nb + k];
      }
      out[((int64_t)e * np + p) * nc + c] = fh;
    }
    return;
  }
  double X[4][3];
  for (int a = 0; a <= dim; ++a) {
    const int v = cells[(int64_t)e * (dim + 1) + a];
    for (int c = 0; c < 3; ++c) X[a][c] = coords[3 * (int64_t)v + c];
  }
  double J[3][3];
  for (int c = 0; c < 3; ++c)
    for (int d = 0; d < dim; ++d) J[c][d] = X[d + 1][c] - X[0][c];
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
  for (int c = 0; c < nc; ++c) {
    double ref[3] = {0, 0, 0};
    for (int a = 0; a < nb; ++a) {
      const int code = rows[((int64_t)e * nb + a) * nc + c];
      double al = alpha[code & kSlotIndexMask];
      if (code & kSlotNeg) al = -al;
      for (int d = 0; d < dim; ++d) ref[d] += al * tab[(p * nb + a) * dim + d];
    }
    for (int d = 0; d < dim; ++d) {
      double fh = 0;
      for (int d2 = 0; d2 < dim; ++d2) fh += J[d][d2] * ref[d2];
      out[(((int64_t)e * np + p) * nc + c) * dim + d] = fh / detJ;
    }
  }
}

}  // namespace nfx

using namespace nfx;

extern "C" int nfx_field_values(int kind, int dim, int nb, int n_points, int n_comp, int n_elem,
                                const int32_t* d_cells, const double* d_coords64, const double* d_tables64,
                                const int32_t* d_rows, const double* d_alpha, double* d_out, void* stream) {
  NFX_REQUIRE(kind == 0 || kind == 1, "nfx_field_values: kind");
  NFX_REQUIRE((dim == 2 || dim == 3) && nb > 0 && n_points > 0 && n_comp > 0 && n_elem >= 0,
              "nfx_field_values: sizes");
  if (n_elem == 0) return NFX_OK;
  NFX_REQUIRE(d_cells && d_coords64 && d_tables64 && d_rows && d_alpha && d_out, "nfx_field_values: null pointer");
  hipLaunchKernelGGL(k_field_values, dim3(grid_for((int64_t)n_elem * n_points, kNB)), dim3(kNB), 0,
                     static_cast<hipStream_t>(stream), kind, dim, nb, n_points, n_comp, n_elem, d_cells, d_coords64,
                     d_tables64, d_rows, d_alpha, d_out);
  NFX_LAUNCH_CHECK();
  return NFX_OK;
}

extern "C" int nfx_field_error(int kind, int dim, int nb, int nq, int n_comp, int n_elem,
                               const int32_t* d_cells, const double* d_coords64,
                               const double* d_tables64, const int32_t* d_rows,
                               const double* d_alpha, const double* d_exact, double* h_out,
                               void* stream) {
  NFX_REQUIRE(kind >= 0 && kind <= 2, "nfx_field_error: kind");
  NFX_REQUIRE((dim == 2 || dim == 3) && nb > 0 && nq > 0 && n_comp > 0 && n_elem >= 0,
              "nfx_field_error: sizes");
  NFX_REQUIRE(h_out, "nfx_field_error: null output");
  if (n_elem == 0) {
    *h_out = 0;
    return NFX_OK;
  }
  NFX_REQUIRE(d_cells && d_coords64 && d_tables64 && d_rows && d_alpha && d_exact,
              "nfx_field_error: null pointer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  int blocks = grid_for(n_elem, kNB);
  if (blocks > 1024) blocks = 1024;
  double* partial = nullptr;
  NFX_CHECK_HIP(hipMalloc(&partial, sizeof(double) * blocks));
  hipLaunchKernelGGL(k_field_error, dim3(blocks), dim3(kNB), 0, s, kind, dim, nb, nq, n_comp, n_elem,
                     d_cells, d_coords64, d_tables64, d_rows, d_alpha, d_exact, partial);
  hipError_t e = hipGetLastError();
  std::vector<double> h(blocks);
  if (e == hipSuccess)
    e = hipMemcpyAsync(h.data(), partial, sizeof(double) * blocks, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFree(partial);
  if (e != hipSuccess) {
    set_error(std::string("nfx_field_error: ") + hipGetErrorString(e));
    return NFX_E_HIP;
  }
  double acc = 0;
  for (int i = 0; i < blocks; ++i) acc += h[i];  // block order: deterministic
  *h_out = acc;
  return NFX_OK;
}
