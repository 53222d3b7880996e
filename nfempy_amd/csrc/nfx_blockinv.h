// Small dense SPD block inverse shared by the elasticity preconditioner kernels.
#pragma once

#include "nfx_common.h"

namespace nfx {

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

}  // namespace nfx
