// Shared helpers of libnfx: status codes, thread-local error text, the CSR
// slot-code encoding used by every assembly kernel, and small device math.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/nfx.h"

namespace nfx {

void set_error(const std::string& msg);

#define NFX_CHECK_HIP(expr)                                                     \
  do {                                                                          \
    hipError_t _e = (expr);                                                     \
    if (_e != hipSuccess) {                                                     \
      ::nfx::set_error(std::string(#expr) + ": " + hipGetErrorString(_e));     \
      return NFX_E_HIP;                                                         \
    }                                                                           \
  } while (0)

#define NFX_REQUIRE(cond, msg)                                                  \
  do {                                                                          \
    if (!(cond)) {                                                              \
      ::nfx::set_error(msg);                                                    \
      return NFX_E_INVALID;                                                     \
    }                                                                           \
  } while (0)

#define NFX_LAUNCH_CHECK()                                                      \
  do {                                                                          \
    hipError_t _e = hipGetLastError();                                          \
    if (_e != hipSuccess) {                                                     \
      ::nfx::set_error(std::string("kernel launch: ") + hipGetErrorString(_e)); \
      return NFX_E_HIP;                                                         \
    }                                                                           \
  } while (0)

// Slot codes: one int32 per element-local matrix entry, in reference-basis order.
//   code < 0            entry dropped (value filter / lower triangle)
//   bits 0..27          index into CSR values, or into the stage array
//   bit 28  (ZERO)      contribute 0.0 instead of the computed value
//   bit 29  (STAGE)     index is into the stage array (multi-contributor slot)
//   bit 30  (NEG)       contribute -value (facet orientation sign)
constexpr int kSlotIndexMask = (1 << 28) - 1;
constexpr int kSlotZero = 1 << 28;
constexpr int kSlotStage = 1 << 29;
constexpr int kSlotNeg = 1 << 30;

template <typename R>
__device__ __forceinline__ void put_slot(int code, R v, R* __restrict__ vals,
                                         R* __restrict__ stage) {
  if (code < 0) return;
  if (code & kSlotNeg) v = -v;
  if (code & kSlotZero) v = R(0);
  const int idx = code & kSlotIndexMask;
  if (code & kSlotStage)
    stage[idx] = v;
  else
    vals[idx] = v;
}

// Row codes for residual vectors: bits 0..27 row, bit 30 NEG, code < 0 skip.
template <typename R>
__device__ __forceinline__ void put_row(int code, R v, R* __restrict__ rhs) {
  if (code < 0) return;
  if (code & kSlotNeg) v = -v;
  rhs[code & kSlotIndexMask] = v;
}

inline int grid_for(int64_t n, int block) { return static_cast<int>((n + block - 1) / block); }

// Two-grid block preconditioner z = M^-1 r (nfx_twogrid_apply), with the output
// divided by *dscale when dscale is not null (FGMRES passes the norm of an
// unnormalised Arnoldi vector).
int twogrid_apply_scaled(const nfx_twogrid_pc* pc, const float* r, float* z, const double* dscale,
                         hipStream_t stream);

// Elasticity two-grid block preconditioner z = P^-1 r (nfx_elasticity_tg_apply), output
// divided by *dscale when dscale is not null.
int el_twogrid_apply_scaled(const nfx_elasticity_tg* pc, const float* r, float* z, const double* dscale,
                            hipStream_t stream);

// Batched two-grid Schur cycle on the rows of one group or of two groups of different nb
// in shared launches (nfx_batch_tg_apply); mode as there.
int batch_tg_apply(const nfx_batch_tg* pc, int n_groups, const int32_t* mode, const float* r, float* z,
                   hipStream_t stream);

}  // namespace nfx
