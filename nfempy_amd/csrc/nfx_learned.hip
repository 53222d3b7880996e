// Learned correction of the Darcy block preconditioner (RT1-P0):
//   * the P0 pressure Schur approximation S = B diag(M)^-1 B^T of the assembled
//     matrix in a 4-wide ELL layout (row p lists the cells across its faces in
//     ascending face order; -1 = boundary face);
//   * a 4-layer message-passing network (width 64) on that graph that outputs the
//     entries of a symmetric sparse smoother W with the pattern of S
//     (W_pp = (0.65 + o_p) / S_pp, W_pq = o_pq / sqrt(S_pp S_qq));
//     weights and stored activations in bfloat16, accumulation in FP32, the layer
//     GEMMs on MFMA (v_mfma_f32_16x16x16_bf16), backward in FP32;
//   * the training loss kernels (ELL products over sample blocks, aggregate sums,
//     the loss gradient with respect to the entries of W);
//   * one symmetric two-grid cycle on S: pre-smooth W, exact coarse correction on
//     aggregates (dense inverse), post-smooth W.
// Every reduction runs in a fixed order: results are bitwise reproducible.
#include <vector>

#include "nfx_common.h"

namespace nfx {
namespace {

constexpr int kIn = 8;
constexpr int kW = 64;
constexpr int kLayers = 4;
constexpr float kOmega0 = 0.65f;
constexpr int kOffW0 = 0;
constexpr int kOffB0 = kIn * kW;
constexpr int kLayerBase = kOffB0 + kW;
constexpr int kLayerSize = 2 * kW * kW + kW;
constexpr int kOffUd = kLayerBase + kLayers * kLayerSize;
constexpr int kOffCd = kOffUd + kW;
constexpr int kOffUe = kOffCd + 1;
constexpr int kOffCe = kOffUe + kW;
constexpr int kNumParams = kOffCe + 1;
static_assert(kNumParams == 33730, "parameter layout");

typedef short short4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float(uint32_t(b) << 16); }
__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return uint16_t(u >> 16);
}

// ---------------------------------------------------------------------------
// S in ELL form from the assembled CSR values (FP64).
// pos_b[p][j]: value index of B_{p f_j}; pos_qb: value index of B_{q f_j} in the
// neighbour's row (-1 if boundary); pos_m: value index of M_{f_j f_j} (-1 = no face).
// ---------------------------------------------------------------------------
__global__ void k_schur_ell(int E, const int32_t* __restrict__ pos_b, const int32_t* __restrict__ pos_qb,
                            const int32_t* __restrict__ pos_m, const double* __restrict__ vals,
                            double* __restrict__ sdiag, double* __restrict__ soff) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= E) return;
  double d = 0;
  for (int j = 0; j < 4; ++j) {
    const int ib = pos_b[4 * p + j], im = pos_m[4 * p + j], iq = pos_qb[4 * p + j];
    double o = 0;
    if (ib >= 0 && im >= 0) {
      const double b = vals[ib], m = vals[im];
      d += b * b / m;
      if (iq >= 0) o = b * vals[iq] / m;
    }
    soff[4 * p + j] = o;
  }
  sdiag[p] = d;
}

// Node features x[p][0..7] and couplings a[p][j] = -S_pq / sqrt(S_pp S_qq).
__global__ void k_features(int E, const int32_t* __restrict__ nbr, const double* __restrict__ sdiag,
                           const double* __restrict__ soff, float* __restrict__ a_out,
                           float* __restrict__ x_out, float* __restrict__ sd32,
                           float* __restrict__ so32) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= E) return;
  const double dp = sdiag[p];
  double sa = 0, mx = 0, mn = 0, rs = dp, lsum = 0, lmax = 0;
  int cnt = 0;
  for (int j = 0; j < 4; ++j) {
    const int q = nbr[4 * p + j];
    const double o = soff[4 * p + j];
    rs += o;
    so32[4 * p + j] = float(o);
    if (q < 0) {
      a_out[4 * p + j] = 0.f;
      continue;
    }
    const double dq = sdiag[q];
    const double a = -o / sqrt(dp * dq);
    a_out[4 * p + j] = float(a);
    const double lr = fabs(log(dq / dp));
    sa += a;
    mx = (cnt == 0 || a > mx) ? a : mx;
    mn = (cnt == 0 || a < mn) ? a : mn;
    lsum += lr;
    lmax = lr > lmax ? lr : lmax;
    ++cnt;
  }
  // max over all four slots includes the zeros of boundary slots
  if (cnt < 4 && mx < 0) mx = 0;
  float* x = x_out + (int64_t)kIn * p;
  x[0] = float(sa);
  x[1] = float(mx);
  x[2] = float(cnt > 0 ? mn : 0.0);
  x[3] = float(rs / dp);
  x[4] = float(cnt > 0 ? lsum / cnt : 0.0);
  x[5] = float(lmax);
  x[6] = float(cnt / 4.0);
  x[7] = 1.f;
  sd32[p] = float(dp);
}

// Input layer: z0 = x W0 + b0, h0 = bf16(relu(z0)). One thread per (node, channel).
__global__ void k_gnn_input(int E, const float* __restrict__ x, const uint16_t* __restrict__ th16,
                            const float* __restrict__ th32, float* __restrict__ z0,
                            uint16_t* __restrict__ h0) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)E * kW) return;
  const int p = int(i / kW), c = int(i % kW);
  float acc = 0.f;
  for (int k = 0; k < kIn; ++k) acc += x[(int64_t)kIn * p + k] * bf2f(th16[kOffW0 + k * kW + c]);
  acc += th32[kOffB0 + c];
  z0[i] = acc;
  h0[i] = f2bf(fmaxf(acc, 0.f));
}

// One message-passing layer:
//   m_p = bf16(sum_j a_pj h_q);  z = [h_p, m_p] [Wh; Wm] + b;  h'_p = bf16(h_p + relu(z)).
// Block = 4 waves, tile = 64 nodes; wave w computes nodes 16w..16w+15 x 64 channels
// with 4 x 8 MFMA 16x16x16 (bf16 in, fp32 accumulate).
constexpr int kTile = 64;
constexpr int kKA = 2 * kW;     // 128
constexpr int kLdA = kKA + 8;   // padded LDS row (bf16 elements)

__global__ __launch_bounds__(256) void k_gnn_layer(int E, const int32_t* __restrict__ nbr,
                                                   const float* __restrict__ a,
                                                   const uint16_t* __restrict__ th16,
                                                   const float* __restrict__ th32, int layer,
                                                   const uint16_t* __restrict__ h_in,
                                                   uint16_t* __restrict__ m_out,
                                                   float* __restrict__ z_out,
                                                   uint16_t* __restrict__ h_out) {
  __shared__ uint16_t sA[kTile * kLdA];
  __shared__ uint16_t sBt[kW * kLdA];  // transposed weights: sBt[c][k]
  const int tid = threadIdx.x;
  const int base = kLayerBase + layer * kLayerSize;
  // weights: W[k][c] for k < 64 from Wh, k >= 64 from Wm
  for (int i = tid; i < kKA * kW; i += 256) {
    const int k = i / kW, c = i % kW;
    sBt[c * kLdA + k] = th16[base + i];  // Wh block then Wm block are contiguous
  }
  const int64_t tile0 = (int64_t)blockIdx.x * kTile;
  // gather: thread -> (node = tid / 4, 16-channel chunk)
  {
    const int ln = tid >> 2, ch = (tid & 3) * 16;
    const int64_t p = tile0 + ln;
    float hv[16], mv[16];
    for (int c = 0; c < 16; ++c) hv[c] = mv[c] = 0.f;
    if (p < E) {
      for (int c = 0; c < 16; ++c) hv[c] = bf2f(h_in[p * kW + ch + c]);
      for (int j = 0; j < 4; ++j) {
        const int q = nbr[4 * p + j];
        if (q < 0) continue;
        const float w = a[4 * p + j];
        for (int c = 0; c < 16; ++c) mv[c] += w * bf2f(h_in[(int64_t)q * kW + ch + c]);
      }
    }
    for (int c = 0; c < 16; ++c) {
      const uint16_t mb = f2bf(mv[c]);
      sA[ln * kLdA + ch + c] = f2bf(hv[c]);
      sA[ln * kLdA + kW + ch + c] = mb;
      if (p < E) m_out[p * kW + ch + c] = mb;
    }
  }
  __syncthreads();
  const int wave = tid >> 6, lane = tid & 63;
  const int row = lane & 15, kq = (lane >> 4) * 4;
  float4_t acc[4];
  for (int nt = 0; nt < 4; ++nt) acc[nt] = float4_t{0.f, 0.f, 0.f, 0.f};
  for (int kk = 0; kk < kKA / 16; ++kk) {
    short4_t av;
    const uint16_t* ap = &sA[(wave * 16 + row) * kLdA + kk * 16 + kq];
    av[0] = (short)ap[0];
    av[1] = (short)ap[1];
    av[2] = (short)ap[2];
    av[3] = (short)ap[3];
    for (int nt = 0; nt < 4; ++nt) {
      const uint16_t* bp = &sBt[(nt * 16 + row) * kLdA + kk * 16 + kq];
      short4_t bv;
      bv[0] = (short)bp[0];
      bv[1] = (short)bp[1];
      bv[2] = (short)bp[2];
      bv[3] = (short)bp[3];
      acc[nt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc[nt], 0, 0, 0);
    }
  }
  // epilogue: lane holds D[kq + i][row] of each 16x16 tile
  for (int nt = 0; nt < 4; ++nt) {
    const int c = nt * 16 + row;
    const float bias = th32[base + 2 * kW * kW + c];
    for (int i = 0; i < 4; ++i) {
      const int ln = wave * 16 + kq + i;
      const int64_t p = tile0 + ln;
      if (p >= E) continue;
      const float z = acc[nt][i] + bias;
      z_out[p * kW + c] = z;
      h_out[p * kW + c] = f2bf(bf2f(sA[ln * kLdA + c]) + fmaxf(z, 0.f));
    }
  }
}

// Heads: o_p = h_p.ud + cd, o_pj = a_pj ((h_p + h_q).ue + ce);
// Wd = (omega0 + o_p) / S_pp, Wo_pj = o_pj / sqrt(S_pp S_qq).
__global__ void k_gnn_heads(int E, const int32_t* __restrict__ nbr, const float* __restrict__ a,
                            const uint16_t* __restrict__ th16, const float* __restrict__ th32,
                            const uint16_t* __restrict__ h, const double* __restrict__ sdiag,
                            float* __restrict__ Wd, float* __restrict__ Wo) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= E) return;
  const uint16_t* hp = h + (int64_t)p * kW;
  float od = 0.f;
  for (int c = 0; c < kW; ++c) od += bf2f(hp[c]) * bf2f(th16[kOffUd + c]);
  od += th32[kOffCd];
  const double dp = sdiag[p];
  Wd[p] = float((kOmega0 + od) / dp);
  for (int j = 0; j < 4; ++j) {
    const int q = nbr[4 * p + j];
    if (q < 0) {
      Wo[4 * p + j] = 0.f;
      continue;
    }
    const uint16_t* hq = h + (int64_t)q * kW;
    float s = 0.f;
    for (int c = 0; c < kW; ++c) s += (bf2f(hp[c]) + bf2f(hq[c])) * bf2f(th16[kOffUe + c]);
    const float oe = a[4 * p + j] * (s + th32[kOffCe]);
    Wo[4 * p + j] = float(oe / sqrt(dp * sdiag[q]));
  }
}

// ---------------------------------------------------------------------------
// Backward (FP32). Weight-gradient partials per block, summed in block order.
// ---------------------------------------------------------------------------
constexpr int kBwdBlocks = 256;
constexpr int kHeadPartial = 2 * kW + 2;            // ud, ue, cd, ce
constexpr int kLayerPartial = 2 * kW * kW + kW;     // Wh, Wm, b
constexpr int kInPartial = kIn * kW + kW;           // W0, b0

// dh_p = gN_p ud + (sum_j a_pj gE_pj) ue; partials of ud, ue, cd, ce.
__global__ __launch_bounds__(256) void k_heads_bwd(int E, const int32_t* __restrict__ nbr,
                                                   const float* __restrict__ a,
                                                   const uint16_t* __restrict__ th16,
                                                   const uint16_t* __restrict__ h,
                                                   const double* __restrict__ sdiag,
                                                   const float* __restrict__ gam_d,
                                                   const float* __restrict__ gam_o,
                                                   float* __restrict__ dh,
                                                   float* __restrict__ partial) {
  __shared__ float sGn[256], sAg[256][4];
  __shared__ int sQ[256][4];
  const int tid = threadIdx.x;
  const int c = tid & 63, grp = tid >> 6;
  float pud = 0.f, pue = 0.f, pcd = 0.f, pce = 0.f;
  for (int64_t t0 = (int64_t)blockIdx.x * 256; t0 < E; t0 += (int64_t)gridDim.x * 256) {
    const int64_t p = t0 + tid;
    float gn = 0.f, agsum = 0.f;
    float ag[4] = {0.f, 0.f, 0.f, 0.f};
    int qq[4] = {-1, -1, -1, -1};
    if (p < E) {
      const double dp = sdiag[p];
      gn = float(gam_d[p] / dp);
      for (int j = 0; j < 4; ++j) {
        const int q = nbr[4 * p + j];
        qq[j] = q;
        if (q < 0) continue;
        const float ge = float(gam_o[4 * p + j] / sqrt(dp * sdiag[q]));
        ag[j] = a[4 * p + j] * ge;
        agsum += ag[j];
      }
    }
    sGn[tid] = gn;
    for (int j = 0; j < 4; ++j) {
      sAg[tid][j] = ag[j];
      sQ[tid][j] = qq[j];
    }
    __syncthreads();
    // channel c, nodes grp, grp+4, ...
    const float udc = bf2f(th16[kOffUd + c]), uec = bf2f(th16[kOffUe + c]);
    for (int ln = grp; ln < 256; ln += 4) {
      const int64_t pp = t0 + ln;
      if (pp >= E) break;
      const float hp = bf2f(h[pp * kW + c]);
      float asum = 0.f;
      for (int j = 0; j < 4; ++j) {
        const int q = sQ[ln][j];
        if (q < 0) continue;
        asum += sAg[ln][j];
        pue += 0.5f * sAg[ln][j] * (hp + bf2f(h[(int64_t)q * kW + c]));
      }
      pud += sGn[ln] * hp;
      dh[pp * kW + c] = sGn[ln] * udc + asum * uec;
      if (c == 0) {
        pcd += sGn[ln];
        pce += 0.5f * asum;
      }
    }
    __syncthreads();
  }
  // reduce the 4 groups in fixed order through LDS
  __shared__ float red[4][2 * kW + 2];
  red[grp][c] = pud;
  red[grp][kW + c] = pue;
  if (c == 0) {
    red[grp][2 * kW] = pcd;
    red[grp][2 * kW + 1] = pce;
  }
  __syncthreads();
  for (int i = tid; i < kHeadPartial; i += 256)
    partial[(int64_t)blockIdx.x * kHeadPartial + i] = red[0][i] + red[1][i] + red[2][i] + red[3][i];
}

// Layer backward, part 1 (per tile of 64 nodes):
//   dz = dh * (z > 0); partial dWh += h^T dz, dWm += m^T dz, db += dz;
//   dh <- dh + dz Wh^T (in place); dm = dz Wm^T.
__global__ __launch_bounds__(256) void k_layer_bwd(int E, const uint16_t* __restrict__ th16, int layer,
                                                   const uint16_t* __restrict__ h,
                                                   const uint16_t* __restrict__ m,
                                                   const float* __restrict__ z,
                                                   float* __restrict__ dh, float* __restrict__ dm,
                                                   float* __restrict__ partial) {
  __shared__ float sIn[kTile][kKA + 1];
  __shared__ float sDz[kTile][kW + 1];
  __shared__ float sWt[kKA][kW + 1];  // W[k][c]
  const int tid = threadIdx.x;
  const int base = kLayerBase + layer * kLayerSize;
  for (int i = tid; i < kKA * kW; i += 256) sWt[i / kW][i % kW] = bf2f(th16[base + i]);
  const int c = tid & 63, g = tid >> 6;  // grad ownership: channel c, k in [32g, 32g+32)
  float accW[32];
  for (int i = 0; i < 32; ++i) accW[i] = 0.f;
  float accB = 0.f;
  const int64_t n_tiles = (E + kTile - 1) / kTile;
  for (int64_t t = blockIdx.x; t < n_tiles; t += gridDim.x) {
    const int64_t p0 = t * kTile;
    __syncthreads();
    for (int i = tid; i < kTile * kW; i += 256) {
      const int ln = i / kW, cc = i % kW;
      const int64_t p = p0 + ln;
      float hv = 0.f, mv = 0.f, dz = 0.f;
      if (p < E) {
        hv = bf2f(h[p * kW + cc]);
        mv = bf2f(m[p * kW + cc]);
        dz = z[p * kW + cc] > 0.f ? dh[p * kW + cc] : 0.f;
      }
      sIn[ln][cc] = hv;
      sIn[ln][kW + cc] = mv;
      sDz[ln][cc] = dz;
    }
    __syncthreads();
    for (int ln = 0; ln < kTile; ++ln) {
      const float d = sDz[ln][c];
      for (int i = 0; i < 32; ++i) accW[i] += sIn[ln][32 * g + i] * d;
      if (g == 0) accB += d;
    }
    // d_in[ln][k] = sum_c dz[ln][c] W[k][c]; thread -> node ln = tid / 4, k chunk of 32
    {
      const int ln = tid >> 2, k0 = (tid & 3) * 32;
      const int64_t p = p0 + ln;
      float out[32];
      for (int i = 0; i < 32; ++i) out[i] = 0.f;
      for (int cc = 0; cc < kW; ++cc) {
        const float d = sDz[ln][cc];
        for (int i = 0; i < 32; ++i) out[i] += d * sWt[k0 + i][cc];
      }
      if (p < E) {
        for (int i = 0; i < 32; ++i) {
          const int k = k0 + i;
          if (k < kW)
            dh[p * kW + k] += out[i];
          else
            dm[p * kW + (k - kW)] = out[i];
        }
      }
    }
  }
  float* part = partial + (int64_t)blockIdx.x * kLayerPartial;
  for (int i = 0; i < 32; ++i) part[(32 * g + i) * kW + c] = accW[i];
  if (g == 0) part[2 * kW * kW + c] = accB;
}

// Layer backward, part 2: dh_p += sum_j a_pj dm_q (transpose of the message sum; a symmetric).
__global__ void k_msg_bwd(int E, const int32_t* __restrict__ nbr, const float* __restrict__ a,
                          const float* __restrict__ dm, float* __restrict__ dh) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)E * kW) return;
  const int64_t p = i / kW;
  const int c = int(i % kW);
  float s = 0.f;
  for (int j = 0; j < 4; ++j) {
    const int q = nbr[4 * p + j];
    if (q < 0) continue;
    s += a[4 * p + j] * dm[(int64_t)q * kW + c];
  }
  dh[i] += s;
}

// Input layer backward: dz0 = dh * (z0 > 0); partial dW0 += x^T dz0, db0 += dz0.
__global__ __launch_bounds__(256) void k_input_bwd(int E, const float* __restrict__ x,
                                                   const float* __restrict__ z0,
                                                   const float* __restrict__ dh,
                                                   float* __restrict__ partial) {
  const int tid = threadIdx.x, c = tid & 63, g = tid >> 6;
  float acc[kIn + 1];
  for (int i = 0; i <= kIn; ++i) acc[i] = 0.f;
  for (int64_t p = (int64_t)blockIdx.x * 4 + g; p < E; p += (int64_t)gridDim.x * 4) {
    const float d = z0[p * kW + c] > 0.f ? dh[p * kW + c] : 0.f;
    for (int k = 0; k < kIn; ++k) acc[k] += x[p * kIn + k] * d;
    acc[kIn] += d;
  }
  __shared__ float red[4][kIn + 1][kW];
  for (int k = 0; k <= kIn; ++k) red[g][k][c] = acc[k];
  __syncthreads();
  for (int i = tid; i < kInPartial; i += 256) {
    const int k = i / kW, cc = i % kW;
    partial[(int64_t)blockIdx.x * kInPartial + i] = red[0][k][cc] + red[1][k][cc] + red[2][k][cc] + red[3][k][cc];
  }
}

// out[i] = sum_b partial[b][i] (fixed block order)
__global__ void k_reduce_partials(int nblocks, int stride, int n, const float* __restrict__ partial,
                                  float* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float s = 0.f;
  for (int b = 0; b < nblocks; ++b) s += partial[(int64_t)b * stride + i];
  out[i] = s;
}

// ---------------------------------------------------------------------------
// Loss-side kernels. Sample blocks are stored [ns][E].
// ---------------------------------------------------------------------------
// Y = beta * Y0 + alpha * (diag + ELL) X
__global__ void k_ell_spmm(int E, int ns, const int32_t* __restrict__ nbr, const float* __restrict__ d,
                           const float* __restrict__ off, float alpha, const float* __restrict__ X,
                           float beta, const float* __restrict__ Y0, float* __restrict__ Y) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)E * ns) return;
  const int64_t s = i / E, p = i % E;
  const float* x = X + s * E;
  float v = d[p] * x[p];
  for (int j = 0; j < 4; ++j) {
    const int q = nbr[4 * p + j];
    if (q >= 0) v += off[4 * p + j] * x[q];
  }
  float out = alpha * v;
  if (Y0) out += beta * Y0[i];
  Y[i] = out;
}

// gamma_d[p] = sum_s w_s Su_p Se_p; gamma_o[p][j] = sum_s w_s (Su_p Se_q + Su_q Se_p)
__global__ void k_smoother_grad(int E, int ns, const int32_t* __restrict__ nbr,
                                const float* __restrict__ Su, const float* __restrict__ Se,
                                const float* __restrict__ w, float* __restrict__ gd,
                                float* __restrict__ go) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= E) return;
  float d = 0.f, o[4] = {0.f, 0.f, 0.f, 0.f};
  int q[4];
  for (int j = 0; j < 4; ++j) q[j] = nbr[4 * p + j];
  for (int s = 0; s < ns; ++s) {
    const float* su = Su + (int64_t)s * E;
    const float* se = Se + (int64_t)s * E;
    const float ws = w[s], sup = su[p] * ws, sep = se[p] * ws;
    d += sup * se[p];
    for (int j = 0; j < 4; ++j)
      if (q[j] >= 0) o[j] += sup * se[q[j]] + sep * su[q[j]];
  }
  gd[p] = d;
  for (int j = 0; j < 4; ++j) go[4 * p + j] = o[j];
}

// C[s][a] = sum over members p of aggregate a (member order) of X[s][p]
__global__ void k_agg_sum(int n_agg, int ns, int E, const int32_t* __restrict__ agg_ptr,
                          const int32_t* __restrict__ agg_mem, const float* __restrict__ X,
                          float* __restrict__ C) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)n_agg * ns) return;
  const int64_t s = i / n_agg;
  const int ag = int(i % n_agg);
  float acc = 0.f;
  for (int k = agg_ptr[ag]; k < agg_ptr[ag + 1]; ++k) acc += X[s * E + agg_mem[k]];
  C[i] = acc;
}

// Dense coarse operator row a: Ac[a][agg(q)] += S_pq for members p of a (FP64, member order).
// Ac must be zero on entry.
__global__ void k_coarse_dense(int n_agg, const int32_t* __restrict__ agg_ptr,
                               const int32_t* __restrict__ agg_mem, const int32_t* __restrict__ agg_of,
                               const int32_t* __restrict__ nbr, const double* __restrict__ sdiag,
                               const double* __restrict__ soff, double* __restrict__ Ac) {
  const int ag = blockIdx.x * blockDim.x + threadIdx.x;
  if (ag >= n_agg) return;
  double* row = Ac + (int64_t)ag * n_agg;
  for (int k = agg_ptr[ag]; k < agg_ptr[ag + 1]; ++k) {
    const int p = agg_mem[k];
    row[ag] += sdiag[p];
    for (int j = 0; j < 4; ++j) {
      const int q = nbr[4 * p + j];
      if (q >= 0) row[agg_of[q]] += soff[4 * p + j];
    }
  }
}

// ---------------------------------------------------------------------------
// Two-grid cycle z = TG(r) on S, with the flux block z_u = dMinv r_u, and an
// overall sign on the pressure part (psign = -1 gives a real spectrum for the
// unsymmetrised [[M, -B^T], [B, 0]]).
// ---------------------------------------------------------------------------
__device__ __forceinline__ float ell_row(int p, const int32_t* __restrict__ nbr, const float* __restrict__ d,
                                         const float* __restrict__ off, const float* __restrict__ v) {
  float s = d[p] * v[p];
  for (int j = 0; j < 4; ++j) {
    const int q = nbr[4 * p + j];
    if (q >= 0) s += off[4 * p + j] * v[q];
  }
  return s;
}

// x1 = W r (all cells)
__global__ void k_tg_pre(int E, const int32_t* __restrict__ nbr, const float* __restrict__ Wd,
                         const float* __restrict__ Wo, const float* __restrict__ r,
                         float* __restrict__ x1) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= E) return;
  x1[p] = ell_row(p, nbr, Wd, Wo, r);
}

// Blocks [0, cell_blocks): x1 = W r on the cells, as k_tg_pre. Blocks after them: the
// flux rows z_u = C r_u, C in slot-major ELL form (width slots per row, one thread per
// row, coalesced over rows), scaled by 1 / *dscale like the pressure rows.
__global__ __launch_bounds__(256) void k_tg_pre_flux(int E, int cell_blocks, const int32_t* __restrict__ nbr,
                                                     const float* __restrict__ Wd, const float* __restrict__ Wo,
                                                     const float* __restrict__ rp, float* __restrict__ x1, int n_u,
                                                     int width, const int32_t* __restrict__ fcol,
                                                     const float* __restrict__ fcoef, const float* __restrict__ r,
                                                     const double* __restrict__ dscale, float* __restrict__ z) {
  if ((int)blockIdx.x < cell_blocks) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < E) x1[p] = ell_row(p, nbr, Wd, Wo, rp);
    return;
  }
  const int i = (blockIdx.x - cell_blocks) * blockDim.x + threadIdx.x;
  if (i >= n_u) return;
  float acc = 0.f;
  for (int s = 0; s < width; ++s) {
    const int64_t k = (int64_t)s * n_u + i;
    acc += fcoef[k] * r[fcol[k]];
  }
  float sc = 1.f;
  if (dscale) {
    const double d = *dscale;
    sc = d > 0 ? (float)(1.0 / d) : 0.f;
  }
  z[i] = sc * acc;
}

__global__ void k_flux_bound_init(float* __restrict__ bound) {
  if (threadIdx.x == 0) bound[0] = 0.f;
}

// bound = max_i sum_j |M_ij| / M_ii over the flux-row ELL slots (the row-sum norm of
// D^-1 M, an upper bound of its spectral radius); positive floats order as their bits
__global__ __launch_bounds__(256) void k_flux_bound(int n_u, int width, const int32_t* __restrict__ pos,
                                                    const int32_t* __restrict__ diag_pos,
                                                    const double* __restrict__ vals, float* __restrict__ bound) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  float b = 0.f;
  if (i < n_u) {
    double s = 0.0;
    for (int k = 0; k < width; ++k) {
      const int q = pos[(int64_t)k * n_u + i];
      if (q >= 0) s += fabs(vals[q]);
    }
    b = (float)(s / vals[diag_pos[i]]);
  }
  for (int off = 32; off > 0; off >>= 1) b = fmaxf(b, __shfl_xor(b, off, 64));
  if ((threadIdx.x & 63) == 0 && b > 0.f)
    atomicMax(reinterpret_cast<unsigned int*>(bound), __float_as_uint(b));
}

// C = 2 omega D^-1 - omega^2 D^-1 M D^-1 on the flux-row ELL slots (FP64 arithmetic,
// FP32 store); one thread per (slot, row). omega <= 0: omega = 2 / (0.5 + *bound).
__global__ void k_flux_sweep_coef(int n_u, int width, const int32_t* __restrict__ pos,
                                  const int32_t* __restrict__ fcol, const int32_t* __restrict__ diag_pos,
                                  const double* __restrict__ vals, double omega, const float* __restrict__ bound,
                                  float* __restrict__ coef) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (int64_t)width * n_u) return;
  const int i = (int)(k % n_u);
  const int q = pos[k];
  if (q < 0) {
    coef[k] = 0.f;
    return;
  }
  const double w = omega > 0.0 ? omega : 2.0 / (0.5 + (double)bound[0]);
  const int j = fcol[k];
  const double di = 1.0 / vals[diag_pos[i]], dj = 1.0 / vals[diag_pos[j]];
  double c = -w * w * di * vals[q] * dj;
  if (j == i) c += 2.0 * w * di;
  coef[k] = (float)c;
}

// c[a] = sum_{p in a} (r - S x1)_p   (one wave per aggregate)
__global__ __launch_bounds__(256) void k_tg_restrict(int n_agg, const int32_t* __restrict__ agg_ptr,
                                                     const int32_t* __restrict__ agg_mem,
                                                     const int32_t* __restrict__ nbr,
                                                     const float* __restrict__ sd,
                                                     const float* __restrict__ so,
                                                     const float* __restrict__ r,
                                                     const float* __restrict__ x1,
                                                     float* __restrict__ c) {
  const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (wave >= n_agg) return;
  float acc = 0.f;
  for (int k = agg_ptr[wave] + lane; k < agg_ptr[wave + 1]; k += 64) {
    const int p = agg_mem[k];
    acc += r[p] - ell_row(p, nbr, sd, so, x1);
  }
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if (lane == 0) c[wave] = acc;
}

// y = Ainv c  (one wave per row, dense row-major FP32; 16-byte loads, 4 in flight per lane)
__global__ __launch_bounds__(256) void k_gemv(int n, const float* __restrict__ A,
                                              const float* __restrict__ x, float* __restrict__ y) {
  const int row = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (row >= n) return;
  const float* a = A + (int64_t)row * n;
  float acc = 0.f;
  if ((n & 3) == 0) {
    const float4* a4 = reinterpret_cast<const float4*>(a);
    const float4* x4 = reinterpret_cast<const float4*>(x);
    const int n4 = n >> 2;
    int k = lane;
    for (; k + 192 < n4; k += 256) {
      const float4 p0 = a4[k], p1 = a4[k + 64], p2 = a4[k + 128], p3 = a4[k + 192];
      const float4 q0 = x4[k], q1 = x4[k + 64], q2 = x4[k + 128], q3 = x4[k + 192];
      acc += p0.x * q0.x + p0.y * q0.y + p0.z * q0.z + p0.w * q0.w;
      acc += p1.x * q1.x + p1.y * q1.y + p1.z * q1.z + p1.w * q1.w;
      acc += p2.x * q2.x + p2.y * q2.y + p2.z * q2.z + p2.w * q2.w;
      acc += p3.x * q3.x + p3.y * q3.y + p3.z * q3.z + p3.w * q3.w;
    }
    for (; k < n4; k += 64) {
      const float4 p = a4[k], q = x4[k];
      acc += p.x * q.x + p.y * q.y + p.z * q.z + p.w * q.w;
    }
  } else {
    for (int k = lane; k < n; k += 64) acc += a[k] * x[k];
  }
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if (lane == 0) y[row] = acc;
}

// In-place inverse of an SPD block (n <= 128, no pivoting), row stride lda, by block
// Gauss-Jordan over 16 x 16 pivot blocks. 512 threads = 8 waves; wave w holds block
// row w (rows 16w .. 16w + 15 of the block padded to 128 with the identity) as eight
// 16 x 16 accumulator tiles of v_mfma_f32_16x16x4_f32 (lane l: rows 4 (l >> 4) + i,
// column l & 15). Per pivot block k the pivot block row and column go through LDS
// (double-buffered, one barrier per step); every wave inverts the 16 x 16 pivot
// block P itself (scalar Gauss-Jordan across the wave, lane l holding row l >> 2,
// columns 4 (l & 3) .. + 3); the pivot wave forms A[k, j] <- P^-1 A[k, j] and
// A[k, k] <- P^-1, every other wave X = A[I, k] P^-1, A[I, j] -= X A[k, j] and
// A[I, k] <- -X. Steps past the last pivot block holding a row < n are skipped: those
// blocks are identity with no coupling.
constexpr int kInvN = 128;
constexpr int kInvRS = 144;  // pivot-row LDS stride (16 mod 32 banks)
constexpr int kInvCS = 18;   // pivot-column and per-wave scratch stride

__device__ __forceinline__ float4_t mfma_f32_4(float a, float b, float4_t c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

__global__ __launch_bounds__(512) void k_spd_inverse_small(int n, float* __restrict__ A, int lda) {
  __shared__ float rowp[2][16][kInvRS];
  __shared__ float colp[2][kInvN][kInvCS];
  __shared__ float pinv[8][16][kInvCS];
  __shared__ float xs[8][16][kInvCS];
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63, lr = l & 15, lq = l >> 4;
  float4_t acc[8];
#pragma unroll
  for (int J = 0; J < 8; ++J)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int r = 16 * w + 4 * lq + i, c = 16 * J + lr;
      acc[J][i] = (r < n && c < n) ? A[(int64_t)r * lda + c] : (r == c ? 1.f : 0.f);
    }
  const int nb = (n + 15) >> 4;
  float(*pv)[kInvCS] = pinv[w];
  float(*xv)[kInvCS] = xs[w];
  const int pr = l >> 2, pc = 4 * (l & 3);
#pragma unroll
  for (int kb = 0; kb < 8; ++kb) {
    if (kb >= nb) break;
    const int buf = kb & 1;
    if (w == kb) {
#pragma unroll
      for (int J = 0; J < 8; ++J)
#pragma unroll
        for (int i = 0; i < 4; ++i) rowp[buf][4 * lq + i][16 * J + lr] = acc[J][i];
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) colp[buf][16 * w + 4 * lq + i][lr] = acc[kb][i];
    __syncthreads();
    float p[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) p[t] = rowp[buf][pr][16 * kb + pc + t];
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      float prow[4];
#pragma unroll
      for (int t = 0; t < 4; ++t) prow[t] = __shfl(p[t], 4 * k + (l & 3), 64);
      const float f = __shfl(p[k & 3], 4 * pr + (k >> 2), 64);
      const float ip = 1.f / __shfl(p[k & 3], 4 * k + (k >> 2), 64);
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const bool ck = (pc + t) == k;
        const float pj = (ck ? 1.f : prow[t]) * ip;
        p[t] = (pr == k) ? pj : (ck ? 0.f : p[t]) - f * pj;
      }
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) pv[pr][pc + t] = p[t];
    __builtin_amdgcn_wave_barrier();
    if (w == kb) {
      float4_t nacc[8];
#pragma unroll
      for (int J = 0; J < 8; ++J) nacc[J] = float4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const float a = pv[lr][4 * s + lq];
#pragma unroll
        for (int J = 0; J < 8; ++J)
          if (J != kb) nacc[J] = mfma_f32_4(a, rowp[buf][4 * s + lq][16 * J + lr], nacc[J]);
      }
#pragma unroll
      for (int J = 0; J < 8; ++J)
        if (J != kb) acc[J] = nacc[J];
#pragma unroll
      for (int i = 0; i < 4; ++i) acc[kb][i] = pv[4 * lq + i][lr];
    } else {
      float4_t x = float4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int s = 0; s < 4; ++s) x = mfma_f32_4(colp[buf][16 * w + lr][4 * s + lq], pv[4 * s + lq][lr], x);
#pragma unroll
      for (int i = 0; i < 4; ++i) xv[4 * lq + i][lr] = x[i];
      __builtin_amdgcn_wave_barrier();
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        const float a = -xv[lr][4 * s + lq];
#pragma unroll
        for (int J = 0; J < 8; ++J)
          if (J != kb) acc[J] = mfma_f32_4(a, rowp[buf][4 * s + lq][16 * J + lr], acc[J]);
      }
#pragma unroll
      for (int i = 0; i < 4; ++i) acc[kb][i] = -x[i];
    }
  }
#pragma unroll
  for (int J = 0; J < 8; ++J)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int r = 16 * w + 4 * lq + i, c = 16 * J + lr;
      if (r < n && c < n) A[(int64_t)r * lda + c] = acc[J][i];
    }
}

// C = beta C + alpha op(A) op(B), and C2 = C^T when C2 is not null (FP32, row-major
// with leading dimensions; op(A) is M x K: A[i*lda + k], or A[k*lda + i] when transA;
// op(B) is K x N: B[k*ldb + j], or B[j*ldb + k] when transB). TILE x TILE output
// tiles (64, or 32 when 64 x 64 tiles would leave CUs idle), 4 waves each holding a
// (TILE/2)^2 block of 16 x 16 v_mfma_f32_16x16x4_f32 tiles. K is taken in panels of
// NP chunks of 32: every load of a panel is issued before its first product, so one
// memory latency covers the panel (the products of the coarse inversion have K <= 256:
// one or two panels); the panel moves through LDS chunk by chunk (double-buffered, one
// barrier per chunk). 64-wide tiles use 4-chunk panels so that two workgroups fit on a
// CU and overlap one's loads with the other's products. Beta = 0 does not read C.
constexpr int kGK = 32;

template <int TILE, int kGP, bool transA, bool transB>
__global__ __launch_bounds__(256) void k_sgemm(int M, int N, int K, float alpha, const float* __restrict__ A,
                                               int lda, const float* __restrict__ B, int ldb, float beta,
                                               float* __restrict__ C, int ldc, float* __restrict__ C2, int ldc2) {
  constexpr int EPT = TILE * kGK / 256;  // elements of each operand per thread and chunk
  constexpr int WT = TILE / 32;          // 16 x 16 tiles per wave side
  __shared__ float sA[2][TILE][kGK + 2];
  __shared__ float sB[2][kGK][TILE + 16];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, lr = lane & 15, lq = lane >> 4;
  const int i0 = blockIdx.y * TILE, j0 = blockIdx.x * TILE;
  const int wr = (wave >> 1) * (TILE / 2), wc = (wave & 1) * (TILE / 2);
  float4_t acc[WT][WT];
#pragma unroll
  for (int a = 0; a < WT; ++a)
#pragma unroll
    for (int b = 0; b < WT; ++b) acc[a][b] = float4_t{0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < K; k0 += kGK * kGP) {
    const int nch = min(kGP, (K - k0 + kGK - 1) / kGK);
    float ra[kGP][EPT], rb[kGP][EPT];
#pragma unroll
    for (int c = 0; c < kGP; ++c) {
      if (c < nch) {
#pragma unroll
        for (int u = 0; u < EPT; ++u) {
          const int e = tid + 256 * u;
          const int ar = transA ? e % TILE : e / kGK, ak = transA ? e / TILE : e % kGK;
          const int gi = i0 + ar, gka = k0 + c * kGK + ak;
          ra[c][u] = (gi < M && gka < K) ? (transA ? A[(int64_t)gka * lda + gi] : A[(int64_t)gi * lda + gka]) : 0.f;
          const int bk = transB ? e % kGK : e / TILE, bj = transB ? e / kGK : e % TILE;
          const int gj = j0 + bj, gkb = k0 + c * kGK + bk;
          rb[c][u] = (gj < N && gkb < K) ? (transB ? B[(int64_t)gj * ldb + gkb] : B[(int64_t)gkb * ldb + gj]) : 0.f;
        }
      }
    }
#pragma unroll
    for (int c = 0; c < kGP; ++c) {
      if (c < nch) {
        const int buf = c & 1;
#pragma unroll
        for (int u = 0; u < EPT; ++u) {
          const int e = tid + 256 * u;
          const int ar = transA ? e % TILE : e / kGK, ak = transA ? e / TILE : e % kGK;
          sA[buf][ar][ak] = ra[c][u];
          const int bk = transB ? e % kGK : e / TILE, bj = transB ? e / kGK : e % TILE;
          sB[buf][bk][bj] = rb[c][u];
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < kGK; ks += 4) {
          const int kl = ks + lq;
          float av[WT], bv[WT];
#pragma unroll
          for (int t = 0; t < WT; ++t) {
            av[t] = sA[buf][wr + 16 * t + lr][kl];
            bv[t] = sB[buf][kl][wc + 16 * t + lr];
          }
#pragma unroll
          for (int a = 0; a < WT; ++a)
#pragma unroll
            for (int b = 0; b < WT; ++b) acc[a][b] = mfma_f32_4(av[a], bv[b], acc[a][b]);
        }
      }
    }
    __syncthreads();
  }
  // lane holds D[4 (lane >> 4) + i][lane & 15] of each 16 x 16 tile
#pragma unroll
  for (int a = 0; a < WT; ++a)
#pragma unroll
    for (int b = 0; b < WT; ++b) {
      const int gj = j0 + wc + 16 * b + lr;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int gi = i0 + wr + 16 * a + 4 * lq + i;
        if (gi < M && gj < N) {
          float* cp = C + (int64_t)gi * ldc + gj;
          const float v = alpha * acc[a][b][i] + (beta == 0.f ? 0.f : beta * *cp);
          *cp = v;
          if (C2) C2[(int64_t)gj * ldc2 + gi] = v;
        }
      }
    }
}

template <int TILE, int NP>
void launch_sgemm(int ta, int tb, int M, int N, int K, float alpha, const float* A, int lda, const float* B, int ldb,
                  float beta, float* C, int ldc, float* C2, int ldc2, hipStream_t st) {
  const dim3 grid(grid_for(N, TILE), grid_for(M, TILE)), block(256);
  if (ta && tb)
    hipLaunchKernelGGL((k_sgemm<TILE, NP, true, true>), grid, block, 0, st, M, N, K, alpha, A, lda, B, ldb, beta, C,
                       ldc, C2, ldc2);
  else if (ta)
    hipLaunchKernelGGL((k_sgemm<TILE, NP, true, false>), grid, block, 0, st, M, N, K, alpha, A, lda, B, ldb, beta, C,
                       ldc, C2, ldc2);
  else if (tb)
    hipLaunchKernelGGL((k_sgemm<TILE, NP, false, true>), grid, block, 0, st, M, N, K, alpha, A, lda, B, ldb, beta, C,
                       ldc, C2, ldc2);
  else
    hipLaunchKernelGGL((k_sgemm<TILE, NP, false, false>), grid, block, 0, st, M, N, K, alpha, A, lda, B, ldb, beta,
                       C, ldc, C2, ldc2);
}

// x2 = x1 + y[agg]; t2 = r - S x2 (x2 of neighbours recomputed from x1 and y)
__global__ void k_tg_post1(int E, const int32_t* __restrict__ nbr, const int32_t* __restrict__ agg_of,
                           const float* __restrict__ sd, const float* __restrict__ so,
                           const float* __restrict__ r, const float* __restrict__ x1,
                           const float* __restrict__ y, float* __restrict__ x2,
                           float* __restrict__ t2) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= E) return;
  const float xp = x1[p] + y[agg_of[p]];
  float s = sd[p] * xp;
  for (int j = 0; j < 4; ++j) {
    const int q = nbr[4 * p + j];
    if (q >= 0) s += so[4 * p + j] * (x1[q] + y[agg_of[q]]);
  }
  x2[p] = xp;
  t2[p] = r[p] - s;
}

// z_p = psign (x2 + W t2) on pressure rows; z_u = dMinv r_u on flux rows; tail rows
// z_f = tail_dinv (r_f - sum_k tail_coef z_p(k)) with z of the two cells recomputed here
// (same arithmetic as their own rows), or tail_dinv r_f when tail_cells is null
// (dscale: optional device scalar; the output is divided by it, 0 gives 0)
// (rows from i0 on: i0 = n_u when the flux rows were written by k_tg_pre_flux, else 0)
__global__ void k_tg_post2(int i0, int n_u, int E, const int32_t* __restrict__ nbr, const float* __restrict__ Wd,
                           const float* __restrict__ Wo, const float* __restrict__ x2,
                           const float* __restrict__ t2, const float* __restrict__ dMinv,
                           const float* __restrict__ r, float psign, const double* __restrict__ dscale,
                           int n_tail, const float* __restrict__ tail_dinv, const int32_t* __restrict__ tail_cells,
                           const float* __restrict__ tail_coef, float* __restrict__ z) {
  const int64_t i = i0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  float sc = 1.f;
  if (dscale) {
    const double d = *dscale;
    sc = d > 0 ? (float)(1.0 / d) : 0.f;
  }
  if (i < n_u) {
    z[i] = sc * dMinv[i] * r[i];
    return;
  }
  const int64_t p = i - n_u;
  if (p >= E) {
    const int64_t f = p - E;
    if (f >= n_tail) return;
    if (!tail_cells) {
      z[i] = sc * tail_dinv[f] * r[i];
      return;
    }
    float v = sc * r[i];
    for (int k = 0; k < 2; ++k) {
      const int c = tail_cells[2 * f + k];
      if (c >= 0) v -= tail_coef[2 * f + k] * (sc * psign * (x2[c] + ell_row(c, nbr, Wd, Wo, t2)));
    }
    z[i] = tail_dinv[f] * v;
    return;
  }
  z[i] = sc * psign * (x2[p] + ell_row(int(p), nbr, Wd, Wo, t2));
}

// Block LU with a diagonal tail block, before the cycle: rq_p = r_p - sum_j c_pj r_f(p,j)
// over the slots of cell p that couple to a tail row.
__global__ void k_tg_tail_pre(int E, int n_u, const int32_t* __restrict__ slot_tail,
                              const float* __restrict__ slot_coef, const float* __restrict__ r,
                              float* __restrict__ rq) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= E) return;
  const float* rf = r + (int64_t)n_u + E;
  float v = r[(int64_t)n_u + p];
  for (int j = 0; j < 4; ++j) {
    const int f = slot_tail[4 * p + j];
    if (f >= 0) v -= slot_coef[4 * p + j] * rf[f];
  }
  rq[p] = v;
}

// Row-scaled FP16 copy of a dense row-major matrix: scale[r] = max |A[r][:]|,
// A16[r][k] = half(A[r][k] / scale[r]). One wave per row.
__global__ __launch_bounds__(256) void k_half_rows(int n, const float* __restrict__ A, uint16_t* __restrict__ A16,
                                                   float* __restrict__ scale) {
  const int row = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (row >= n) return;
  const float* a = A + (int64_t)row * n;
  float mx = 0.f;
  for (int k = lane; k < n; k += 64) mx = fmaxf(mx, fabsf(a[k]));
  for (int off = 32; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off, 64));
  const float inv = mx > 0.f ? 1.f / mx : 0.f;
  for (int k = lane; k < n; k += 64) {
    const _Float16 h = (_Float16)(a[k] * inv);
    A16[(int64_t)row * n + k] = __builtin_bit_cast(uint16_t, h);
  }
  if (lane == 0) scale[row] = mx;
}

// y = diag(scale) A16 c (FP16 storage, FP32 arithmetic); one wave per row, 16-byte loads.
__global__ __launch_bounds__(256) void k_gemv_h(int n, const uint16_t* __restrict__ A16,
                                                const float* __restrict__ scale, const float* __restrict__ x,
                                                float* __restrict__ y) {
  const int row = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = threadIdx.x & 63;
  if (row >= n) return;
  const uint16_t* a = A16 + (int64_t)row * n;
  float acc = 0.f;
  if ((n & 7) == 0) {
    const uint4* a8 = reinterpret_cast<const uint4*>(a);
    const float4* x4 = reinterpret_cast<const float4*>(x);
    for (int k = lane; k < (n >> 3); k += 64) {
      const uint4 u = a8[k];
      const float4 xa = x4[2 * k], xb = x4[2 * k + 1];
      const uint32_t w[4] = {u.x, u.y, u.z, u.w};
      const float xs[8] = {xa.x, xa.y, xa.z, xa.w, xb.x, xb.y, xb.z, xb.w};
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const _Float16 lo = __builtin_bit_cast(_Float16, (uint16_t)(w[t] & 0xFFFFu));
        const _Float16 hi = __builtin_bit_cast(_Float16, (uint16_t)(w[t] >> 16));
        acc += (float)lo * xs[2 * t] + (float)hi * xs[2 * t + 1];
      }
    }
  } else {
    for (int k = lane; k < n; k += 64) acc += (float)__builtin_bit_cast(_Float16, a[k]) * x[k];
  }
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if (lane == 0) y[row] = acc * scale[row];
}

}  // namespace
}  // namespace nfx

using namespace nfx;

extern "C" {

int nfx_gnn_num_params(void) { return kNumParams; }

int nfx_flux_sweep_coef(int n_u, int width, const int32_t* pos, const int32_t* fcol, const int32_t* diag_pos,
                        const double* vals64, double omega, float* coef, float* bound, void* stream) {
  NFX_REQUIRE(n_u >= 0 && width > 0, "nfx_flux_sweep_coef: bad sizes");
  if (n_u == 0) return NFX_OK;
  NFX_REQUIRE(pos && fcol && diag_pos && vals64 && coef, "nfx_flux_sweep_coef: null pointer");
  NFX_REQUIRE(omega > 0.0 || bound, "nfx_flux_sweep_coef: the automatic weight needs the bound workspace");
  hipStream_t st = (hipStream_t)stream;
  if (bound) {
    hipLaunchKernelGGL(k_flux_bound_init, dim3(1), dim3(64), 0, st, bound);
    hipLaunchKernelGGL(k_flux_bound, dim3(grid_for(n_u, 256)), dim3(256), 0, st, n_u, width, pos, diag_pos, vals64,
                       bound);
  }
  hipLaunchKernelGGL(k_flux_sweep_coef, dim3(grid_for((int64_t)width * n_u, 256)), dim3(256), 0, st, n_u, width,
                     pos, fcol, diag_pos, vals64, omega, bound, coef);
  NFX_LAUNCH_CHECK();
  return NFX_OK;
}

int nfx_schur_ell(int E, const int32_t* pos_b, const int32_t* pos_qb, const int32_t* pos_m,
                  const double* vals64, double* sdiag, double* soff, void* stream) {
  NFX_REQUIRE(E >= 0, "nfx_schur_ell: E < 0");
  if (E == 0) return NFX_OK;
  NFX_REQUIRE(pos_b && pos_qb && pos_m && vals64 && sdiag && soff, "nfx_schur_ell: null pointer");
  hipLaunchKernelGGL(k_schur_ell, dim3(grid_for(E, 256)), dim3(256), 0, (hipStream_t)stream, E, pos_b,
                     pos_qb, pos_m, vals64, sdiag, soff);
  NFX_LAUNCH_CHECK();
  return NFX_OK;
}

int nfx_gnn_features(int E, const int32_t* nbr, const double* sdiag, const double* soff, float* a,
                     float* x, float* sd32, float* so32, void* stream) {
  NFX_REQUIRE(E >= 0, "nfx_gnn_features: E < 0");
  if (E == 0) return NFX_OK;
  NFX_REQUIRE(nbr && sdiag && soff && a && x && sd32 && so32, "nfx_gnn_features: null pointer");
  hipLaunchKernelGGL(k_features, dim3(grid_for(E, 256)), dim3(256), 0, (hipStream_t)stream, E, nbr, sdiag,
                     soff, a, x, sd32, so32);
  NFX_LAUNCH_CHECK();
  return NFX_OK;
}

int nfx_gnn_forward(int E, const int32_t* nbr, const float* a, const float* x, const float* theta32,
                    const uint16_t* theta16, uint16_t* h, uint16_t* m, float* z, const double* sdiag,
                    float* Wd, float* Wo, void* stream) {
  NFX_REQUIRE(E > 0 && (int64_t)E * kW < (int64_t)1 << 31, "nfx_gnn_forward: E out of range");
  NFX_REQUIRE(nbr && a && x && theta32 && theta16 && h && m && z && sdiag && Wd && Wo,
              "nfx_gnn_forward: null pointer");
  hipStream_t st = (hipStream_t)stream;
  const int64_t EW = (int64_t)E * kW;
  hipLaunchKernelGGL(k_gnn_input, dim3(grid_for(EW, 256)), dim3(256), 0, st, E, x, theta16, theta32, z, h);
  NFX_LAUNCH_CHECK();
  const int tiles = grid_for(E, kTile);
  for (int l = 0; l < kLayers; ++l) {
    hipLaunchKernelGGL(k_gnn_layer, dim3(tiles), dim3(256), 0, st, E, nbr, a, theta16, theta32, l,
                       h + l * EW, m + l * EW, z + (l + 1) * EW, h + (l + 1) * EW);
    NFX_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(k_gnn_heads, dim3(grid_for(E, 256)), dim3(256), 0, st, E, nbr, a, theta16, theta32,
                     h + kLayers * EW, sdiag, Wd, Wo);
  NFX_LAUNCH_CHECK();
  return NFX_OK;
}

int nfx_gnn_backward_workspace(int E) {
  (void)E;
  // floats of the per-block partial buffer
  return kBwdBlocks * kLayerPartial;
}

int nfx_gnn_backward(int E, const int32_t* nbr, const float* a, const float* x, const uint16_t* theta16,
                     const uint16_t* h, const uint16_t* m, const float* z, const double* sdiag,
                     const float* gamma_d, const float* gamma_o, float* dh, float* dm, float* partial,
                     float* grad, void* stream) {
  NFX_REQUIRE(E > 0 && (int64_t)E * kW < (int64_t)1 << 31, "nfx_gnn_backward: E out of range");
  NFX_REQUIRE(nbr && a && x && theta16 && h && m && z && sdiag && gamma_d && gamma_o && dh && dm && partial &&
                  grad,
              "nfx_gnn_backward: null pointer");
  hipStream_t st = (hipStream_t)stream;
  const int64_t EW = (int64_t)E * kW;
  hipLaunchKernelGGL(k_heads_bwd, dim3(kBwdBlocks), dim3(256), 0, st, E, nbr, a, theta16, h + kLayers * EW,
                     sdiag, gamma_d, gamma_o, dh, partial);
  NFX_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(256), 0, st, kBwdBlocks, kHeadPartial, kW, partial,
                     grad + kOffUd);
  hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(256), 0, st, kBwdBlocks, kHeadPartial, kW, partial + kW,
                     grad + kOffUe);
  hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(64), 0, st, kBwdBlocks, kHeadPartial, 1,
                     partial + 2 * kW, grad + kOffCd);
  hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(64), 0, st, kBwdBlocks, kHeadPartial, 1,
                     partial + 2 * kW + 1, grad + kOffCe);
  NFX_LAUNCH_CHECK();
  for (int l = kLayers - 1; l >= 0; --l) {
    hipLaunchKernelGGL(k_layer_bwd, dim3(kBwdBlocks), dim3(256), 0, st, E, theta16, l, h + l * EW, m + l * EW,
                       z + (l + 1) * EW, dh, dm, partial);
    NFX_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_reduce_partials, dim3(grid_for(kLayerPartial, 256)), dim3(256), 0, st, kBwdBlocks,
                       kLayerPartial, kLayerPartial, partial, grad + kLayerBase + l * kLayerSize);
    hipLaunchKernelGGL(k_msg_bwd, dim3(grid_for(EW, 256)), dim3(256), 0, st, E, nbr, a, dm, dh);
    NFX_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL(k_input_bwd, dim3(kBwdBlocks), dim3(256), 0, st, E, x, z, dh, partial);
  hipLaunchKernelGGL(k_reduce_partials, dim3(grid_for(kInPartial, 256)), dim3(256), 0, st, kBwdBlocks,
                     kInPartial, kInPartial, partial, grad + kOffW0);
  NFX_LAUNCH_CHECK();
  return NFX_OK;
}

int nfx_ell_spmm(int E, int ns, const int32_t* nbr, const float* d, const float* off, float alpha,
                 const float* X, float beta, const float* Y0, float* Y, void* stream) {
  NFX_REQUIRE(E >= 0 && ns >= 0, "nfx_ell_spmm: negative size");
  if ((int64_t)E * ns == 0) return NFX_OK;
  NFX_REQUIRE(nbr && d && off && X && Y, "nfx_ell_spmm: null pointer");
  hipLaunchKernelGGL(k_ell_spmm, dim3(grid_for((int64_t)E * ns, 256)), dim3(256), 0, (hipStream_t)stream, E,
                     ns, nbr, d, off, alpha, X, beta, Y0, Y);
  NFX_LAUNCH_CHECK();
  return NFX_OK;
}

int nfx_smoother_grad(int E, int ns, const int32_t* nbr, const float* Su, const float* Se, const float* w,
                      float* gamma_d, float* gamma_o, void* stream) {
  NFX_REQUIRE(E >= 0 && ns >= 0, "nfx_smoother_grad: negative size");
  if (E == 0) return NFX_OK;
  NFX_REQUIRE(nbr && Su && Se && w && gamma_d && gamma_o, "nfx_smoother_grad: null pointer");
  hipLaunchKernelGGL(k_smoother_grad, dim3(grid_for(E, 256)), dim3(256), 0, (hipStream_t)stream, E, ns, nbr,
                     Su, Se, w, gamma_d, gamma_o);
  NFX_LAUNCH_CHECK();
  return NFX_OK;
}

int nfx_agg_sum(int n_agg, int ns, int E, const int32_t* agg_ptr, const int32_t* agg_mem, const float* X,
                float* C, void* stream) {
  NFX_REQUIRE(n_agg >= 0 && ns >= 0 && E >= 0, "nfx_agg_sum: negative size");
  if ((int64_t)n_agg * ns == 0) return NFX_OK;
  NFX_REQUIRE(agg_ptr && agg_mem && X && C, "nfx_agg_sum: null pointer");
  hipLaunchKernelGGL(k_agg_sum, dim3(grid_for((int64_t)n_agg * ns, 256)), dim3(256), 0, (hipStream_t)stream,
                     n_agg, ns, E, agg_ptr, agg_mem, X, C);
  NFX_LAUNCH_CHECK();
  return NFX_OK;
}

int nfx_coarse_dense(int n_agg, const int32_t* agg_ptr, const int32_t* agg_mem, const int32_t* agg_of,
                     const int32_t* nbr, const double* sdiag, const double* soff, double* Ac, void* stream) {
  NFX_REQUIRE(n_agg >= 0, "nfx_coarse_dense: n_agg < 0");
  if (n_agg == 0) return NFX_OK;
  NFX_REQUIRE(agg_ptr && agg_mem && agg_of && nbr && sdiag && soff && Ac, "nfx_coarse_dense: null pointer");
  hipLaunchKernelGGL(k_coarse_dense, dim3(grid_for(n_agg, 64)), dim3(64), 0, (hipStream_t)stream, n_agg,
                     agg_ptr, agg_mem, agg_of, nbr, sdiag, soff, Ac);
  NFX_LAUNCH_CHECK();
  return NFX_OK;
}

int nfx_spd_inverse_small(int n, float* A, int lda, void* stream) {
  NFX_REQUIRE(n >= 1 && n <= kInvN && lda >= n, "nfx_spd_inverse_small: need 1 <= n <= 128 and lda >= n");
  NFX_REQUIRE(A, "nfx_spd_inverse_small: null pointer");
  hipLaunchKernelGGL(k_spd_inverse_small, dim3(1), dim3(512), 0, (hipStream_t)stream, n, A, lda);
  NFX_LAUNCH_CHECK();
  return NFX_OK;
}

int nfx_sgemm(int M, int N, int K, float alpha, const float* A, int lda, int transA, const float* B, int ldb,
              int transB, float beta, float* C, int ldc, float* C2, int ldc2, void* stream) {
  NFX_REQUIRE(M >= 0 && N >= 0 && K >= 0, "nfx_sgemm: negative size");
  if (M == 0 || N == 0) return NFX_OK;
  NFX_REQUIRE(A && B && C, "nfx_sgemm: null pointer");
  NFX_REQUIRE(lda >= (transA ? M : K) && ldb >= (transB ? K : N) && ldc >= N, "nfx_sgemm: leading dimension");
  NFX_REQUIRE(!C2 || ldc2 >= M, "nfx_sgemm: leading dimension of the transposed output");
  // 64 x 64 tiles unless fewer than two per CU (256 CUs) would result
  const bool big = (int64_t)grid_for(M, 64) * grid_for(N, 64) >= 512;
  const int ta = transA ? 1 : 0, tb = transB ? 1 : 0;
  if (big) {
    launch_sgemm<64, 4>(ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, C2, ldc2, (hipStream_t)stream);
  } else {
    launch_sgemm<32, 8>(ta, tb, M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, C2, ldc2, (hipStream_t)stream);
  }
  NFX_LAUNCH_CHECK();
  return NFX_OK;
}

int nfx_half_rows(int n, const float* A, uint16_t* A16, float* scale, void* stream) {
  NFX_REQUIRE(n > 0, "nfx_half_rows: n <= 0");
  NFX_REQUIRE(A && A16 && scale, "nfx_half_rows: null pointer");
  hipLaunchKernelGGL(k_half_rows, dim3(grid_for((int64_t)n * 64, 256)), dim3(256), 0, (hipStream_t)stream, n, A,
                     A16, scale);
  NFX_LAUNCH_CHECK();
  return NFX_OK;
}

int nfx_twogrid_apply(const nfx_twogrid_pc* pc, const float* r, float* z, void* stream) {
  return nfx::twogrid_apply_scaled(pc, r, z, nullptr, (hipStream_t)stream);
}

}  // extern "C"

int nfx::twogrid_apply_scaled(const nfx_twogrid_pc* pc, const float* r, float* z, const double* dscale,
                              hipStream_t st) {
  NFX_REQUIRE(pc, "nfx_twogrid_apply: null descriptor");
  const int n_u = pc->n_u, E = pc->n_cells, n_agg = pc->n_agg;
  NFX_REQUIRE(n_u >= 0 && E > 0 && n_agg > 0 && pc->n_tail >= 0, "nfx_twogrid_apply: bad sizes");
  NFX_REQUIRE(pc->n_tail == 0 || pc->tail_dinv, "nfx_twogrid_apply: null tail diagonal");
  NFX_REQUIRE(pc->nbr && pc->Wd && pc->Wo && pc->sd && pc->so && pc->agg_ptr && pc->agg_mem && pc->agg_of &&
                  (pc->Ainv || (pc->Ainv16 && pc->Ascale)) && r && z && pc->work && (n_u == 0 || pc->dMinv),
              "nfx_twogrid_apply: null pointer");
  NFX_REQUIRE(!pc->flux_coef || (pc->flux_col && pc->flux_width > 0),
              "nfx_twogrid_apply: incomplete flux-row operator");
  const bool lu = pc->slot_tail != nullptr;
  NFX_REQUIRE(!lu || (pc->n_tail > 0 && pc->slot_coef && pc->tail_cells && pc->tail_coef),
              "nfx_twogrid_apply: incomplete tail coupling");
  // work: x1[E] | x2[E] | t2[E] | c[n_agg] | y[n_agg] (| rq[E] with the tail coupling)
  float* x1 = pc->work;
  float* x2 = x1 + E;
  float* t2 = x1 + 2 * (int64_t)E;
  float* c = x1 + 3 * (int64_t)E;
  float* y = c + n_agg;
  const float* rp = r + n_u;
  if (lu) {
    float* rq = y + n_agg;
    hipLaunchKernelGGL(k_tg_tail_pre, dim3(grid_for(E, 256)), dim3(256), 0, st, E, n_u, pc->slot_tail,
                       pc->slot_coef, r, rq);
    rp = rq;
  }
  // (a fused x1 + restriction kernel, recomputing x1 at the members' neighbours, measured
  // 18.6 us against 4.8 + 5.9 us for the two launches at 196,608 cells)
  const bool sweeps = pc->flux_coef != nullptr && n_u > 0;
  if (sweeps) {
    const int cb = grid_for(E, 256);
    const int fb = grid_for(n_u, 256);
    hipLaunchKernelGGL(k_tg_pre_flux, dim3(cb + fb), dim3(256), 0, st, E, cb, pc->nbr, pc->Wd, pc->Wo, rp, x1, n_u,
                       pc->flux_width, pc->flux_col, pc->flux_coef, r, dscale, z);
  } else {
    hipLaunchKernelGGL(k_tg_pre, dim3(grid_for(E, 256)), dim3(256), 0, st, E, pc->nbr, pc->Wd, pc->Wo, rp, x1);
  }
  hipLaunchKernelGGL(k_tg_restrict, dim3(grid_for((int64_t)n_agg * 64, 256)), dim3(256), 0, st, n_agg, pc->agg_ptr,
                     pc->agg_mem, pc->nbr, pc->sd, pc->so, rp, x1, c);
  if (pc->Ainv16)
    hipLaunchKernelGGL(k_gemv_h, dim3(grid_for((int64_t)n_agg * 64, 256)), dim3(256), 0, st, n_agg, pc->Ainv16,
                       pc->Ascale, c, y);
  else
    hipLaunchKernelGGL(k_gemv, dim3(grid_for((int64_t)n_agg * 64, 256)), dim3(256), 0, st, n_agg, pc->Ainv, c, y);
  hipLaunchKernelGGL(k_tg_post1, dim3(grid_for(E, 256)), dim3(256), 0, st, E, pc->nbr, pc->agg_of, pc->sd, pc->so,
                     rp, x1, y, x2, t2);
  const int i0 = sweeps ? n_u : 0;
  hipLaunchKernelGGL(k_tg_post2, dim3(grid_for((int64_t)n_u - i0 + E + pc->n_tail, 256)), dim3(256), 0, st, i0, n_u, E,
                     pc->nbr, pc->Wd, pc->Wo, x2, t2, pc->dMinv, r, pc->psign, dscale, pc->n_tail, pc->tail_dinv,
                     lu ? pc->tail_cells : nullptr, lu ? pc->tail_coef : nullptr, z);
  NFX_LAUNCH_CHECK();
  return NFX_OK;
}
