// Learned density surrogate: an MLP 2 -> 128 -> 128 -> 128 -> 1 (SiLU)
// trained on (query, log-density) pairs from the exact direct sum and served
// with a sampled check against the exact sum (bk_mlp_*, declared in
// include/bbrcit_hip.h).  The reference has no counterpart; the exact direct
// sum stays both the fallback and the arbiter of the surrogate's error.
//
// Weights are BF16 (FP32 master copies updated by Adam; every forward pass,
// in training and serving, uses the BF16-rounded weights), accumulation is
// FP32, biases are FP32.
//
// Training kernels (FP32 MFMA v_mfma_f32_16x16x4f32, all hand-written):
//   gather -> layer 1 (VALU) -> GEMM+bias+SiLU x 2 -> head (output, loss
//   gradient, dZ3) -> weight-gradient GEMMs (split over the batch, reduced
//   in fixed order) -> input-gradient GEMMs fused with SiLU' -> column sums
//   for biases and the 2-wide first layer -> Adam.
// Serving kernel (BF16 MFMA v_mfma_f32_16x16x32_bf16): one wavefront
// carries 16 queries through the whole network in registers.  Queries sit
// on the N axis of the MFMA, units on the M axis, so a layer's output
// fragment (lane l: query l%16, units 16t + 4(l/16) + i) is directly the
// next layer's B operand once the K order inside each 32-wide step is
// permuted to match; the packed weights carry the same permutation.
// Activations enter the BF16 MFMA as a hi + lo pair (a = hi + lo, both
// BF16), so only the weights are rounded to BF16.
#include "../../include/bbrcit_hip.h"
#include "bk_common.h"
#include "bk_internal.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <limits>
#include <memory>
#include <string>
#include <vector>

namespace {

using bk::DevBuf;
using bk::DeviceGuard;
using bk::fail;
using bk::guarded;
using bk::hip_check;

typedef __attribute__((__vector_size__(4 * sizeof(float)))) float v4f;
typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 v8bf;

constexpr int kWidth = 128;
// parameter layout (floats)
constexpr int kOffW1 = 0;                       // [128][2]
constexpr int kOffB1 = kOffW1 + kWidth * 2;     // [128]
constexpr int kOffW2 = kOffB1 + kWidth;         // [128][128] (out, in)
constexpr int kOffB2 = kOffW2 + kWidth * kWidth;
constexpr int kOffW3 = kOffB2 + kWidth;
constexpr int kOffB3 = kOffW3 + kWidth * kWidth;
constexpr int kOffW4 = kOffB3 + kWidth;         // [128]
constexpr int kOffB4 = kOffW4 + kWidth;         // [1]
constexpr int kNumParams = kOffB4 + 1;
static_assert(kNumParams == BK_MLP_NUM_PARAMS, "parameter count");

constexpr int kGemmTile = 64;
constexpr int kGemmK = 16;
constexpr int kSplitK = 64;   // batch splits of the weight-gradient GEMMs (4 output tiles x 64 = 256 workgroups)
constexpr int kMaxBatch = 65536;

__device__ __forceinline__ float sigmoidf(float z) { return 1.0f / (1.0f + __expf(-z)); }
__device__ __forceinline__ float silu(float z) { return z * sigmoidf(z); }
__device__ __forceinline__ float dsilu(float z) {
  const float s = sigmoidf(z);
  return s * (1.0f + z * (1.0f - s));
}

__device__ __forceinline__ unsigned long long splitmix64(unsigned long long x) {
  x += 0x9E3779B97F4A7C15ULL;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
  return x ^ (x >> 31);
}

// ---- training -------------------------------------------------------------------
// Batch b of step `step`: sample index splitmix64(seed + step * B + b) mod n.
__global__ void mlp_gather_kernel(const float2 *__restrict__ xin, const float *__restrict__ tin, long long n,
                                  const long long *__restrict__ step, unsigned long long seed, int B,
                                  float2 *__restrict__ X, float *__restrict__ T) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  const unsigned long long h = splitmix64(seed + (unsigned long long)(*step) * (unsigned long long)B + b);
  const long long i = (long long)(h % (unsigned long long)n);
  X[b] = xin[i];
  T[b] = tin[i];
}

// Z1 = X W1^T + b1, H1 = SiLU(Z1); one thread per (sample, unit).
__global__ void mlp_layer1_kernel(const float2 *__restrict__ X, const float *__restrict__ wq, int B,
                                  float *__restrict__ Z1, float *__restrict__ H1) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= B * kWidth) return;
  const int b = k / kWidth, o = k % kWidth;
  const float2 x = X[b];
  const float z = __builtin_fmaf(wq[kOffW1 + 2 * o], x.x, __builtin_fmaf(wq[kOffW1 + 2 * o + 1], x.y, wq[kOffB1 + o]));
  Z1[k] = z;
  H1[k] = silu(z);
}

enum Epi : int { EPI_STORE = 0, EPI_BIAS_SILU = 1, EPI_DSILU = 2 };

// C[M][N] = op(A)[M][K] op(B)[K][N] (FP32 MFMA 16x16x4).  op(A)[m][k] is
// A[k*lda + m] if TA else A[m*lda + k]; op(B)[k][n] is B[n*ldb + k] if TB
// else B[k*ldb + n].  M, N multiples of 64; the K range of split z is
// [z*K/gridDim.z, (z+1)*K/gridDim.z), a multiple of 16; split z writes
// C + z*M*ldc.  Epilogues: EPI_BIAS_SILU stores Z = C + bias[n] to C and
// SiLU(Z) to H; EPI_DSILU stores C * SiLU'(Zaux).
template <bool TA, bool TB, int EPI>
__global__ __launch_bounds__(256) void gemm_kernel(int M, int N, int K, const float *__restrict__ A, int lda,
                                                   const float *__restrict__ Bm, int ldb, float *__restrict__ C,
                                                   int ldc, const float *__restrict__ bias, float *__restrict__ H,
                                                   const float *__restrict__ Zaux) {
  __shared__ float sA[kGemmK][kGemmTile + 1];
  __shared__ float sB[kGemmK][kGemmTile + 1];
  const int m0 = blockIdx.y * kGemmTile, n0 = blockIdx.x * kGemmTile;
  const int kc = K / gridDim.z;
  const int kb = blockIdx.z * kc, ke = kb + kc;
  C += (long long)blockIdx.z * M * ldc;
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int wm = (wave / 2) * 32, wn = (wave % 2) * 32;
  v4f acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = v4f{0.f, 0.f, 0.f, 0.f};
  for (int k0 = kb; k0 < ke; k0 += kGemmK) {
    for (int e = threadIdx.x; e < kGemmTile * kGemmK; e += 256) {
      if (TA) {
        const int kk = e / kGemmTile, mm = e % kGemmTile;
        sA[kk][mm] = A[(long long)(k0 + kk) * lda + m0 + mm];
      } else {
        const int mm = e / kGemmK, kk = e % kGemmK;
        sA[kk][mm] = A[(long long)(m0 + mm) * lda + k0 + kk];
      }
      if (TB) {
        const int nn = e / kGemmK, kk = e % kGemmK;
        sB[kk][nn] = Bm[(long long)(n0 + nn) * ldb + k0 + kk];
      } else {
        const int kk = e / kGemmTile, nn = e % kGemmTile;
        sB[kk][nn] = Bm[(long long)(k0 + kk) * ldb + n0 + nn];
      }
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < kGemmK; ks += 4) {
      const int kr = ks + lane / 16;
      const float a0 = sA[kr][wm + lane % 16], a1 = sA[kr][wm + 16 + lane % 16];
      const float b0 = sB[kr][wn + lane % 16], b1 = sB[kr][wn + 16 + lane % 16];
      acc[0][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, b0, acc[0][0], 0, 0, 0);
      acc[0][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, b1, acc[0][1], 0, 0, 0);
      acc[1][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b0, acc[1][0], 0, 0, 0);
      acc[1][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, b1, acc[1][1], 0, 0, 0);
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + wm + 16 * i + 4 * (lane / 16) + r;
        const int n = n0 + wn + 16 * j + lane % 16;
        const long long off = (long long)m * ldc + n;
        float v = acc[i][j][r];
        if (EPI == EPI_BIAS_SILU) {
          v += bias[n];
          H[off] = silu(v);
        } else if (EPI == EPI_DSILU) {
          v *= dsilu(Zaux[off]);
        }
        C[off] = v;
      }
}

// Sum of `splits` partial matrices of n floats (fixed order).
__global__ void sum_splits_kernel(const float *__restrict__ part, int splits, int n, float *__restrict__ out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  float s = 0.0f;
#pragma unroll 8
  for (int z = 0; z < splits; ++z) s += part[(long long)z * n + k];
  out[k] = s;
}

// Output layer and loss gradient, one wavefront per sample:
// y = W4 . H3 + b4, r = y - t, sq[b] = r^2, dy[b] = 2 r / B,
// dZ3[b][k] = dy W4[k] SiLU'(Z3[b][k]).
__global__ __launch_bounds__(256) void mlp_head_kernel(const float *__restrict__ H3, const float *__restrict__ Z3,
                                                       const float *__restrict__ wq, const float *__restrict__ T,
                                                       int B, float *__restrict__ sq, float *__restrict__ dy,
                                                       float *__restrict__ dZ3) {
  const int b = blockIdx.x * 4 + threadIdx.x / 64;
  const int lane = threadIdx.x % 64;
  if (b >= B) return;
  const float *h = H3 + (long long)b * kWidth;
  float p = wq[kOffW4 + lane] * h[lane] + wq[kOffW4 + 64 + lane] * h[64 + lane];
  for (int o = 32; o > 0; o >>= 1) p += __shfl_xor(p, o, 64);
  const float y = p + wq[kOffB4];
  const float r = y - T[b];
  const float g = 2.0f * r / (float)B;
  if (lane == 0) {
    sq[b] = r * r;
    dy[b] = g;
  }
  const float *z = Z3 + (long long)b * kWidth;
  float *d = dZ3 + (long long)b * kWidth;
  d[lane] = g * wq[kOffW4 + lane] * dsilu(z[lane]);
  d[64 + lane] = g * wq[kOffW4 + 64 + lane] * dsilu(z[64 + lane]);
}

// The backward pass's column sums in two passes, each sum in a fixed order.
// Family f of output column c (kGradFamilies of them):
//   0: sum_b dy[b] H3[b][c] (dW4)    1: sum_b dZ3[b][c] (db3)
//   2: sum_b dZ2[b][c] (db2)         3: sum_b dZ1[b][c] x_b (dW1[c][0])
//   4: sum_b dZ1[b][c] y_b (dW1[c][1])  5: sum_b dZ1[b][c] (db1)
//   6: sum_b dy[b] (db4; column 0 only)
// Pass 1: workgroup (chunk, matrix) reads rows [chunk * kGradRows, ...) of
// one matrix (H3 with dy: families 0 and 6; dZ3: 1; dZ2: 2; dZ1: 3, 4, 5)
// and sums its families in FP64, thread t on column t % 128 and the rows of
// parity t / 128 (each row a coalesced 512-byte line, four rows of a thread
// in flight), the two parities added in order.  Pass 2: per (f, c), the
// chunks added in order.
constexpr int kGradFamilies = 7;
constexpr int kGradRows = 256;

struct GradCols {
  const float *H3, *dZ3, *dZ2, *dZ1, *dy, *X;
};

__global__ __launch_bounds__(256) void grad_colsum_partial_kernel(GradCols a, int B, double *__restrict__ part) {
  __shared__ double sh[3][kWidth];
  const int chunk = blockIdx.x, mat = blockIdx.y;
  const int c = threadIdx.x % kWidth, par = threadIdx.x / kWidth;
  const int r0 = chunk * kGradRows, r1 = min(B, r0 + kGradRows);
  const float *A = mat == 0 ? a.H3 : mat == 1 ? a.dZ3 : mat == 2 ? a.dZ2 : a.dZ1;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  int r = r0 + par;
  for (; r + 6 < r1; r += 8) {
    float v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) v[u] = A[(long long)(r + 2 * u) * kWidth + c];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int ru = r + 2 * u;
      if (mat == 0) {
        s0 += (double)v[u] * (double)a.dy[ru];
        s1 += c == 0 ? (double)a.dy[ru] : 0.0;
      } else if (mat == 3) {
        s0 += (double)v[u] * (double)a.X[2 * ru];
        s1 += (double)v[u] * (double)a.X[2 * ru + 1];
        s2 += (double)v[u];
      } else {
        s0 += (double)v[u];
      }
    }
  }
  for (; r < r1; r += 2) {
    const float v = A[(long long)r * kWidth + c];
    if (mat == 0) {
      s0 += (double)v * (double)a.dy[r];
      s1 += c == 0 ? (double)a.dy[r] : 0.0;
    } else if (mat == 3) {
      s0 += (double)v * (double)a.X[2 * r];
      s1 += (double)v * (double)a.X[2 * r + 1];
      s2 += (double)v;
    } else {
      s0 += (double)v;
    }
  }
  if (par == 1) {
    sh[0][c] = s0;
    sh[1][c] = s1;
    sh[2][c] = s2;
  }
  __syncthreads();
  if (par == 0) {
    double *o = part + (long long)chunk * kGradFamilies * kWidth;
    if (mat == 0) {
      o[0 * kWidth + c] = s0 + sh[0][c];
      o[6 * kWidth + c] = s1 + sh[1][c];
    } else if (mat == 3) {
      o[3 * kWidth + c] = s0 + sh[0][c];
      o[4 * kWidth + c] = s1 + sh[1][c];
      o[5 * kWidth + c] = s2 + sh[2][c];
    } else {
      o[mat * kWidth + c] = s0 + sh[0][c];
    }
  }
}

__global__ void grad_colsum_final_kernel(const double *__restrict__ part, int chunks, float *__restrict__ g) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= kGradFamilies * kWidth) return;
  const int f = k / kWidth, c = k % kWidth;
  if (f == 6 && c != 0) return;
  double s = 0.0;
#pragma unroll 8
  for (int ch = 0; ch < chunks; ++ch) s += part[((long long)ch * kGradFamilies + f) * kWidth + c];
  const float v = (float)s;
  switch (f) {
    case 0: g[kOffW4 + c] = v; break;
    case 1: g[kOffB3 + c] = v; break;
    case 2: g[kOffB2 + c] = v; break;
    case 3: g[kOffW1 + 2 * c] = v; break;
    case 4: g[kOffW1 + 2 * c + 1] = v; break;
    case 5: g[kOffB1 + c] = v; break;
    default: g[kOffB4] = v; break;
  }
}

__device__ __forceinline__ float bf16_round(float x) { return (float)(__bf16)x; }

// Adam on the FP32 master parameters, then the BF16-rounded working copy
// (weights rounded, biases copied).  *step indexes this call's learning-rate
// schedule; it is advanced by mlp_step_kernel after the update, since no
// thread of this grid can know when all others have read it.
__global__ void mlp_adam_kernel(float *__restrict__ theta, float *__restrict__ m, float *__restrict__ v,
                                const float *__restrict__ g, float *__restrict__ wq,
                                const long long *__restrict__ step, long long t_offset,
                                const float *__restrict__ lr_sched, float beta1, float beta2, float eps) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= kNumParams) return;
  const long long t = t_offset + *step + 1;
  const float lr = lr_sched[*step];
  const float gk = g[k];
  const float mk = beta1 * m[k] + (1.0f - beta1) * gk;
  const float vk = beta2 * v[k] + (1.0f - beta2) * gk * gk;
  m[k] = mk;
  v[k] = vk;
  const float mh = mk / (1.0f - powf(beta1, (float)t));
  const float vh = vk / (1.0f - powf(beta2, (float)t));
  const float th = theta[k] - lr * mh / (sqrtf(vh) + eps);
  theta[k] = th;
  const bool is_bias = (k >= kOffB1 && k < kOffW2) || (k >= kOffB2 && k < kOffW3) || (k >= kOffB3 && k < kOffW4) ||
                       k == kOffB4;
  wq[k] = is_bias ? th : bf16_round(th);
}

__global__ void mlp_step_kernel(long long *step) {
  if (threadIdx.x == 0) *step += 1;
}

__global__ void mlp_quantize_kernel(const float *__restrict__ theta, float *__restrict__ wq) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= kNumParams) return;
  const bool is_bias = (k >= kOffB1 && k < kOffW2) || (k >= kOffB2 && k < kOffW3) || (k >= kOffB3 && k < kOffW4) ||
                       k == kOffB4;
  wq[k] = is_bias ? theta[k] : bf16_round(theta[k]);
}

// ---- serving -----------------------------------------------------------------------
// Packed BF16 A operands of layers 2 and 3: [layer][tile t][kstep s][lane][8].
// Lane l of (t, s) holds W[16t + l%16][perm(s, l/16, j)], j = 0..7, with
// perm(s, g, j) = 32s + 4g + j (j < 4), 32s + 16 + 4g + (j - 4) (j >= 4): the
// units lane group g holds of output tiles 2s and 2s+1.
__device__ __forceinline__ int kperm(int s, int g, int j) { return 32 * s + (j < 4 ? 4 * g + j : 16 + 4 * g + j - 4); }

__global__ void mlp_pack_kernel(const float *__restrict__ wq, v8bf *__restrict__ pk) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;  // over 2 * 8 * 4 * 64
  if (idx >= 2 * 8 * 4 * 64) return;
  const int lane = idx % 64, s = (idx / 64) % 4, t = (idx / 256) % 8, layer = idx / 2048;
  const float *W = wq + (layer == 0 ? kOffW2 : kOffW3);
  const int o = 16 * t + lane % 16, g = lane / 16;
  v8bf v;
#pragma unroll
  for (int j = 0; j < 8; ++j) v[j] = (__bf16)W[o * kWidth + kperm(s, g, j)];
  pk[idx] = v;
}

struct ServeArgs {
  const float *qx, *qy;
  long long n;
  const v8bf *pk;
  const float *wq;
  float cx, cy, inv_s;      // input map: ((x - cx) inv_s, (y - cy) inv_s)
  double mu, sigma, eps;    // density = max(exp(mu + sigma y) - eps, 0)
  double *out;              // density (FP64), may be null
  float *yout;              // normalised output, may be null
};

__device__ __forceinline__ void split_bf16(float a, __bf16 &hi, __bf16 &lo) {
  hi = (__bf16)a;
  lo = (__bf16)(a - (float)hi);
}

__global__ __launch_bounds__(256) void mlp_infer_kernel(ServeArgs a) {
  __shared__ v8bf sw[2 * 8 * 4 * 64];  // 64 KB
  for (int i = threadIdx.x; i < 2 * 8 * 4 * 64; i += 256) sw[i] = a.pk[i];
  __syncthreads();
  const int lane = threadIdx.x % 64, g = lane / 16, qn = lane % 16;
  const long long nblocks = (a.n + 15) / 16;
  const float *wq = a.wq;
  for (long long blk = (long long)blockIdx.x * 4 + threadIdx.x / 64; blk < nblocks; blk += (long long)gridDim.x * 4) {
    const long long qi = blk * 16 + qn;
    const bool valid = qi < a.n;
    const float x = valid ? (a.qx[qi] - a.cx) * a.inv_s : 0.0f;
    const float y = valid ? (a.qy[qi] - a.cy) * a.inv_s : 0.0f;
    float h[8][4];
#pragma unroll
    for (int t = 0; t < 8; ++t)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int u = 16 * t + 4 * g + i;
        h[t][i] = silu(__builtin_fmaf(wq[kOffW1 + 2 * u], x, __builtin_fmaf(wq[kOffW1 + 2 * u + 1], y, wq[kOffB1 + u])));
      }
#pragma unroll
    for (int layer = 0; layer < 2; ++layer) {
      v8bf bhi[4], blo[4];
#pragma unroll
      for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          const float v = j < 4 ? h[2 * s][j] : h[2 * s + 1][j - 4];
          __bf16 hi, lo;
          split_bf16(v, hi, lo);
          bhi[s][j] = hi;
          blo[s][j] = lo;
        }
      const float *bias = wq + (layer == 0 ? kOffB2 : kOffB3);
#pragma unroll
      for (int t = 0; t < 8; ++t) {
        v4f acc;
#pragma unroll
        for (int i = 0; i < 4; ++i) acc[i] = bias[16 * t + 4 * g + i];
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          const v8bf w = sw[((layer * 8 + t) * 4 + s) * 64 + lane];
          acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w, bhi[s], acc, 0, 0, 0);
          acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w, blo[s], acc, 0, 0, 0);
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) h[t][i] = silu(acc[i]);
      }
    }
    float p = 0.0f;
#pragma unroll
    for (int t = 0; t < 8; ++t)
#pragma unroll
      for (int i = 0; i < 4; ++i) p = __builtin_fmaf(wq[kOffW4 + 16 * t + 4 * g + i], h[t][i], p);
    p += __shfl_xor(p, 16, 64);
    p += __shfl_xor(p, 32, 64);
    const float yv = p + wq[kOffB4];
    if (valid && g == 0) {
      if (a.yout) a.yout[qi] = yv;
      if (a.out) a.out[qi] = fmax(exp(a.mu + a.sigma * (double)yv) - a.eps, 0.0);
    }
  }
}

// ---- sampled check and tile fallback ----------------------------------------------
__device__ __forceinline__ int cell_of(float x, float y, float lo, float inv_w, int G) {
  int cx = (int)floorf((x - lo) * inv_w), cy = (int)floorf((y - lo) * inv_w);
  cx = min(max(cx, 0), G - 1);
  cy = min(max(cy, 0), G - 1);
  return cy * G + cx;
}

// Deterministic sample: query i is checked if splitmix64(seed ^ i) mod `period` == 0.
__global__ void select_samples_kernel(long long n, unsigned long long seed, unsigned long long period,
                                      int *__restrict__ count, long long *__restrict__ list, long long cap) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (splitmix64(seed ^ (unsigned long long)i) % period == 0) {
    const int k = atomicAdd(count, 1);
    if (k < cap) list[k] = i;
  }
}

__global__ void gather_kernel(const float *__restrict__ qx, const float *__restrict__ qy,
                              const long long *__restrict__ list, const int *__restrict__ count, long long cap,
                              float *__restrict__ gx, float *__restrict__ gy) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= min((long long)*count, cap)) return;
  const long long i = list[k];
  gx[k] = qx[i];
  gy[k] = qy[i];
}

// err = |s - e| / max(e, floor) for each sample; per-cell max (non-negative
// floats compare as their bit patterns); the exact value replaces the
// surrogate's.
__global__ void check_samples_kernel(const float *__restrict__ qx, const float *__restrict__ qy,
                                     const long long *__restrict__ list, const int *__restrict__ count,
                                     long long cap, const double *__restrict__ exact, double *__restrict__ out,
                                     double floor, float lo, float inv_w, int G, unsigned int *__restrict__ cell_err,
                                     unsigned int *__restrict__ cell_cnt, unsigned int *__restrict__ max_err) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= min((long long)*count, cap)) return;
  const long long i = list[k];
  const double e = exact[k];
  const double s = out[i];
  const float err = (float)(fabs(s - e) / fmax(e, floor));
  const unsigned int bits = __float_as_uint(fminf(err, 3.0e38f));
  const int c = cell_of(qx[i], qy[i], lo, inv_w, G);
  atomicMax(cell_err + c, bits);
  atomicAdd(cell_cnt + c, 1u);
  atomicMax(max_err, bits);
  out[i] = e;
}

// A cell falls back if a sample in it failed or if it holds no sample; its
// sampled queries already hold their exact sums and are left out.
__global__ void select_fallback_kernel(const float *__restrict__ qx, const float *__restrict__ qy, long long n,
                                       const unsigned int *__restrict__ cell_err,
                                       const unsigned int *__restrict__ cell_cnt, float tol, float lo, float inv_w,
                                       int G, unsigned long long seed, unsigned long long period,
                                       int *__restrict__ count, long long *__restrict__ list, long long cap) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (splitmix64(seed ^ (unsigned long long)i) % period == 0) return;
  const int c = cell_of(qx[i], qy[i], lo, inv_w, G);
  if (__uint_as_float(cell_err[c]) > tol || cell_cnt[c] == 0u) {
    const int k = atomicAdd(count, 1);
    if (k < cap) list[k] = i;
  }
}

__global__ void scatter_kernel(const long long *__restrict__ list, const int *__restrict__ count, long long cap,
                               const double *__restrict__ vals, double *__restrict__ out) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= min((long long)*count, cap)) return;
  out[list[k]] = vals[k];
}

__global__ void zero_u32_kernel(unsigned int *p, int n) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) p[k] = 0u;
}

unsigned grid_for(long long n, int threads) { return (unsigned)std::max<long long>(1, (n + threads - 1) / threads); }

}  // namespace

struct bk_mlp {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev[8] = {};
  DevBuf theta, m, v, grad, wq, pk, state, lr;
  // training set
  DevBuf xin, tin;
  long long n_train = 0;
  double lo = -4.0, hi = 4.0, mu = 0.0, sigma = 1.0, eps = 0.0;
  bool have_params = false;
  // training workspace
  int ws_batch = 0;
  DevBuf X, T, Z1, H1, Z2, H2, Z3, H3, dZ3, dZ2, dZ1, dH, sq, dy, part, gpart;
  // serving workspace
  DevBuf slist, flist, cnt, cell_err, gx, gy, exact, fx, fy, fexact, keys_a, keys_b, sort_tmp;
  long long serve_cap = 0, fb_cap = 0;
  long long steps_done = 0;
};

namespace {

bk_mlp *require_mlp(bk_mlp *p) {
  if (!p) fail(BK_ERR_INVALID_ARG, "null surrogate");
  return p;
}

void ensure_workspace(bk_mlp *p, int B) {
  if (p->ws_batch >= B) return;
  const size_t mat = (size_t)B * kWidth * sizeof(float);
  for (DevBuf *b : {&p->Z1, &p->H1, &p->Z2, &p->H2, &p->Z3, &p->H3, &p->dZ3, &p->dZ2, &p->dZ1, &p->dH})
    b->reserve(mat);
  p->X.reserve((size_t)B * sizeof(float2));
  p->T.reserve((size_t)B * sizeof(float));
  p->sq.reserve((size_t)B * sizeof(float));
  p->dy.reserve((size_t)B * sizeof(float));
  p->part.reserve((size_t)kSplitK * kWidth * kWidth * sizeof(float));
  p->gpart.reserve((size_t)((B + kGradRows - 1) / kGradRows) * kGradFamilies * kWidth * sizeof(double));
  p->ws_batch = B;
}

template <bool TA, bool TB, int EPI>
void gemm(hipStream_t s, int M, int N, int K, int splits, const float *A, int lda, const float *B, int ldb, float *C,
          int ldc, const float *bias = nullptr, float *H = nullptr, const float *Z = nullptr) {
  if (M % kGemmTile || N % kGemmTile || K % (kGemmK * splits)) fail(BK_ERR_STATE, "internal: GEMM shape");
  hipLaunchKernelGGL((gemm_kernel<TA, TB, EPI>), dim3(N / kGemmTile, M / kGemmTile, splits), dim3(256), 0, s, M, N,
                     K, A, lda, B, ldb, C, ldc, bias, H, Z);
  hip_check(hipGetLastError(), "GEMM launch");
}

// The backward pass's column sums (grad_colsum_*_kernel) into g.
void grad_colsums(bk_mlp *p, int B, float *g) {
  const hipStream_t s = p->stream;
  const int chunks = (B + kGradRows - 1) / kGradRows;  // gpart sized by ensure_workspace
  GradCols a{p->H3.as<float>(), p->dZ3.as<float>(), p->dZ2.as<float>(), p->dZ1.as<float>(), p->dy.as<float>(),
             p->X.as<float>()};
  hipLaunchKernelGGL(grad_colsum_partial_kernel, dim3(chunks, 4), dim3(256), 0, s, a, B,
                     p->gpart.as<double>());
  hip_check(hipGetLastError(), "column-sum launch");
  hipLaunchKernelGGL(grad_colsum_final_kernel, dim3(grid_for(kGradFamilies * kWidth, 256)), dim3(256), 0, s,
                     p->gpart.as<double>(), chunks, g);
  hip_check(hipGetLastError(), "column-sum launch");
}

// One training step: forward and backward of batch X, T (already gathered)
// leaving the loss terms in sq and the gradient in p->grad.
void forward_backward(bk_mlp *p, int B) {
  const hipStream_t s = p->stream;
  float *wq = p->wq.as<float>();
  float *g = p->grad.as<float>();
  hipLaunchKernelGGL(mlp_layer1_kernel, dim3(grid_for((long long)B * kWidth, 256)), dim3(256), 0, s,
                     p->X.as<float2>(), wq, B, p->Z1.as<float>(), p->H1.as<float>());
  hip_check(hipGetLastError(), "layer-1 launch");
  gemm<false, true, EPI_BIAS_SILU>(s, B, kWidth, kWidth, 1, p->H1.as<float>(), kWidth, wq + kOffW2, kWidth,
                                   p->Z2.as<float>(), kWidth, wq + kOffB2, p->H2.as<float>());
  gemm<false, true, EPI_BIAS_SILU>(s, B, kWidth, kWidth, 1, p->H2.as<float>(), kWidth, wq + kOffW3, kWidth,
                                   p->Z3.as<float>(), kWidth, wq + kOffB3, p->H3.as<float>());
  hipLaunchKernelGGL(mlp_head_kernel, dim3(grid_for(B, 4)), dim3(256), 0, s, p->H3.as<float>(), p->Z3.as<float>(),
                     wq, p->T.as<float>(), B, p->sq.as<float>(), p->dy.as<float>(), p->dZ3.as<float>());
  hip_check(hipGetLastError(), "head launch");
  // layer 3: dW3 = dZ3^T H2, db3, dZ2 = (dZ3 W3) * SiLU'(Z2)
  const int splits = B % (kGemmK * kSplitK) == 0 ? kSplitK : 1;
  gemm<true, false, EPI_STORE>(s, kWidth, kWidth, B, splits, p->dZ3.as<float>(), kWidth, p->H2.as<float>(), kWidth,
                               splits > 1 ? p->part.as<float>() : g + kOffW3, kWidth);
  if (splits > 1) {
    hipLaunchKernelGGL(sum_splits_kernel, dim3(grid_for(kWidth * kWidth, 256)), dim3(256), 0, s, p->part.as<float>(),
                       splits, kWidth * kWidth, g + kOffW3);
    hip_check(hipGetLastError(), "split-sum launch");
  }
  gemm<false, false, EPI_DSILU>(s, B, kWidth, kWidth, 1, p->dZ3.as<float>(), kWidth, wq + kOffW3, kWidth,
                                p->dZ2.as<float>(), kWidth, nullptr, nullptr, p->Z2.as<float>());
  // layer 2
  gemm<true, false, EPI_STORE>(s, kWidth, kWidth, B, splits, p->dZ2.as<float>(), kWidth, p->H1.as<float>(), kWidth,
                               splits > 1 ? p->part.as<float>() : g + kOffW2, kWidth);
  if (splits > 1) {
    hipLaunchKernelGGL(sum_splits_kernel, dim3(grid_for(kWidth * kWidth, 256)), dim3(256), 0, s, p->part.as<float>(),
                       splits, kWidth * kWidth, g + kOffW2);
    hip_check(hipGetLastError(), "split-sum launch");
  }
  gemm<false, false, EPI_DSILU>(s, B, kWidth, kWidth, 1, p->dZ2.as<float>(), kWidth, wq + kOffW2, kWidth,
                                p->dZ1.as<float>(), kWidth, nullptr, nullptr, p->Z1.as<float>());
  // dW4, db4, db3, db2 and layer 1 (dW1[o][0] = sum_b dZ1[b][o] x_b,
  // dW1[o][1] = sum_b dZ1[b][o] y_b, db1) as column sums
  grad_colsums(p, B, g);
}

void upload_theta(bk_mlp *p, const float *theta) {
  hip_check(hipMemcpyAsync(p->theta.ptr, theta, kNumParams * sizeof(float), hipMemcpyHostToDevice, p->stream),
            "upload parameters");
  hipLaunchKernelGGL(mlp_quantize_kernel, dim3(grid_for(kNumParams, 256)), dim3(256), 0, p->stream,
                     p->theta.as<float>(), p->wq.as<float>());
  hip_check(hipGetLastError(), "quantize launch");
  hipLaunchKernelGGL(mlp_pack_kernel, dim3(grid_for(2 * 8 * 4 * 64, 256)), dim3(256), 0, p->stream,
                     p->wq.as<float>(), p->pk.as<v8bf>());
  hip_check(hipGetLastError(), "pack launch");
  hip_check(hipStreamSynchronize(p->stream), "hipStreamSynchronize");
  p->have_params = true;
}

ServeArgs serve_args(const bk_mlp *p, const float *qx, const float *qy, long long n, double *out, float *yout) {
  ServeArgs a{};
  a.qx = qx;
  a.qy = qy;
  a.n = n;
  a.pk = p->pk.as<v8bf>();
  a.wq = p->wq.as<float>();
  a.cx = (float)(0.5 * (p->lo + p->hi));
  a.cy = a.cx;
  a.inv_s = (float)(2.0 / (p->hi - p->lo));
  a.mu = p->mu;
  a.sigma = p->sigma;
  a.eps = p->eps;
  a.out = out;
  a.yout = yout;
  return a;
}

void launch_infer(bk_mlp *p, const float *qx, const float *qy, long long n, double *out, float *yout) {
  if (n <= 0) return;
  const long long blocks16 = (n + 15) / 16;
  const unsigned grid = (unsigned)std::min<long long>((blocks16 + 3) / 4, 8LL * bk::device_cu_count());
  hipLaunchKernelGGL(mlp_infer_kernel, dim3(grid), dim3(256), 0, p->stream, serve_args(p, qx, qy, n, out, yout));
  hip_check(hipGetLastError(), "inference launch");
}

float elapsed(hipEvent_t a, hipEvent_t b) {
  float ms = 0.f;
  hip_check(hipEventElapsedTime(&ms, a, b), "hipEventElapsedTime");
  return ms;
}

}  // namespace

extern "C" {

int bk_mlp_create(bk_ctx *ctx, unsigned long long seed, bk_mlp **out) {
  return guarded([&] {
    if (!out) fail(BK_ERR_INVALID_ARG, "null output pointer");
    *out = nullptr;
    if (!ctx) fail(BK_ERR_INVALID_ARG, "null context");
    DeviceGuard dg(bk::ctx_device(ctx));
    std::unique_ptr<bk_mlp> p(new bk_mlp());
    p->device = bk::ctx_device(ctx);
    hip_check(hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking), "hipStreamCreateWithFlags");
    for (hipEvent_t &e : p->ev) hip_check(hipEventCreate(&e), "hipEventCreate");
    for (DevBuf *b : {&p->theta, &p->m, &p->v, &p->grad, &p->wq}) b->reserve(kNumParams * sizeof(float));
    p->pk.reserve(2 * 8 * 4 * 64 * sizeof(v8bf));
    p->state.reserve(sizeof(long long));
    // Glorot-uniform weights from a splitmix64 stream, zero biases
    std::vector<float> th(kNumParams, 0.0f);
    unsigned long long st = seed;
    auto uni = [&st]() {
      st += 0x9E3779B97F4A7C15ULL;
      unsigned long long z = st;
      z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
      z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
      z ^= z >> 31;
      return (double)(z >> 11) * (1.0 / 9007199254740992.0);
    };
    auto fill = [&](int off, int fan_out, int fan_in) {
      const double a = std::sqrt(6.0 / (fan_in + fan_out));
      for (int k = 0; k < fan_out * fan_in; ++k) th[off + k] = (float)((2.0 * uni() - 1.0) * a);
    };
    fill(kOffW1, kWidth, 2);
    fill(kOffW2, kWidth, kWidth);
    fill(kOffW3, kWidth, kWidth);
    fill(kOffW4, 1, kWidth);
    std::vector<float> zero(kNumParams, 0.0f);
    hip_check(hipMemcpy(p->m.ptr, zero.data(), kNumParams * sizeof(float), hipMemcpyHostToDevice), "init");
    hip_check(hipMemcpy(p->v.ptr, zero.data(), kNumParams * sizeof(float), hipMemcpyHostToDevice), "init");
    upload_theta(p.get(), th.data());
    *out = p.release();
  });
}

int bk_mlp_destroy(bk_mlp *p) {
  return guarded([&] {
    if (!p) return;
    {
      DeviceGuard dg(p->device);
      if (p->stream) (void)hipStreamSynchronize(p->stream);
      for (DevBuf *b : {&p->theta, &p->m, &p->v, &p->grad, &p->wq, &p->pk, &p->state, &p->lr, &p->xin, &p->tin,
                        &p->X, &p->T, &p->Z1, &p->H1, &p->Z2, &p->H2, &p->Z3, &p->H3, &p->dZ3, &p->dZ2, &p->dZ1,
                        &p->dH, &p->sq, &p->dy, &p->part, &p->gpart, &p->slist, &p->flist, &p->cnt, &p->cell_err, &p->gx,
                        &p->gy, &p->exact, &p->fx, &p->fy, &p->fexact, &p->keys_a, &p->keys_b, &p->sort_tmp})
        b->release();
      for (hipEvent_t e : p->ev)
        if (e) (void)hipEventDestroy(e);
      if (p->stream) (void)hipStreamDestroy(p->stream);
    }
    delete p;
  });
}

int bk_mlp_set_params(bk_mlp *p, const float *theta, double lo, double hi, double mu, double sigma, double eps) {
  return guarded([&] {
    require_mlp(p);
    if (!theta) fail(BK_ERR_INVALID_ARG, "null parameter array");
    if (!(hi > lo) || !std::isfinite(lo) || !std::isfinite(hi)) fail(BK_ERR_INVALID_ARG, "need lo < hi");
    if (!(sigma > 0.0) || !std::isfinite(mu) || !(eps >= 0.0)) fail(BK_ERR_INVALID_ARG, "bad output normalisation");
    for (int k = 0; k < kNumParams; ++k)
      if (!std::isfinite(theta[k])) fail(BK_ERR_INVALID_ARG, "parameters must be finite");
    DeviceGuard dg(p->device);
    p->lo = lo;
    p->hi = hi;
    p->mu = mu;
    p->sigma = sigma;
    p->eps = eps;
    upload_theta(p, theta);
  });
}

int bk_mlp_get_params(const bk_mlp *p, float *theta, double *lo, double *hi, double *mu, double *sigma,
                      double *eps) {
  return guarded([&] {
    if (!p) fail(BK_ERR_INVALID_ARG, "null surrogate");
    DeviceGuard dg(p->device);
    if (theta) {
      hip_check(hipStreamSynchronize(p->stream), "hipStreamSynchronize");
      hip_check(hipMemcpy(theta, p->theta.ptr, kNumParams * sizeof(float), hipMemcpyDeviceToHost), "copy parameters");
    }
    if (lo) *lo = p->lo;
    if (hi) *hi = p->hi;
    if (mu) *mu = p->mu;
    if (sigma) *sigma = p->sigma;
    if (eps) *eps = p->eps;
  });
}

}  // extern "C"

namespace {

// Checkpoint file (little-endian):
//   char[8] magic "BKMLPCKP", uint32 format version, uint32 parameter count,
//   uint32 width, uint32 layers, int64 Adam steps taken,
//   double lo, hi, mu, sigma, eps,
//   float theta[P] (FP32 master), float wq[P] (BF16-rounded working copy),
//   float m[P], float v[P] (Adam moments),
//   uint64 FNV-1a 64 of every byte before it.
constexpr char kCkptMagic[8] = {'B', 'K', 'M', 'L', 'P', 'C', 'K', 'P'};
constexpr uint32_t kCkptVersion = 1;
constexpr uint32_t kLayers = 4;

uint64_t fnv1a(const std::vector<char> &b) {
  uint64_t h = 0xcbf29ce484222325ULL;
  for (char c : b) {
    h ^= (uint64_t)(unsigned char)c;
    h *= 0x100000001b3ULL;
  }
  return h;
}

template <typename T>
void put(std::vector<char> &b, const T &v) {
  const char *c = reinterpret_cast<const char *>(&v);
  b.insert(b.end(), c, c + sizeof(T));
}

template <typename T>
T take(const std::vector<char> &b, size_t &off) {
  if (off + sizeof(T) > b.size()) fail(BK_ERR_INVALID_ARG, "checkpoint truncated");
  T v;
  std::memcpy(&v, b.data() + off, sizeof(T));
  off += sizeof(T);
  return v;
}

void copy_params(DevBuf &d, std::vector<float> &h) {
  h.resize(kNumParams);
  hip_check(hipMemcpy(h.data(), d.ptr, kNumParams * sizeof(float), hipMemcpyDeviceToHost), "copy state");
}

}  // namespace

extern "C" {

int bk_mlp_save(const bk_mlp *cp, const char *path) {
  return guarded([&] {
    if (!cp) fail(BK_ERR_INVALID_ARG, "null surrogate");
    if (!path || !*path) fail(BK_ERR_INVALID_ARG, "null or empty path");
    bk_mlp *p = const_cast<bk_mlp *>(cp);
    DeviceGuard dg(p->device);
    hip_check(hipStreamSynchronize(p->stream), "hipStreamSynchronize");
    std::vector<float> theta, wq, m, v;
    copy_params(p->theta, theta);
    copy_params(p->wq, wq);
    copy_params(p->m, m);
    copy_params(p->v, v);
    std::vector<char> b;
    b.insert(b.end(), kCkptMagic, kCkptMagic + 8);
    put(b, kCkptVersion);
    put(b, (uint32_t)kNumParams);
    put(b, (uint32_t)kWidth);
    put(b, kLayers);
    put(b, (int64_t)p->steps_done);
    for (double x : {p->lo, p->hi, p->mu, p->sigma, p->eps}) put(b, x);
    for (const std::vector<float> *a : {&theta, &wq, &m, &v})
      b.insert(b.end(), reinterpret_cast<const char *>(a->data()),
               reinterpret_cast<const char *>(a->data() + kNumParams));
    put(b, fnv1a(b));
    const std::string tmp = std::string(path) + ".tmp";
    {
      std::ofstream out(tmp, std::ios::binary | std::ios::trunc);
      if (!out) fail(BK_ERR_INVALID_ARG, std::string("cannot write ") + tmp);
      out.write(b.data(), (std::streamsize)b.size());
      if (!out) fail(BK_ERR_INVALID_ARG, std::string("cannot write ") + tmp);
    }
    if (std::rename(tmp.c_str(), path) != 0) fail(BK_ERR_INVALID_ARG, std::string("cannot replace ") + path);
  });
}

int bk_mlp_load(bk_mlp *p, const char *path) {
  return guarded([&] {
    require_mlp(p);
    if (!path || !*path) fail(BK_ERR_INVALID_ARG, "null or empty path");
    std::ifstream in(path, std::ios::binary);
    if (!in) fail(BK_ERR_INVALID_ARG, std::string("cannot read ") + path);
    std::vector<char> b((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
    if (b.size() < 8 + sizeof(uint64_t) || std::memcmp(b.data(), kCkptMagic, 8) != 0)
      fail(BK_ERR_INVALID_ARG, "not a surrogate checkpoint");
    size_t off = b.size() - sizeof(uint64_t);
    const uint64_t sum = take<uint64_t>(b, off);
    b.resize(b.size() - sizeof(uint64_t));
    if (fnv1a(b) != sum) fail(BK_ERR_INVALID_ARG, "checkpoint checksum mismatch (corrupted file)");
    off = 8;
    if (take<uint32_t>(b, off) != kCkptVersion) fail(BK_ERR_INVALID_ARG, "checkpoint format version not supported");
    if (take<uint32_t>(b, off) != (uint32_t)kNumParams || take<uint32_t>(b, off) != (uint32_t)kWidth ||
        take<uint32_t>(b, off) != kLayers)
      fail(BK_ERR_INVALID_ARG, "checkpoint of another network shape");
    const int64_t steps = take<int64_t>(b, off);
    double nv[5];
    for (double &x : nv) x = take<double>(b, off);
    if (b.size() - off != 4 * (size_t)kNumParams * sizeof(float)) fail(BK_ERR_INVALID_ARG, "checkpoint size mismatch");
    std::vector<float> a[4];
    for (auto &x : a) {
      x.resize(kNumParams);
      std::memcpy(x.data(), b.data() + off, kNumParams * sizeof(float));
      off += kNumParams * sizeof(float);
    }
    if (steps < 0) fail(BK_ERR_INVALID_ARG, "bad step count in checkpoint");
    if (!(nv[1] > nv[0]) || !(nv[3] > 0.0)) fail(BK_ERR_INVALID_ARG, "bad normalisation in checkpoint");
    DeviceGuard dg(p->device);
    hip_check(hipStreamSynchronize(p->stream), "hipStreamSynchronize");
    std::vector<float> before;
    copy_params(p->theta, before);
    upload_theta(p, a[0].data());
    // the working copy follows from the master weights; a file whose stored
    // copy differs was written under another rounding rule (the surrogate is
    // left as it was)
    std::vector<float> wq;
    copy_params(p->wq, wq);
    if (std::memcmp(wq.data(), a[1].data(), kNumParams * sizeof(float)) != 0) {
      upload_theta(p, before.data());
      fail(BK_ERR_INVALID_ARG, "checkpoint working copy does not match its master weights");
    }
    hip_check(hipMemcpy(p->m.ptr, a[2].data(), kNumParams * sizeof(float), hipMemcpyHostToDevice), "restore m");
    hip_check(hipMemcpy(p->v.ptr, a[3].data(), kNumParams * sizeof(float), hipMemcpyHostToDevice), "restore v");
    p->steps_done = (long long)steps;
    p->lo = nv[0];
    p->hi = nv[1];
    p->mu = nv[2];
    p->sigma = nv[3];
    p->eps = nv[4];
  });
}

int bk_mlp_set_training_data(bk_mlp *p, size_t n, const double *q, const double *f, double lo, double hi,
                             double eps) {
  return guarded([&] {
    require_mlp(p);
    if (n == 0 || !q || !f) fail(BK_ERR_INVALID_ARG, "need a non-empty training set");
    if (!(hi > lo) || !std::isfinite(lo) || !std::isfinite(hi)) fail(BK_ERR_INVALID_ARG, "need lo < hi");
    if (!(eps > 0.0) || !std::isfinite(eps)) fail(BK_ERR_INVALID_ARG, "eps must be positive");
    DeviceGuard dg(p->device);
    // targets t = (ln(f + eps) - mu) / sigma, inputs mapped onto [-1, 1]^2
    std::vector<double> t(n);
    double s1 = 0.0;
    for (size_t i = 0; i < n; ++i) {
      if (!(f[i] >= 0.0) || !std::isfinite(f[i])) fail(BK_ERR_INVALID_ARG, "densities must be finite and >= 0");
      t[i] = std::log(f[i] + eps);
      s1 += t[i];
    }
    const double mu = s1 / (double)n;
    double s2 = 0.0;
    for (size_t i = 0; i < n; ++i) s2 += (t[i] - mu) * (t[i] - mu);
    const double sigma = std::max(std::sqrt(s2 / (double)n), 1e-6);
    const double c = 0.5 * (lo + hi), is = 2.0 / (hi - lo);
    std::vector<float2> x(n);
    std::vector<float> tf(n);
    for (size_t i = 0; i < n; ++i) {
      x[i] = make_float2((float)((q[2 * i] - c) * is), (float)((q[2 * i + 1] - c) * is));
      tf[i] = (float)((t[i] - mu) / sigma);
    }
    p->xin.reserve(n * sizeof(float2));
    p->tin.reserve(n * sizeof(float));
    hip_check(hipMemcpy(p->xin.ptr, x.data(), n * sizeof(float2), hipMemcpyHostToDevice), "upload inputs");
    hip_check(hipMemcpy(p->tin.ptr, tf.data(), n * sizeof(float), hipMemcpyHostToDevice), "upload targets");
    p->n_train = (long long)n;
    p->lo = lo;
    p->hi = hi;
    p->mu = mu;
    p->sigma = sigma;
    p->eps = eps;
  });
}

int bk_mlp_train(bk_mlp *p, size_t steps, size_t batch, double lr_max, double lr_min, unsigned long long seed,
                 double *mean_loss_out, double *ms_out) {
  return guarded([&] {
    require_mlp(p);
    if (p->n_train == 0) fail(BK_ERR_STATE, "no training data");
    if (batch < 256 || batch > (size_t)kMaxBatch || batch % 256 != 0)
      fail(BK_ERR_INVALID_ARG, "batch must be a multiple of 256 between 256 and 65536");
    if (!(lr_max > 0.0) || !(lr_min >= 0.0) || lr_min > lr_max) fail(BK_ERR_INVALID_ARG, "bad learning rates");
    if (steps == 0) return;
    DeviceGuard dg(p->device);
    const int B = (int)batch;
    ensure_workspace(p, B);
    // cosine schedule from lr_max to lr_min over this call's steps
    std::vector<float> lr(steps);
    for (size_t k = 0; k < steps; ++k)
      lr[k] = (float)(lr_min + 0.5 * (lr_max - lr_min) * (1.0 + std::cos(M_PI * (double)k / (double)steps)));
    p->lr.reserve(steps * sizeof(float));
    hip_check(hipMemcpy(p->lr.ptr, lr.data(), steps * sizeof(float), hipMemcpyHostToDevice), "upload schedule");
    // the device step counter indexes the schedule (reset per call); Adam's
    // bias correction continues from the steps already taken
    const long long zero = 0;
    hip_check(hipMemcpy(p->state.ptr, &zero, sizeof(long long), hipMemcpyHostToDevice), "reset step");
    const hipStream_t s = p->stream;
    double loss_sum = 0.0;
    long long loss_n = 0;
    std::vector<float> sq(B);
    hip_check(hipEventRecord(p->ev[0], s), "hipEventRecord");
    for (size_t k = 0; k < steps; ++k) {
      hipLaunchKernelGGL(mlp_gather_kernel, dim3(grid_for(B, 256)), dim3(256), 0, s, p->xin.as<float2>(),
                         p->tin.as<float>(), p->n_train, p->state.as<long long>(), seed, B, p->X.as<float2>(),
                         p->T.as<float>());
      hip_check(hipGetLastError(), "gather launch");
      forward_backward(p, B);
      hipLaunchKernelGGL(mlp_adam_kernel, dim3(grid_for(kNumParams, 256)), dim3(256), 0, s, p->theta.as<float>(),
                         p->m.as<float>(), p->v.as<float>(), p->grad.as<float>(), p->wq.as<float>(),
                         p->state.as<long long>(), p->steps_done, p->lr.as<float>(), 0.9f, 0.999f, 1e-8f);
      hip_check(hipGetLastError(), "Adam launch");
      hipLaunchKernelGGL(mlp_step_kernel, dim3(1), dim3(64), 0, s, p->state.as<long long>());
      hip_check(hipGetLastError(), "step launch");
      if (k + 1 == steps || (steps > 16 && k + 16 >= steps)) {
        hip_check(hipMemcpyAsync(sq.data(), p->sq.ptr, B * sizeof(float), hipMemcpyDeviceToHost, s), "copy loss");
        hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
        for (int b = 0; b < B; ++b) loss_sum += sq[b];
        loss_n += B;
      }
    }
    hip_check(hipEventRecord(p->ev[1], s), "hipEventRecord");
    hip_check(hipEventSynchronize(p->ev[1]), "hipEventSynchronize");
    p->steps_done += (long long)steps;
    hipLaunchKernelGGL(mlp_pack_kernel, dim3(grid_for(2 * 8 * 4 * 64, 256)), dim3(256), 0, s, p->wq.as<float>(),
                       p->pk.as<v8bf>());
    hip_check(hipGetLastError(), "pack launch");
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    if (mean_loss_out) *mean_loss_out = loss_n ? loss_sum / (double)loss_n : std::nan("");
    if (ms_out) *ms_out = elapsed(p->ev[0], p->ev[1]);
  });
}

int bk_mlp_loss_grad(bk_mlp *p, size_t n, const double *x, const double *t, double *loss_out, float *grad_out) {
  return guarded([&] {
    require_mlp(p);
    if (n == 0 || n % 256 != 0 || n > (size_t)kMaxBatch)
      fail(BK_ERR_INVALID_ARG, "n must be a multiple of 256 up to 65536");
    if (!x || !t) fail(BK_ERR_INVALID_ARG, "null input array");
    DeviceGuard dg(p->device);
    const int B = (int)n;
    ensure_workspace(p, B);
    std::vector<float2> xf(n);
    std::vector<float> tf(n);
    for (size_t i = 0; i < n; ++i) {
      xf[i] = make_float2((float)x[2 * i], (float)x[2 * i + 1]);
      tf[i] = (float)t[i];
    }
    hip_check(hipMemcpy(p->X.ptr, xf.data(), n * sizeof(float2), hipMemcpyHostToDevice), "upload inputs");
    hip_check(hipMemcpy(p->T.ptr, tf.data(), n * sizeof(float), hipMemcpyHostToDevice), "upload targets");
    forward_backward(p, B);
    hip_check(hipStreamSynchronize(p->stream), "hipStreamSynchronize");
    std::vector<float> sq(n);
    hip_check(hipMemcpy(sq.data(), p->sq.ptr, n * sizeof(float), hipMemcpyDeviceToHost), "copy loss");
    double l = 0.0;
    for (float v : sq) l += v;
    if (loss_out) *loss_out = l / (double)n;
    if (grad_out)
      hip_check(hipMemcpy(grad_out, p->grad.ptr, kNumParams * sizeof(float), hipMemcpyDeviceToHost), "copy gradient");
  });
}

int bk_mlp_forward(bk_mlp *p, size_t n, const double *x, double *y_out) {
  return guarded([&] {
    require_mlp(p);
    if (!x || !y_out) fail(BK_ERR_INVALID_ARG, "null array");
    if (n == 0) return;
    DeviceGuard dg(p->device);
    // inputs already on the network's [-1, 1] scale: map them back to query
    // coordinates so the serving kernel's own input map applies
    const double c = 0.5 * (p->lo + p->hi), hs = 0.5 * (p->hi - p->lo);
    std::vector<float> qx(n), qy(n), y(n);
    for (size_t i = 0; i < n; ++i) {
      qx[i] = (float)(c + hs * x[2 * i]);
      qy[i] = (float)(c + hs * x[2 * i + 1]);
    }
    DevBuf dx, dyb, dout;
    dx.reserve(n * 4);
    dyb.reserve(n * 4);
    dout.reserve(n * 4);
    hip_check(hipMemcpy(dx.ptr, qx.data(), n * 4, hipMemcpyHostToDevice), "upload");
    hip_check(hipMemcpy(dyb.ptr, qy.data(), n * 4, hipMemcpyHostToDevice), "upload");
    launch_infer(p, dx.as<float>(), dyb.as<float>(), (long long)n, nullptr, dout.as<float>());
    hip_check(hipStreamSynchronize(p->stream), "hipStreamSynchronize");
    hip_check(hipMemcpy(y.data(), dout.ptr, n * 4, hipMemcpyDeviceToHost), "copy outputs");
    for (size_t i = 0; i < n; ++i) y_out[i] = y[i];
  });
}

int bk_mlp_infer_device(bk_mlp *p, size_t n, const float *qx, const float *qy, double *out) {
  return guarded([&] {
    require_mlp(p);
    if (n == 0) return;
    if (!qx || !qy || !out) fail(BK_ERR_INVALID_ARG, "null array");
    DeviceGuard dg(p->device);
    launch_infer(p, qx, qy, (long long)n, out, nullptr);
    hip_check(hipStreamSynchronize(p->stream), "hipStreamSynchronize");
  });
}

int bk_mlp_serve_device(bk_mlp *p, bk_ctx *ctx, const bk_kernel_desc *desc, size_t n, const float *qx,
                        const float *qy, double *out, double sample_fraction, double rel_tol, double floor_density,
                        int cells_per_axis, unsigned long long seed, bk_serve_stats *stats) {
  return guarded([&] {
    require_mlp(p);
    if (!ctx || !desc) fail(BK_ERR_INVALID_ARG, "null context or descriptor");
    if (!qx || !qy || !out) fail(BK_ERR_INVALID_ARG, "null array");
    if (!(sample_fraction > 0.0 && sample_fraction <= 1.0)) fail(BK_ERR_INVALID_ARG, "sample_fraction in (0, 1]");
    if (!(rel_tol > 0.0) || !(floor_density > 0.0)) fail(BK_ERR_INVALID_ARG, "rel_tol and floor must be positive");
    if (cells_per_axis < 1 || cells_per_axis > 4096) fail(BK_ERR_INVALID_ARG, "cells_per_axis in [1, 4096]");
    if (bk::ctx_device(ctx) != p->device) fail(BK_ERR_INVALID_ARG, "context and surrogate on different devices");
    DeviceGuard dg(p->device);
    bk_serve_stats st{};
    st.n = n;
    if (n == 0) {
      if (stats) *stats = st;
      return;
    }
    const hipStream_t s = p->stream;
    const int G = cells_per_axis;
    const long long N = (long long)n;
    const unsigned long long period = (unsigned long long)std::max(1.0, std::floor(1.0 / sample_fraction + 0.5));
    const long long scap = std::min<long long>(N, N / (long long)period + 64 + 8 * (long long)std::sqrt((double)N));
    p->slist.reserve((size_t)scap * sizeof(long long));
    p->flist.reserve((size_t)N * sizeof(long long));
    p->cnt.reserve(8 * sizeof(int));
    p->cell_err.reserve((size_t)(2 * G * G + 1) * sizeof(unsigned int));
    p->gx.reserve((size_t)N * sizeof(float));
    p->gy.reserve((size_t)N * sizeof(float));
    p->exact.reserve((size_t)N * sizeof(double));
    const float lo = (float)p->lo, inv_w = (float)(G / (p->hi - p->lo));
    int *cnt = p->cnt.as<int>();
    unsigned int *cerr = p->cell_err.as<unsigned int>();
    // work starts after everything queued on the caller's stream
    hip_check(hipStreamSynchronize(bk::ctx_stream(ctx)), "hipStreamSynchronize");
    hip_check(hipEventRecord(p->ev[0], s), "hipEventRecord");
    unsigned int *ccnt = cerr + G * G + 1;
    hipLaunchKernelGGL(zero_u32_kernel, dim3(grid_for(2 * G * G + 1, 256)), dim3(256), 0, s, cerr, 2 * G * G + 1);
    hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(64), 0, s, reinterpret_cast<unsigned int *>(cnt), 8);
    hip_check(hipGetLastError(), "zero launch");
    launch_infer(p, qx, qy, N, out, nullptr);
    hip_check(hipEventRecord(p->ev[1], s), "hipEventRecord");
    // sampled check
    hipLaunchKernelGGL(select_samples_kernel, dim3(grid_for(N, 256)), dim3(256), 0, s, N, seed, period, cnt,
                       p->slist.as<long long>(), scap);
    hip_check(hipGetLastError(), "sample launch");
    int ns = 0;
    hip_check(hipMemcpyAsync(&ns, cnt, sizeof(int), hipMemcpyDeviceToHost, s), "read sample count");
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    if (ns > scap) fail(BK_ERR_STATE, "internal: sample list overflow");
    // Z-curve order: each workgroup of the exact sum gets nearby queries and
    // skips the reference tiles beyond the kernel's reach; the list is sorted
    // before the gather, so exact[k] belongs to query slist[k]
    if (N < (1LL << 32))
      bk::spatial_sort(qx, qy, p->slist.as<long long>(), ns, N, (float)p->lo, (float)p->hi, p->slist.as<long long>(),
                       p->keys_a, p->keys_b, p->sort_tmp, s);
    hipLaunchKernelGGL(gather_kernel, dim3(grid_for(scap, 256)), dim3(256), 0, s, qx, qy, p->slist.as<long long>(),
                       cnt, scap, p->gx.as<float>(), p->gy.as<float>());
    hip_check(hipGetLastError(), "gather launch");
    bk::eval_queries_dev(ctx, desc, p->gx.as<float>(), p->gy.as<float>(), (size_t)ns, p->exact.as<double>(), s);
    hipLaunchKernelGGL(check_samples_kernel, dim3(grid_for(std::max(ns, 1), 256)), dim3(256), 0, s, qx, qy,
                       p->slist.as<long long>(), cnt, scap, p->exact.as<double>(), out, floor_density, lo, inv_w, G,
                       cerr, ccnt, cerr + G * G);
    hip_check(hipGetLastError(), "check launch");
    hip_check(hipEventRecord(p->ev[2], s), "hipEventRecord");
    // fallback of failing cells
    hipLaunchKernelGGL(select_fallback_kernel, dim3(grid_for(N, 256)), dim3(256), 0, s, qx, qy, N, cerr, ccnt,
                       (float)rel_tol, lo, inv_w, G, seed, period, cnt + 1, p->flist.as<long long>(), N);
    hip_check(hipGetLastError(), "fallback selection launch");
    int nf = 0;
    unsigned int maxbits = 0;
    hip_check(hipMemcpyAsync(&nf, cnt + 1, sizeof(int), hipMemcpyDeviceToHost, s), "read fallback count");
    hip_check(hipMemcpyAsync(&maxbits, cerr + G * G, sizeof(unsigned int), hipMemcpyDeviceToHost, s), "read error");
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    if (nf > 0) {
      if (N < (1LL << 32))
        bk::spatial_sort(qx, qy, p->flist.as<long long>(), nf, N, (float)p->lo, (float)p->hi, p->flist.as<long long>(),
                       p->keys_a, p->keys_b, p->sort_tmp, s);
      hipLaunchKernelGGL(gather_kernel, dim3(grid_for(nf, 256)), dim3(256), 0, s, qx, qy, p->flist.as<long long>(),
                         cnt + 1, N, p->gx.as<float>(), p->gy.as<float>());
      hip_check(hipGetLastError(), "gather launch");
      bk::eval_queries_dev(ctx, desc, p->gx.as<float>(), p->gy.as<float>(), (size_t)nf, p->exact.as<double>(), s);
      hipLaunchKernelGGL(scatter_kernel, dim3(grid_for(nf, 256)), dim3(256), 0, s, p->flist.as<long long>(), cnt + 1,
                         N, p->exact.as<double>(), out);
      hip_check(hipGetLastError(), "scatter launch");
    }
    hip_check(hipEventRecord(p->ev[3], s), "hipEventRecord");
    hip_check(hipEventSynchronize(p->ev[3]), "hipEventSynchronize");
    std::vector<unsigned int> ce(2 * G * G + 1);
    hip_check(hipMemcpy(ce.data(), cerr, ce.size() * sizeof(unsigned int), hipMemcpyDeviceToHost), "copy cell errors");
    long long failed = 0;
    for (int c = 0; c < G * G; ++c) {
      float e;
      std::memcpy(&e, &ce[c], 4);
      if (e > (float)rel_tol || ce[G * G + 1 + c] == 0u) ++failed;
    }
    float maxe;
    std::memcpy(&maxe, &maxbits, 4);
    st.n_samples = (size_t)ns;
    st.n_fallback = (size_t)nf;
    st.cells = (size_t)G * G;
    st.cells_failed = (size_t)failed;
    st.max_sample_err = maxe;
    st.ms_infer = elapsed(p->ev[0], p->ev[1]);
    st.ms_check = elapsed(p->ev[1], p->ev[2]);
    st.ms_fallback = elapsed(p->ev[2], p->ev[3]);
    st.ms_total = elapsed(p->ev[0], p->ev[3]);
    if (stats) *stats = st;
  });
}

}  // extern "C"
