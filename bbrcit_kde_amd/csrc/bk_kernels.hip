// HIP kernels of the direct kernel-density sum for gfx950 (MI355X).
//
// The reference runs one thread per query and walks block-sized tiles of
// AoS reference points in shared memory, accumulating
// sum += m * U(q, r; h*abw) in FP32 (include/CudaDirectKde.h:126-173).
// Here the same sum is organised for CDNA4:
//   * queries are read once, coalesced, from SoA arrays and kept in
//     registers, kQpt per thread;
//   * references are pre-transformed once per kernel descriptor into
//     per-point feature vectors (scaled bandwidth factor, mass term), so the
//     per-pair division by h^2 a^2 of the reference becomes a multiply;
//   * a workgroup stages kTile feature vectors in LDS (double-buffered, one
//     barrier per stage); every lane of a wavefront reads the same entry, so
//     each LDS read is a broadcast;
//   * the Gaussian term m*exp(-c*d2/(h a)^2) is evaluated as
//     exp2(log2(m/m_scale) - d2 * w) with one v_exp_f32;
//   * per-query sums are FP32 within one LDS stage (<= kTile terms) and are
//     then added into an FP64 accumulator (block compensation);
//   * large reference ranges are split over workgroups, with a fixed-order
//     FP64 reduction of the partial sums, so results are deterministic.
#include "bk_internal.h"

#include <hip/hip_runtime.h>

namespace bk {

namespace {

constexpr float kMaskedLog = -1.0e30f;  // log2 mass of a masked reference

__device__ __forceinline__ float clamp01(float v) {
  return __builtin_fminf(__builtin_fmaxf(v, 0.0f), 1.0f);
}

template <int FAM>
struct Fam {
  static constexpr int kNf = (FAM == FAM_EPANPROD || FAM == FAM_EPANCONV) ? 2 : 1;
};

// Feature vector of a reference that must contribute nothing.
template <int FAM>
__device__ __forceinline__ float4 masked_f0() {
  if constexpr (FAM == FAM_GAUSS) {
    return make_float4(0.f, 0.f, 0.f, kMaskedLog);
  } else {
    return make_float4(0.f, 0.f, 0.f, 0.f);
  }
}

// Contribution m_r * U(q - r) of one reference with features (a, b).
//   FAM_GAUSS:    a = (x, y, w, log2(m/m_scale))     w = c log2(e) / (hx a_r)^2
//   FAM_EPAN:     a = (x, y, w, m/m_scale)           w = 1 / (h a_r)^2
//   FAM_EPANPROD: a = (x, y, wx, wy), b.x = m/m_scale
//   FAM_EPANCONV: a = (x, y, 1/(hx a_r), 1/(hy a_r)), b.x = m/m_scale * 2^-10
// rho = (weight of y) / (weight of x) for the quadratic forms.
template <int FAM, bool ISO>
__device__ __forceinline__ float pair_term(float qx, float qy, float4 a,
                                           float4 b, float rho) {
  const float dx = qx - a.x;
  const float dy = qy - a.y;
  if constexpr (FAM == FAM_GAUSS) {
    const float d2 = ISO ? __builtin_fmaf(dy, dy, dx * dx)
                         : __builtin_fmaf(dy * rho, dy, dx * dx);
    return __builtin_amdgcn_exp2f(__builtin_fmaf(-a.z, d2, a.w));
  } else if constexpr (FAM == FAM_EPAN) {
    const float d2 = ISO ? __builtin_fmaf(dy, dy, dx * dx)
                         : __builtin_fmaf(dy * rho, dy, dx * dx);
    return a.w * clamp01(__builtin_fmaf(-a.z, d2, 1.0f));
  } else if constexpr (FAM == FAM_EPANPROD) {
    const float tx = clamp01(__builtin_fmaf(-a.z * dx, dx, 1.0f));
    const float ty = clamp01(__builtin_fmaf(-a.w * dy, dy, 1.0f));
    return b.x * (tx * ty);
  } else {
    const float ux = __builtin_fabsf(dx) * a.z;
    const float uy = __builtin_fabsf(dy) * a.w;
    const float cx = __builtin_fmaxf(2.0f - ux, 0.0f);
    const float cy = __builtin_fmaxf(2.0f - uy, 0.0f);
    const float px = __builtin_fmaf(ux, ux + 6.0f, 4.0f);
    const float py = __builtin_fmaf(uy, uy + 6.0f, 4.0f);
    const float kx = cx * cx * cx * px;
    const float ky = cy * cy * cy * py;
    return b.x * (kx * ky);
  }
}

// One work item of the FP32 direct kernel: references [r_begin, r_end)
// against up to kQueriesPerBlock queries starting at q_begin.
template <int FAM, bool ISO>
__device__ __forceinline__ void direct_f32_item(
    const float *__restrict__ qx, const float *__restrict__ qy,
    const float4 *__restrict__ f0, const float4 *__restrict__ f1,
    long long r_begin, long long r_end, long long q_begin, int q_count,
    double *__restrict__ dst, double out_scale, int self_excl, float rho) {
  constexpr int kNf = Fam<FAM>::kNf;
  __shared__ float4 s0[2][kTile];
  __shared__ float4 s1[2][kNf == 2 ? kTile : 1];

  const int t = threadIdx.x;
  float px[kQpt], py[kQpt];
  double acc[kQpt];
#pragma unroll
  for (int k = 0; k < kQpt; ++k) {
    const int qi = t + k * kBlock;
    const long long q = q_begin + (qi < q_count ? qi : 0);
    px[k] = qx[q];
    py[k] = qy[q];
    acc[k] = 0.0;
  }

  const float4 m0 = masked_f0<FAM>();
  const float4 m1 = make_float4(0.f, 0.f, 0.f, 0.f);
  float4 n0 = m0, n1 = m1;
  {
    const long long r = r_begin + t;
    if (r < r_end) {
      n0 = f0[r];
      if constexpr (kNf == 2) n1 = f1[r];
    }
  }

  // The self pair (i, i) can only occur in stages whose reference range
  // overlaps this workgroup's query range.
  const long long q_lo = q_begin, q_hi = q_begin + q_count;

  int buf = 0;
  for (long long tile = r_begin; tile < r_end; tile += kTile) {
    s0[buf][t] = n0;
    if constexpr (kNf == 2) s1[buf][t] = n1;
    __syncthreads();
    {
      const long long r = tile + kTile + t;
      n0 = m0;
      n1 = m1;
      if (r < r_end) {
        n0 = f0[r];
        if constexpr (kNf == 2) n1 = f1[r];
      }
    }
    const int cnt = (int)((r_end - tile) < kTile ? (r_end - tile) : kTile);
    float s[kQpt];
#pragma unroll
    for (int k = 0; k < kQpt; ++k) s[k] = 0.0f;

    const bool diag = self_excl && tile < q_hi && q_lo < tile + cnt;
    if (!diag) {
#pragma unroll 4
      for (int j = 0; j < cnt; ++j) {
        const float4 a = s0[buf][j];
        const float4 b = (kNf == 2) ? s1[buf][j] : a;
#pragma unroll
        for (int k = 0; k < kQpt; ++k) s[k] += pair_term<FAM, ISO>(px[k], py[k], a, b, rho);
      }
    } else {
      for (int j = 0; j < cnt; ++j) {
        const float4 a = s0[buf][j];
        const float4 b = (kNf == 2) ? s1[buf][j] : a;
        const long long rj = tile + j;
#pragma unroll
        for (int k = 0; k < kQpt; ++k) {
          const float v = pair_term<FAM, ISO>(px[k], py[k], a, b, rho);
          s[k] += (rj == q_begin + t + k * kBlock) ? 0.0f : v;
        }
      }
    }
#pragma unroll
    for (int k = 0; k < kQpt; ++k) acc[k] += (double)s[k];
    buf ^= 1;
  }

#pragma unroll
  for (int k = 0; k < kQpt; ++k) {
    const int qi = t + k * kBlock;
    if (qi < q_count) dst[qi] = acc[k] * out_scale;
  }
}

template <int FAM, bool ISO>
__global__ __launch_bounds__(kBlock) void direct_f32_range_kernel(
    const float *__restrict__ qx, const float *__restrict__ qy,
    const float4 *__restrict__ f0, const float4 *__restrict__ f1,
    long long r_i, long long r_end, long long split_len, long long q_i,
    long long nq, double *__restrict__ dst, double out_scale, int self_excl,
    float rho) {
  const long long qb = (long long)blockIdx.x * kQueriesPerBlock;
  const long long rb = r_i + (long long)blockIdx.y * split_len;
  long long re = rb + split_len;
  if (re > r_end) re = r_end;
  const long long left = nq - qb;
  const int q_count = (int)(left < kQueriesPerBlock ? left : kQueriesPerBlock);
  direct_f32_item<FAM, ISO>(qx, qy, f0, f1, rb, re, q_i + qb, q_count,
                            dst + (long long)blockIdx.y * nq + qb, out_scale,
                            self_excl, rho);
}

template <int FAM, bool ISO>
__global__ __launch_bounds__(kBlock) void direct_f32_list_kernel(
    const float *__restrict__ qx, const float *__restrict__ qy,
    const float4 *__restrict__ f0, const float4 *__restrict__ f1,
    const WorkItem *__restrict__ items, double *__restrict__ dst,
    double out_scale, int self_excl, float rho) {
  const WorkItem it = items[blockIdx.x];
  direct_f32_item<FAM, ISO>(qx, qy, f0, f1, it.r_begin, it.r_end, it.q_begin,
                            it.q_count, dst + it.out_index, out_scale,
                            self_excl, rho);
}

// Per-reference features, computed in FP64 from the FP32 point data and
// rounded once.
__global__ void prep_f32_kernel(EvalParams p, const float *__restrict__ x,
                                const float *__restrict__ y,
                                const float *__restrict__ m,
                                const float *__restrict__ abw, long long n,
                                double inv_m_scale, float4 *__restrict__ f0,
                                float4 *__restrict__ f1) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double a = (double)abw[i];
  const double ia2 = 1.0 / (a * a);
  const double mass = (double)m[i] * inv_m_scale;
  const float xi = x[i], yi = y[i];
  switch (p.family) {
    case FAM_GAUSS: {
      const float lm = mass > 0.0 ? (float)log2(mass) : kMaskedLog;
      f0[i] = make_float4(xi, yi, (float)(p.gx * ia2), lm);
      break;
    }
    case FAM_EPAN:
      f0[i] = make_float4(xi, yi, (float)(p.gx * ia2), (float)mass);
      break;
    case FAM_EPANPROD:
      f0[i] = make_float4(xi, yi, (float)(p.gx * ia2), (float)(p.gy * ia2));
      f1[i] = make_float4((float)mass, 0.f, 0.f, 0.f);
      break;
    default:
      f0[i] = make_float4(xi, yi, (float)(p.ix / a), (float)(p.iy / a));
      f1[i] = make_float4((float)(mass * p.mass_fold), 0.f, 0.f, 0.f);
      break;
  }
}

// ---- FP64 path ---------------------------------------------------------
constexpr int kBlock64 = 256;
constexpr int kTile64 = 256;

template <int FAM>
__device__ __forceinline__ double pair_term_f64(double qx, double qy,
                                                double4 a, double4 b,
                                                double rho) {
  const double dx = qx - a.x;
  const double dy = qy - a.y;
  if constexpr (FAM == FAM_GAUSS) {
    // a = (x, y, c/(hx a_r)^2, m)
    return a.w * exp(-a.z * (dx * dx + rho * dy * dy));
  } else if constexpr (FAM == FAM_EPAN) {
    const double t = 1.0 - a.z * (dx * dx + rho * dy * dy);
    return a.w * (t > 0.0 ? t : 0.0);
  } else if constexpr (FAM == FAM_EPANPROD) {
    double tx = 1.0 - a.z * dx * dx;
    double ty = 1.0 - a.w * dy * dy;
    tx = tx > 0.0 ? tx : 0.0;
    ty = ty > 0.0 ? ty : 0.0;
    return b.x * tx * ty;
  } else {
    const double ux = fabs(dx) * a.z, uy = fabs(dy) * a.w;
    const double cx = ux < 2.0 ? 2.0 - ux : 0.0;
    const double cy = uy < 2.0 ? 2.0 - uy : 0.0;
    return b.x * (cx * cx * cx * (ux * ux + 6.0 * ux + 4.0)) *
           (cy * cy * cy * (uy * uy + 6.0 * uy + 4.0));
  }
}

template <int FAM>
__device__ __forceinline__ void direct_f64_item(
    const double *__restrict__ qx, const double *__restrict__ qy,
    const double4 *__restrict__ f0, const double4 *__restrict__ f1,
    long long r_begin, long long r_end, long long q_begin, int q_count,
    double *__restrict__ dst, double out_scale, int self_excl, double rho) {
  constexpr int kNf = Fam<FAM>::kNf;
  __shared__ double4 s0[kTile64];
  __shared__ double4 s1[kNf == 2 ? kTile64 : 1];
  const int t = threadIdx.x;
  const bool valid = t < q_count;
  const long long q = q_begin + (valid ? t : 0);
  const double px = qx[q], py = qy[q];
  double sum = 0.0;
  for (long long tile = r_begin; tile < r_end; tile += kTile64) {
    const long long r = tile + t;
    if (r < r_end) {
      s0[t] = f0[r];
      if constexpr (kNf == 2) s1[t] = f1[r];
    }
    __syncthreads();
    const int cnt = (int)((r_end - tile) < kTile64 ? (r_end - tile) : kTile64);
    for (int j = 0; j < cnt; ++j) {
      const double4 a = s0[j];
      const double4 b = (kNf == 2) ? s1[j] : a;
      const double v = pair_term_f64<FAM>(px, py, a, b, rho);
      if (!(self_excl && tile + j == q)) sum += v;
    }
    __syncthreads();
  }
  if (valid) dst[t] = sum * out_scale;
}

template <int FAM>
__global__ __launch_bounds__(kBlock64) void direct_f64_range_kernel(
    const double *__restrict__ qx, const double *__restrict__ qy,
    const double4 *__restrict__ f0, const double4 *__restrict__ f1,
    long long r_i, long long r_end, long long split_len, long long q_i,
    long long nq, double *__restrict__ dst, double out_scale, int self_excl,
    double rho) {
  const long long qb = (long long)blockIdx.x * kBlock64;
  const long long rb = r_i + (long long)blockIdx.y * split_len;
  long long re = rb + split_len;
  if (re > r_end) re = r_end;
  const long long left = nq - qb;
  const int q_count = (int)(left < kBlock64 ? left : kBlock64);
  direct_f64_item<FAM>(qx, qy, f0, f1, rb, re, q_i + qb, q_count,
                       dst + (long long)blockIdx.y * nq + qb, out_scale,
                       self_excl, rho);
}

template <int FAM>
__global__ __launch_bounds__(kBlock64) void direct_f64_list_kernel(
    const double *__restrict__ qx, const double *__restrict__ qy,
    const double4 *__restrict__ f0, const double4 *__restrict__ f1,
    const WorkItem *__restrict__ items, double *__restrict__ dst,
    double out_scale, int self_excl, double rho) {
  const WorkItem it = items[blockIdx.x];
  // The FP64 kernel covers kBlock64 queries per item.
  direct_f64_item<FAM>(qx, qy, f0, f1, it.r_begin, it.r_end, it.q_begin,
                       it.q_count, dst + it.out_index, out_scale, self_excl,
                       rho);
}

__global__ void prep_f64_kernel(EvalParams p, const double *__restrict__ x,
                                const double *__restrict__ y,
                                const double *__restrict__ m,
                                const double *__restrict__ abw, long long n,
                                double4 *__restrict__ f0,
                                double4 *__restrict__ f1) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double a = abw[i];
  const double ia2 = 1.0 / (a * a);
  switch (p.family) {
    case FAM_GAUSS:
      f0[i] = make_double4(x[i], y[i], p.c * p.gx * ia2, m[i]);
      break;
    case FAM_EPAN:
      f0[i] = make_double4(x[i], y[i], p.gx * ia2, m[i]);
      break;
    case FAM_EPANPROD:
      f0[i] = make_double4(x[i], y[i], p.gx * ia2, p.gy * ia2);
      f1[i] = make_double4(m[i], 0.0, 0.0, 0.0);
      break;
    default:
      f0[i] = make_double4(x[i], y[i], p.ix / a, p.iy / a);
      f1[i] = make_double4(m[i] * p.mass_fold, 0.0, 0.0, 0.0);
      break;
  }
}

// ---- general dimension (FP32, 3 <= D <= 16) -------------------------------
constexpr int kBlockG = 256;
constexpr int kTileG = 128;

template <int D, int FAM>
__device__ __forceinline__ void direct_general_item(
    const float *__restrict__ qc, const float *__restrict__ rc,
    const float2 *__restrict__ wm, long long r_begin, long long r_end,
    long long q_begin, int q_count, double *__restrict__ dst,
    double out_scale, int self_excl) {
  __shared__ float sc[kTileG * D];
  __shared__ float2 sw[kTileG];
  const int t = threadIdx.x;
  const bool valid = t < q_count;
  const long long q = q_begin + (valid ? t : 0);
  float pq[D];
#pragma unroll
  for (int d = 0; d < D; ++d) pq[d] = qc[q * D + d];
  double acc = 0.0;
  for (long long tile = r_begin; tile < r_end; tile += kTileG) {
    const int cnt = (int)((r_end - tile) < kTileG ? (r_end - tile) : kTileG);
    for (int e = t; e < cnt * D; e += kBlockG) sc[e] = rc[tile * D + e];
    for (int e = t; e < cnt; e += kBlockG) sw[e] = wm[tile + e];
    __syncthreads();
    float s = 0.0f;
    for (int j = 0; j < cnt; ++j) {
      float d2 = 0.0f;
#pragma unroll
      for (int d = 0; d < D; ++d) {
        const float df = pq[d] - sc[j * D + d];
        d2 = __builtin_fmaf(df, df, d2);
      }
      const float2 w = sw[j];
      float v;
      if constexpr (FAM == FAM_GAUSS) {
        v = __builtin_amdgcn_exp2f(__builtin_fmaf(-w.x, d2, w.y));
      } else {
        v = w.y * clamp01(__builtin_fmaf(-w.x, d2, 1.0f));
      }
      if (!(self_excl && tile + j == q)) s += v;
    }
    acc += (double)s;
    __syncthreads();
  }
  if (valid) dst[t] = acc * out_scale;
}

template <int D, int FAM>
__global__ __launch_bounds__(kBlockG) void direct_general_range_kernel(
    const float *__restrict__ qc, const float *__restrict__ rc,
    const float2 *__restrict__ wm, long long r_i, long long r_end,
    long long split_len, long long q_i, long long nq, double *__restrict__ dst,
    double out_scale, int self_excl) {
  const long long qb = (long long)blockIdx.x * kBlockG;
  const long long rb = r_i + (long long)blockIdx.y * split_len;
  long long re = rb + split_len;
  if (re > r_end) re = r_end;
  const long long left = nq - qb;
  const int q_count = (int)(left < kBlockG ? left : kBlockG);
  direct_general_item<D, FAM>(qc, rc, wm, rb, re, q_i + qb, q_count,
                              dst + (long long)blockIdx.y * nq + qb, out_scale,
                              self_excl);
}

template <int D, int FAM>
__global__ __launch_bounds__(kBlockG) void direct_general_list_kernel(
    const float *__restrict__ qc, const float *__restrict__ rc,
    const float2 *__restrict__ wm, const WorkItem *__restrict__ items,
    double *__restrict__ dst, double out_scale, int self_excl) {
  const WorkItem it = items[blockIdx.x];
  direct_general_item<D, FAM>(qc, rc, wm, it.r_begin, it.r_end, it.q_begin,
                              it.q_count, dst + it.out_index, out_scale,
                              self_excl);
}

__global__ void prep_general_kernel(EvalParams p, const float *__restrict__ abw,
                                    const float *__restrict__ m, long long n,
                                    double inv_m_scale,
                                    float2 *__restrict__ wm) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double a = (double)abw[i];
  const double w = p.gx / (a * a);
  const double mass = (double)m[i] * inv_m_scale;
  if (p.family == FAM_GAUSS) {
    wm[i] = make_float2((float)w, mass > 0.0 ? (float)log2(mass) : kMaskedLog);
  } else {
    wm[i] = make_float2((float)w, (float)mass);
  }
}

// ---- reductions ----------------------------------------------------------
__global__ void reduce_splits_kernel(const double *__restrict__ partial,
                                     int splits, long long nq,
                                     double *__restrict__ out) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nq) return;
  double s = 0.0;
  for (int k = 0; k < splits; ++k) s += partial[(long long)k * nq + q];
  out[q] = s;
}

constexpr int kReduceThreads = 256;

__device__ double block_sum(double v) {
  __shared__ double sh[kReduceThreads];
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int o = kReduceThreads / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  return sh[0];
}

__global__ __launch_bounds__(kReduceThreads) void weighted_reduce_kernel(
    int mode, const double *__restrict__ w, const double *__restrict__ u,
    const double *__restrict__ v, double a, double b, long long n,
    double *__restrict__ partial) {
  double s = 0.0;
  for (long long i = (long long)blockIdx.x * kReduceThreads + threadIdx.x; i < n;
       i += (long long)gridDim.x * kReduceThreads) {
    if (mode == 0) {
      s += w[i] * log(a * u[i]);
    } else {
      s += w[i] * (a * u[i] - b * v[i]);
    }
  }
  const double tot = block_sum(s);
  if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

__global__ __launch_bounds__(kReduceThreads) void final_reduce_kernel(
    const double *__restrict__ partial, int n, double *__restrict__ result) {
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += kReduceThreads) s += partial[i];
  const double tot = block_sum(s);
  if (threadIdx.x == 0) *result = tot;
}

__global__ void adapt_lambda_kernel(const double *__restrict__ S, long long n,
                                    double norm,
                                    const double *__restrict__ log_g,
                                    double alpha, double *__restrict__ lambda) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double g = exp(*log_g);
  lambda[i] = pow(norm * S[i] / g, -alpha);
}

long long split_length(long long nr, int splits, int tile) {
  long long len = (nr + splits - 1) / splits;
  len = (len + tile - 1) / tile * tile;
  return len;
}

}  // namespace

// ---- launchers -----------------------------------------------------------

hipError_t launch_prep_f32(const EvalParams &p, const float *x, const float *y,
                           const float *m, const float *abw, size_t n,
                           double m_scale, float4 *f0, float4 *f1,
                           hipStream_t s) {
  if (n == 0) return hipSuccess;
  const int th = 256;
  const unsigned blocks = (unsigned)((n + th - 1) / th);
  prep_f32_kernel<<<blocks, th, 0, s>>>(p, x, y, m, abw, (long long)n,
                                        1.0 / m_scale, f0, f1);
  return hipGetLastError();
}

hipError_t launch_direct_f32_range(const EvalParams &p, const float *qx,
                                   const float *qy, const float4 *f0,
                                   const float4 *f1, long long r_i,
                                   long long r_end, long long q_i,
                                   long long nq, int splits, double *dst,
                                   double out_scale, int self_excl,
                                   hipStream_t s) {
  if (nq <= 0) return hipSuccess;
  const long long len = split_length(r_end - r_i, splits, kTile);
  const dim3 grid((unsigned)((nq + kQueriesPerBlock - 1) / kQueriesPerBlock),
                  (unsigned)splits);
  const float rho = (float)(p.gy / p.gx);
  const bool iso = p.iso != 0;
#define BK_RANGE(F, I)                                                        \
  direct_f32_range_kernel<F, I><<<grid, kBlock, 0, s>>>(                      \
      qx, qy, f0, f1, r_i, r_end, len, q_i, nq, dst, out_scale, self_excl, rho)
  switch (p.family) {
    case FAM_GAUSS:
      if (iso) BK_RANGE(FAM_GAUSS, true); else BK_RANGE(FAM_GAUSS, false);
      break;
    case FAM_EPAN:
      if (iso) BK_RANGE(FAM_EPAN, true); else BK_RANGE(FAM_EPAN, false);
      break;
    case FAM_EPANPROD:
      BK_RANGE(FAM_EPANPROD, false);
      break;
    default:
      BK_RANGE(FAM_EPANCONV, false);
      break;
  }
#undef BK_RANGE
  return hipGetLastError();
}

hipError_t launch_direct_f32_list(const EvalParams &p, const float *qx,
                                  const float *qy, const float4 *f0,
                                  const float4 *f1, const WorkItem *items,
                                  int n_items, double *dst, double out_scale,
                                  int self_excl, hipStream_t s) {
  if (n_items <= 0) return hipSuccess;
  const float rho = (float)(p.gy / p.gx);
  const bool iso = p.iso != 0;
#define BK_LIST(F, I)                                                         \
  direct_f32_list_kernel<F, I><<<n_items, kBlock, 0, s>>>(                    \
      qx, qy, f0, f1, items, dst, out_scale, self_excl, rho)
  switch (p.family) {
    case FAM_GAUSS:
      if (iso) BK_LIST(FAM_GAUSS, true); else BK_LIST(FAM_GAUSS, false);
      break;
    case FAM_EPAN:
      if (iso) BK_LIST(FAM_EPAN, true); else BK_LIST(FAM_EPAN, false);
      break;
    case FAM_EPANPROD:
      BK_LIST(FAM_EPANPROD, false);
      break;
    default:
      BK_LIST(FAM_EPANCONV, false);
      break;
  }
#undef BK_LIST
  return hipGetLastError();
}

hipError_t launch_prep_f64(const EvalParams &p, const double *x,
                           const double *y, const double *m,
                           const double *abw, size_t n, double4 *f0,
                           double4 *f1, hipStream_t s) {
  if (n == 0) return hipSuccess;
  const int th = 256;
  prep_f64_kernel<<<(unsigned)((n + th - 1) / th), th, 0, s>>>(
      p, x, y, m, abw, (long long)n, f0, f1);
  return hipGetLastError();
}

hipError_t launch_direct_f64(const EvalParams &p, const double *qx,
                             const double *qy, const double4 *f0,
                             const double4 *f1, const WorkItem *items,
                             int n_items, long long r_i, long long r_end,
                             long long q_i, long long nq, int splits,
                             double *dst, double out_scale, int self_excl,
                             hipStream_t s) {
  const double rho = p.gy / p.gx;
  if (items) {
    if (n_items <= 0) return hipSuccess;
#define BK_L64(F)                                                             \
  direct_f64_list_kernel<F><<<n_items, kBlock64, 0, s>>>(                     \
      qx, qy, f0, f1, items, dst, out_scale, self_excl, rho)
    switch (p.family) {
      case FAM_GAUSS: BK_L64(FAM_GAUSS); break;
      case FAM_EPAN: BK_L64(FAM_EPAN); break;
      case FAM_EPANPROD: BK_L64(FAM_EPANPROD); break;
      default: BK_L64(FAM_EPANCONV); break;
    }
#undef BK_L64
  } else {
    if (nq <= 0) return hipSuccess;
    const long long len = split_length(r_end - r_i, splits, kTile64);
    const dim3 grid((unsigned)((nq + kBlock64 - 1) / kBlock64), (unsigned)splits);
#define BK_R64(F)                                                             \
  direct_f64_range_kernel<F><<<grid, kBlock64, 0, s>>>(                       \
      qx, qy, f0, f1, r_i, r_end, len, q_i, nq, dst, out_scale, self_excl, rho)
    switch (p.family) {
      case FAM_GAUSS: BK_R64(FAM_GAUSS); break;
      case FAM_EPAN: BK_R64(FAM_EPAN); break;
      case FAM_EPANPROD: BK_R64(FAM_EPANPROD); break;
      default: BK_R64(FAM_EPANCONV); break;
    }
#undef BK_R64
  }
  return hipGetLastError();
}

hipError_t launch_prep_general(const EvalParams &p, const float *abw,
                               const float *m, size_t n, double m_scale,
                               float2 *wm, hipStream_t s) {
  if (n == 0) return hipSuccess;
  const int th = 256;
  prep_general_kernel<<<(unsigned)((n + th - 1) / th), th, 0, s>>>(
      p, abw, m, (long long)n, 1.0 / m_scale, wm);
  return hipGetLastError();
}

namespace {
template <int D, int FAM>
void general_dispatch(const float *qc, const float *rc, const float2 *wm,
                      const WorkItem *items, int n_items, long long r_i,
                      long long r_end, long long q_i, long long nq, int splits,
                      double *dst, double out_scale, int self_excl,
                      hipStream_t s) {
  if (items) {
    direct_general_list_kernel<D, FAM><<<n_items, kBlockG, 0, s>>>(
        qc, rc, wm, items, dst, out_scale, self_excl);
  } else {
    const long long len = split_length(r_end - r_i, splits, kTileG);
    const dim3 grid((unsigned)((nq + kBlockG - 1) / kBlockG), (unsigned)splits);
    direct_general_range_kernel<D, FAM><<<grid, kBlockG, 0, s>>>(
        qc, rc, wm, r_i, r_end, len, q_i, nq, dst, out_scale, self_excl);
  }
}

template <int D>
void general_dispatch_fam(int fam, const float *qc, const float *rc,
                          const float2 *wm, const WorkItem *items, int n_items,
                          long long r_i, long long r_end, long long q_i,
                          long long nq, int splits, double *dst,
                          double out_scale, int self_excl, hipStream_t s) {
  if (fam == FAM_GAUSS) {
    general_dispatch<D, FAM_GAUSS>(qc, rc, wm, items, n_items, r_i, r_end, q_i,
                                   nq, splits, dst, out_scale, self_excl, s);
  } else {
    general_dispatch<D, FAM_EPAN>(qc, rc, wm, items, n_items, r_i, r_end, q_i,
                                  nq, splits, dst, out_scale, self_excl, s);
  }
}
}  // namespace

hipError_t launch_direct_general(const EvalParams &p, const float *qc,
                                 const float *rc, const float2 *wm,
                                 const WorkItem *items, int n_items,
                                 long long r_i, long long r_end, long long q_i,
                                 long long nq, int splits, double *dst,
                                 double out_scale, int self_excl,
                                 hipStream_t s) {
  if ((items && n_items <= 0) || (!items && nq <= 0)) return hipSuccess;
#define BK_GD(DV)                                                             \
  case DV:                                                                    \
    general_dispatch_fam<DV>(p.family, qc, rc, wm, items, n_items, r_i, r_end, \
                             q_i, nq, splits, dst, out_scale, self_excl, s);  \
    break
  switch (p.dim) {
    BK_GD(3); BK_GD(4); BK_GD(5); BK_GD(6); BK_GD(7); BK_GD(8); BK_GD(9);
    BK_GD(10); BK_GD(11); BK_GD(12); BK_GD(13); BK_GD(14); BK_GD(15); BK_GD(16);
    default:
      return hipErrorInvalidValue;
  }
#undef BK_GD
  return hipGetLastError();
}

hipError_t launch_reduce_splits(const double *partial, int splits,
                                long long nq, double *out, hipStream_t s) {
  if (nq <= 0) return hipSuccess;
  const int th = 256;
  reduce_splits_kernel<<<(unsigned)((nq + th - 1) / th), th, 0, s>>>(
      partial, splits, nq, out);
  return hipGetLastError();
}

hipError_t launch_weighted_reduce(int mode, const double *w, const double *u,
                                  const double *v, double a, double b,
                                  long long n, double *scratch,
                                  double *result, hipStream_t s) {
  weighted_reduce_kernel<<<kReduceBlocks, kReduceThreads, 0, s>>>(
      mode, w, u, v, a, b, n, scratch);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  final_reduce_kernel<<<1, kReduceThreads, 0, s>>>(scratch, kReduceBlocks,
                                                   result);
  return hipGetLastError();
}

hipError_t launch_adapt_lambda(const double *S, long long n, double norm,
                               const double *log_g_dev, double alpha,
                               double *lambda, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const int th = 256;
  adapt_lambda_kernel<<<(unsigned)((n + th - 1) / th), th, 0, s>>>(
      S, n, norm, log_g_dev, alpha, lambda);
  return hipGetLastError();
}

int device_cu_count() {
  static int cached[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  if (cached[dev] == 0) {
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) !=
            hipSuccess ||
        cus <= 0)
      cus = 256;
    cached[dev] = cus;
  }
  return cached[dev];
}

}  // namespace bk
