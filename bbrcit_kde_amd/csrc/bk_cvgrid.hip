// Batched cross-validation over datasets x bandwidth grids (bk_cv_grid_*,
// declared in include/bbrcit_hip.h).
//
// The reference's grid_search tool (grid_search/grid_search.cc:366-402)
// sets each bandwidth pair of a grid in turn and runs one cross-validation
// per pair, each a full self-evaluation of the dataset.  Here every dataset
// of a plan and every bandwidth pair of a cross-product grid hx[ix] x hy[iy]
// are evaluated by ONE persistent launch:
//   * work items are (dataset, block of 256 queries, reference split,
//     bandwidth sub-grid of A x B pairs), ordered longest first and pulled
//     from a device-side queue by resident workgroups until it is empty;
//   * the product kernels separate over the axes, U = ux(dx) uy(dy), so a
//     thread holding one query and A x B bandwidth pairs computes A x-factors
//     and B y-factors per reference and accumulates the A*B products
//     (rank-1 update): 16 exponentials serve 64 bandwidth pairs;
//   * references are staged 256 at a time in LDS (double-buffered, one
//     barrier per stage, every LDS read a broadcast); sums are FP32 within a
//     stage and FP64 across stages; the self pair is excluded by index;
//   * FP32 Gaussian leave-one-out sums below 2^-40 of the dataset's largest
//     mass are recomputed in FP64, then LCV and LSCV are reduced in FP64
//     with a fixed reduction tree, so scores are deterministic;
//   * the whole device pipeline (queue reset, pair kernel, split combine,
//     rescue, score reductions) is captured once per grid shape in a
//     hipGraph and replayed; bandwidth constants live in a device table
//     uploaded before each replay.
// The schedule (register tiling A x B, unroll of the reference loop) comes
// from the autotuner cache (bk_tune.hip, which can be saved to a file and
// reloaded) when it holds the run's (device, kind, grid shape, dataset
// sizes), else from a default; bk_cv_grid_tune fills the cache by timing
// every candidate on whole runs of the plan's own grid.  Every candidate sums
// in the same order, so the choice changes no result bit.
#include "../../include/bbrcit_hip.h"
#include "bk_common.h"
#include "bk_internal.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
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

constexpr int kCvBlock = 256;   // threads per workgroup = queries per work item
constexpr int kCvTile = 256;    // references per LDS stage
constexpr int kCvChunk = 64;    // references per wavefront-level test (one wavefront's stage writes)
constexpr int kMaxGridAxis = 64;
constexpr float kMaskedLog2 = -1.0e30f;  // base-2 log mass of a masked reference
constexpr double kLog2e = 1.4426950408889634074;
constexpr double kPi = 3.14159265358979323846;
constexpr double kRescueFraction = 9.094947017729282e-13;  // 2^-40
// base-2 exponent beyond which a Gaussian stage's products ex ey are all
// below 2^-152 (with the rounding of their FP32 exponents: below 2^-151)
constexpr double kCvGaussCullExp = 152.0;
constexpr int kScoreChunk = 4096;                            // points per score block

enum CvFamily : int { CV_GAUSS = 0, CV_EPAN = 1 };

// One work item: queries [q_begin, q_begin + q_count) against references
// [r_begin, r_end) of the same dataset, bandwidth sub-grid starting at
// (sub_x, sub_y), partial sums into slot `slot` (the reference split).
struct CvItem {
  int q_begin, q_count;
  int r_begin, r_end;
  int sub_x, sub_y;
  int slot, set;
};

struct CvArgs {
  const float4 *pts;    // Gaussian, per point: (x, y, log2(m/m_scale), 1/abw^2)
  const double4 *pts64; // Epanechnikov, per point: (x, y, m/m_scale, 1/abw)
  const CvItem *items;
  int n_items;
  int *queue;           // queue[0]: next item
  const float *kx;      // Gaussian, per hx index (padded): log2(e)/(2 hx^2)
  const float *ky;
  const double *kx64;   // Epanechnikov, per hx index (padded): 1/hx
  const double *ky64;
  double *P;            // partial sums [slot][plane][g][point], g = ix * n_hy + iy
  long long slot_stride;
  long long plane_stride;
  long long n_points;
  int n_hx, n_hy;
  unsigned long long *count;  // when set, the pair kernel adds the units it evaluates
};

// ---- Gaussian product kernel: 8 or 4 x-factors, 8 y-factors per reference ----
// U: unroll of the stage loop over references (a tuned parameter; it does
// not change the order of any sum).
// One chunk of kCvChunk references of the stage, starting at k0.
template <int A, int B, bool ADAPT, bool DIAG, int U>
__device__ __forceinline__ void gauss_stage(const float4 *__restrict__ tl, int k0, float qx, float qy, int self,
                                            const float (&kx)[A], const float (&ky)[B], float (&acc)[A][B]) {
#pragma unroll U
  for (int k = k0; k < k0 + kCvChunk; ++k) {
    const float4 r = tl[k];
    const float dx = qx - r.x, dy = qy - r.y;
    float tx = dx * dx, ty = dy * dy;
    if (ADAPT) {
      tx *= r.w;
      ty *= r.w;
    }
    float lm = r.z;
    if (DIAG) lm = (k == self) ? kMaskedLog2 : lm;
    float ex[A], ey[B];
#pragma unroll
    for (int a = 0; a < A; ++a) ex[a] = __builtin_amdgcn_exp2f(__builtin_fmaf(-tx, kx[a], lm));
#pragma unroll
    for (int b = 0; b < B; ++b) ey[b] = __builtin_amdgcn_exp2f(-ty * ky[b]);
#pragma unroll
    for (int a = 0; a < A; ++a)
#pragma unroll
      for (int b = 0; b < B; ++b) acc[a][b] = __builtin_fmaf(ex[a], ey[b], acc[a][b]);
  }
}

// ---- Epanechnikov product kernel and its convolution kernel (FP64) --------
// ux = max(1 - u^2, 0) cancels near the support edge, which FP32 cannot
// resolve to 1e-5 relative, so these sums are FP64 throughout (exact
// differences of the FP32 points, FP64 accumulators, no stage split; an FP64
// instruction costs gfx950 about 1.8 FP32 ones).  The convolution factor
// (2-u)^3 (u^2 + 6u + 4) for u < 2 carries 32 times the reference's k(u);
// the 1/1024 of both axes is applied in the score reduction.
template <int A, int B, bool ADAPT, bool DIAG, int U>
__device__ __forceinline__ void epan_stage(const double4 *__restrict__ tl, int k0, double qx, double qy, int self,
                                           const double (&kx)[A], const double (&ky)[B], double (&accK)[A][B],
                                           double (&accT)[A][B]) {
#pragma unroll U
  for (int k = k0; k < k0 + kCvChunk; ++k) {
    const double4 r = tl[k];
    double ax = fabs(qx - r.x), ay = fabs(qy - r.y);
    if (ADAPT) {
      ax *= r.w;
      ay *= r.w;
    }
    double m = r.z;
    if (DIAG) m = (k == self) ? 0.0 : m;
    double ex[A], px[A], ey[B], py[B];
#pragma unroll
    for (int a = 0; a < A; ++a) {
      const double u = ax * kx[a];
      const double e = bk::clamp_unit(fma(-u, u, 1.0));
      const double v = fmax(2.0 - u, 0.0);
      const double c = v * v * v * fma(u, u + 6.0, 4.0);
      ex[a] = m * e;
      px[a] = m * c;
    }
#pragma unroll
    for (int b = 0; b < B; ++b) {
      const double u = ay * ky[b];
      ey[b] = bk::clamp_unit(fma(-u, u, 1.0));
      const double v = fmax(2.0 - u, 0.0);
      py[b] = v * v * v * fma(u, u + 6.0, 4.0);
    }
#pragma unroll
    for (int a = 0; a < A; ++a)
#pragma unroll
      for (int b = 0; b < B; ++b) {
        accK[a][b] = fma(ex[a], ey[b], accK[a][b]);
        accT[a][b] = fma(px[a], py[b], accT[a][b]);
      }
  }
}

// Wave-wide minimum of a double (every lane gets it).
__device__ __forceinline__ double wave_min_d(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
  return v;
}

// Per-wavefront minima of the 5 values v (one set per thread) into
// box[5][wavefront]: written before a barrier, read after it.  Wavefront c of
// a stage's writes holds the box of the stage's c-th chunk of kCvChunk
// references; a wavefront's query box is its own entry.
__device__ __forceinline__ void box_write(double (&box)[5][kCvBlock / 64], const double (&v)[5]) {
#pragma unroll
  for (int i = 0; i < 5; ++i) {
    const double m = wave_min_d(v[i]);
    if ((threadIdx.x & 63) == 0) box[i][threadIdx.x >> 6] = m;
  }
}

// The (query, reference, bandwidth pair) units that wavefront wv of item w
// evaluates over the chunk of references starting at c0 (its lanes that
// hold a real query), added to g.count.
template <int A, int B>
__device__ __forceinline__ void count_units(const CvArgs &g, const CvItem &w, int c0, int wv) {
  const int la = g.n_hx - w.sub_x < A ? g.n_hx - w.sub_x : A;
  const int lb = g.n_hy - w.sub_y < B ? g.n_hy - w.sub_y : B;
  const int nr = w.r_end - c0 < kCvChunk ? w.r_end - c0 : kCvChunk;
  const int left = w.q_count - wv * 64;
  const int nq = left < 0 ? 0 : (left < 64 ? left : 64);
  if (nr > 0 && nq > 0)
    atomicAdd(g.count, (unsigned long long)nr * (unsigned long long)nq * (unsigned long long)(la * lb));
}

// Persistent pair kernels.  Every workgroup pulls items until the queue is
// exhausted; all branches on the item are uniform (the item index is read
// from LDS and made scalar).
//
// The Gaussian pair kernel: a wavefront passes over a chunk of kCvChunk
// references of a stage when every product ex ey of it is below 2^-152 for
// every query of the wavefront: the per-axis gaps between the wavefront's
// query box and the chunk's box give W_min (gx^2 kx_min + gy^2 ky_min) >=
// kCvGaussCullExp, with kx_min, ky_min the sub-grid's smallest exponent
// factors, W_min the chunk's smallest 1/abw^2 and log2(m / m_scale) <= 0.
// Such a product is below half of FP32's smallest denormal, so the FMA that
// would add it leaves its accumulator unchanged: passing over changes no
// result bit.  Small bandwidth sub-grids skip most chunks of Z-ordered data.
template <int A, int B, bool ADAPT, int U>
__global__ __launch_bounds__(kCvBlock) void cv_gauss_kernel(CvArgs g) {
  // the 8 x 8 tiling holds every bandwidth of a typical grid (nothing to
  // skip) and uses every register of 2 waves per SIMD: it tests no stage
  constexpr bool kCull = A * B < 64;
  __shared__ float4 tile[2][kCvTile];
  __shared__ double qbox[5][kCvBlock / 64];
  __shared__ double tbox[2][5][kCvBlock / 64];
  __shared__ int s_item;
  const int t = threadIdx.x;
  const double inf = __builtin_inf();
  // (x0, -x1, y0, -y1, smallest W) of one point; a masked slot is neutral
  auto point_box = [&](const float4 &r, bool valid, double (&v)[5]) {
    v[0] = valid ? (double)r.x : inf;
    v[1] = valid ? -(double)r.x : inf;
    v[2] = valid ? (double)r.y : inf;
    v[3] = valid ? -(double)r.y : inf;
    v[4] = valid ? (ADAPT ? (double)r.w : 1.0) : inf;
  };
  for (;;) {
    __syncthreads();
    if (t == 0) s_item = atomicAdd(g.queue, 1);
    __syncthreads();
    const int it = __builtin_amdgcn_readfirstlane(s_item);
    if (it >= g.n_items) break;
    const CvItem w = g.items[it];
    const bool active = t < w.q_count;
    const int qi = w.q_begin + (active ? t : 0);
    const float4 q = g.pts[qi];
    float kx[A], ky[B];
#pragma unroll
    for (int a = 0; a < A; ++a) kx[a] = g.kx[w.sub_x + a];
#pragma unroll
    for (int b = 0; b < B; ++b) ky[b] = g.ky[w.sub_y + b];
    // padded sub-grid entries repeat the last bandwidth, so the minimum over
    // all A (B) entries is the sub-grid's smallest factor
    float kx_min = kx[0], ky_min = ky[0];
#pragma unroll
    for (int a = 1; a < A; ++a) kx_min = __builtin_fminf(kx_min, kx[a]);
#pragma unroll
    for (int b = 1; b < B; ++b) ky_min = __builtin_fminf(ky_min, ky[b]);
    double tot[A][B];
#pragma unroll
    for (int a = 0; a < A; ++a)
#pragma unroll
      for (int b = 0; b < B; ++b) tot[a][b] = 0.0;

    if constexpr (kCull) {
      double v[5];
      point_box(q, true, v);  // inactive lanes hold the item's first query
      box_write(qbox, v);
    }
    const bool first_valid = w.r_begin + t < w.r_end;
    const float4 first = first_valid ? g.pts[w.r_begin + t] : make_float4(0.f, 0.f, kMaskedLog2, 0.f);
    tile[0][t] = first;
    if constexpr (kCull) {
      double v[5];
      point_box(first, first_valid, v);
      box_write(tbox[0], v);
    }
    __syncthreads();
    int buf = 0;
    const int q_end = w.q_begin + w.q_count;
    for (int r0 = w.r_begin; r0 < w.r_end; r0 += kCvTile) {
      const int rn = r0 + kCvTile;
      const bool nxt_valid = rn + t < w.r_end;
      float4 nxt = make_float4(0.f, 0.f, kMaskedLog2, 0.f);
      if (nxt_valid) nxt = g.pts[rn + t];
      // per chunk of kCvChunk references: this wavefront's query box against
      // the chunk's box (both re-read from LDS at every stage, no registers
      // held across it); a chunk out of reach of all its queries is passed over
      const bool diag = r0 < q_end && w.q_begin < rn;
      float acc[A][B];
#pragma unroll
      for (int a = 0; a < A; ++a)
#pragma unroll
        for (int b = 0; b < B; ++b) acc[a][b] = 0.0f;
      bool any = false;
      for (int c = 0; c < kCvTile / kCvChunk; ++c) {
        if (r0 + c * kCvChunk >= w.r_end) break;
        bool live = true;
        if constexpr (kCull) {
          const int wv = t >> 6;
          const double gx = fmax(0.0, fmax(tbox[buf][0][c] + qbox[1][wv], qbox[0][wv] + tbox[buf][1][c]));
          const double gy = fmax(0.0, fmax(tbox[buf][2][c] + qbox[3][wv], qbox[2][wv] + tbox[buf][3][c]));
          live = !(tbox[buf][4][c] * (gx * gx * (double)kx_min + gy * gy * (double)ky_min) >= kCvGaussCullExp);
        }
        if (!live) continue;
        any = true;
        if (g.count && (t & 63) == 0) count_units<A, B>(g, w, r0 + c * kCvChunk, t >> 6);
        if (diag)
          gauss_stage<A, B, ADAPT, true, U>(tile[buf], c * kCvChunk, q.x, q.y, qi - r0, kx, ky, acc);
        else
          gauss_stage<A, B, ADAPT, false, U>(tile[buf], c * kCvChunk, q.x, q.y, 0, kx, ky, acc);
      }
      if (any) {
#pragma unroll
        for (int a = 0; a < A; ++a)
#pragma unroll
          for (int b = 0; b < B; ++b) tot[a][b] += (double)acc[a][b];
      }
      if (rn < w.r_end) {
        tile[buf ^ 1][t] = nxt;
        if constexpr (kCull) {
          double v[5];
          point_box(nxt, nxt_valid, v);
          box_write(tbox[buf ^ 1], v);
        }
      }
      __syncthreads();
      buf ^= 1;
    }
    if (active) {
      double *dst = g.P + (long long)w.slot * g.slot_stride + qi;
#pragma unroll
      for (int a = 0; a < A; ++a)
#pragma unroll
        for (int b = 0; b < B; ++b)
          if (w.sub_x + a < g.n_hx && w.sub_y + b < g.n_hy)
            dst[(long long)((w.sub_x + a) * g.n_hy + w.sub_y + b) * g.n_points] = tot[a][b];
    }
  }
}

// The Epanechnikov pair kernel: a wavefront passes over a chunk of kCvChunk
// references whose references are all out of the support of every bandwidth
// of the item's sub-grid for every query of the wavefront: the per-axis gap
// between its query box and the chunk's box, times the chunk's smallest
// 1/abw and the sub-grid's smallest 1/h, is at least 2 (the convolution
// kernel's support, wider than the kernel's) on one axis.  The test repeats the pair term's own operations on smaller inputs,
// so every skipped term is exactly zero as computed and no result changes.
template <int A, int B, bool ADAPT, int U>
__global__ __launch_bounds__(kCvBlock) void cv_epan_kernel(CvArgs g) {
  __shared__ double4 tile[2][kCvTile];
  __shared__ double qbox[5][kCvBlock / 64];
  __shared__ double tbox[2][5][kCvBlock / 64];
  __shared__ int s_item;
  const int t = threadIdx.x;
  const double4 zero = make_double4(0.0, 0.0, 0.0, 0.0);
  const double inf = __builtin_inf();
  // (x0, -x1, y0, -y1, smallest 1/abw) of one point; a masked slot is neutral
  auto point_box = [&](const double4 &r, bool valid, double (&v)[5]) {
    v[0] = valid ? r.x : inf;
    v[1] = valid ? -r.x : inf;
    v[2] = valid ? r.y : inf;
    v[3] = valid ? -r.y : inf;
    v[4] = valid ? (ADAPT ? r.w : 1.0) : inf;
  };
  for (;;) {
    __syncthreads();
    if (t == 0) s_item = atomicAdd(g.queue, 1);
    __syncthreads();
    const int it = __builtin_amdgcn_readfirstlane(s_item);
    if (it >= g.n_items) break;
    const CvItem w = g.items[it];
    const bool active = t < w.q_count;
    const int qi = w.q_begin + (active ? t : 0);
    const double4 q = g.pts64[qi];
    double kx[A], ky[B];
#pragma unroll
    for (int a = 0; a < A; ++a) kx[a] = g.kx64[w.sub_x + a];
#pragma unroll
    for (int b = 0; b < B; ++b) ky[b] = g.ky64[w.sub_y + b];
    // padded sub-grid entries repeat the last bandwidth, so the minimum over
    // all A (B) entries is the sub-grid's smallest 1/h
    double kx_min = kx[0], ky_min = ky[0];
#pragma unroll
    for (int a = 1; a < A; ++a) kx_min = fmin(kx_min, kx[a]);
#pragma unroll
    for (int b = 1; b < B; ++b) ky_min = fmin(ky_min, ky[b]);
    double accK[A][B], accT[A][B];
#pragma unroll
    for (int a = 0; a < A; ++a)
#pragma unroll
      for (int b = 0; b < B; ++b) accK[a][b] = accT[a][b] = 0.0;

    {
      double v[5];
      point_box(q, true, v);  // inactive lanes hold the item's first query
      box_write(qbox, v);
    }
    const double4 first = w.r_begin + t < w.r_end ? g.pts64[w.r_begin + t] : zero;
    tile[0][t] = first;
    {
      double v[5];
      point_box(first, w.r_begin + t < w.r_end, v);
      box_write(tbox[0], v);
    }
    __syncthreads();
    int buf = 0;
    const int q_end = w.q_begin + w.q_count;
    const int wv = t >> 6;
    for (int r0 = w.r_begin; r0 < w.r_end; r0 += kCvTile) {
      const int rn = r0 + kCvTile;
      double4 nxt = zero;
      if (rn + t < w.r_end) nxt = g.pts64[rn + t];
      const bool diag = r0 < q_end && w.q_begin < rn;
      // per chunk of kCvChunk references: this wavefront's query box against
      // the chunk's box; a chunk out of the support for all its queries is
      // passed over
      for (int c = 0; c < kCvTile / kCvChunk; ++c) {
        if (r0 + c * kCvChunk >= w.r_end) break;
        const double gx = fmax(0.0, fmax(tbox[buf][0][c] + qbox[1][wv], qbox[0][wv] + tbox[buf][1][c]));
        const double gy = fmax(0.0, fmax(tbox[buf][2][c] + qbox[3][wv], qbox[2][wv] + tbox[buf][3][c]));
        const double ux = (ADAPT ? gx * tbox[buf][4][c] : gx) * kx_min;
        const double uy = (ADAPT ? gy * tbox[buf][4][c] : gy) * ky_min;
        if (ux >= 2.0 || uy >= 2.0) continue;
        if (g.count && (t & 63) == 0) count_units<A, B>(g, w, r0 + c * kCvChunk, wv);
        if (diag)
          epan_stage<A, B, ADAPT, true, U>(tile[buf], c * kCvChunk, q.x, q.y, qi - r0, kx, ky, accK, accT);
        else
          epan_stage<A, B, ADAPT, false, U>(tile[buf], c * kCvChunk, q.x, q.y, 0, kx, ky, accK, accT);
      }
      if (rn < w.r_end) {
        tile[buf ^ 1][t] = nxt;
        double v[5];
        point_box(nxt, rn + t < w.r_end, v);
        box_write(tbox[buf ^ 1], v);
      }
      __syncthreads();
      buf ^= 1;
    }
    if (active) {
      double *dst = g.P + (long long)w.slot * g.slot_stride + qi;
#pragma unroll
      for (int a = 0; a < A; ++a)
#pragma unroll
        for (int b = 0; b < B; ++b)
          if (w.sub_x + a < g.n_hx && w.sub_y + b < g.n_hy) {
            const long long off = (long long)((w.sub_x + a) * g.n_hy + w.sub_y + b) * g.n_points;
            dst[off] = accK[a][b];
            dst[g.plane_stride + off] = accT[a][b];
          }
    }
  }
}

__global__ void cv_reset_kernel(int *queue) {
  if (threadIdx.x < 4) queue[threadIdx.x] = 0;
}

// Sum the reference splits of one dataset into slot 0 (fixed order).
// (P is [slot][plane][g][point], so plane p, bandwidth g sit at (p G + g) n_points.)
__global__ void cv_combine_kernel(double *__restrict__ P, long long slot_stride, long long n_points, int planes,
                                  int G, int splits, int begin, int n) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long total = (long long)planes * G * n;
  if (k >= total) return;
  const long long i = k % n;
  const long long pg = k / n;  // plane * G + g
  const long long off = pg * n_points + begin + i;
  double s = P[off];
  for (int sp = 1; sp < splits; ++sp) s += P[sp * slot_stride + off];
  P[off] = s;
}

// Flag (point, sub-grid) entries of one dataset whose Gaussian sum fell
// below the rescue threshold; queue[1] counts them, list holds (point, sub).
__global__ void cv_flag_kernel(const double *__restrict__ P, long long n_points, int n_hx, int n_hy, int A,
                               int B, int nsx, int nsy, int begin, int n, double threshold, int *queue,
                               int2 *__restrict__ list, int capacity) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (long long)n * nsx * nsy) return;
  const int i = (int)(k % n);
  const int sub = (int)(k / n);
  const int sx = (sub / nsy) * A, sy = (sub % nsy) * B;
  bool small = false;
  for (int a = 0; a < A && sx + a < n_hx; ++a)
    for (int b = 0; b < B && sy + b < n_hy; ++b)
      small |= P[(long long)((sx + a) * n_hy + sy + b) * n_points + begin + i] < threshold;
  if (small) {
    const int slot = atomicAdd(queue + 1, 1);
    if (slot < capacity) list[slot] = make_int2(begin + i, sub);
  }
}

// 2^e for e <= 0 in FP64 range: FP32 exp2 of the fractional part, exact
// power of two by ldexp in FP64 (relative error ~1e-7; values down to
// 2^-1074 survive).
__device__ __forceinline__ double exp2_wide(double e) {
  const double n = fmax(floor(e), -2000.0);
  return ldexp((double)__builtin_amdgcn_exp2f((float)(e - n)), (int)n);
}

__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// FP64 recomputation of flagged Gaussian (point, sub-grid) entries, one
// workgroup per entry.  pd holds (m / m_scale, 1/abw^2) per point in FP64
// from the FP32-rounded values; the products keep FP64 range.
template <int A, int B>
__global__ __launch_bounds__(kCvBlock) void cv_rescue_kernel(const float4 *__restrict__ pts,
                                                             const double2 *__restrict__ pd,
                                                             const int *__restrict__ set_of_point,
                                                             const int *__restrict__ set_begin,
                                                             const double *__restrict__ kx64,
                                                             const double *__restrict__ ky64, int nsy, int n_hx,
                                                             int n_hy, const int *queue, const int2 *list,
                                                             int capacity, double threshold,
                                                             double *__restrict__ P, long long n_points) {
  __shared__ double red[kCvBlock / 64][A * B];
  const int count = min(queue[1], capacity);
  for (int e = blockIdx.x; e < count; e += gridDim.x) {
    const int2 en = list[e];
    const int qi = en.x;
    const int sx = (en.y / nsy) * A, sy = (en.y % nsy) * B;
    const int set = set_of_point[qi];
    const int rb = set_begin[set], re = set_begin[set + 1];
    const double qx = (double)pts[qi].x, qy = (double)pts[qi].y;
    double kx[A], ky[B];
    for (int a = 0; a < A; ++a) kx[a] = kx64[sx + a];
    for (int b = 0; b < B; ++b) ky[b] = ky64[sy + b];
    double s[A][B];
    for (int a = 0; a < A; ++a)
      for (int b = 0; b < B; ++b) s[a][b] = 0.0;
    for (int j = rb + (int)threadIdx.x; j < re; j += kCvBlock) {
      if (j == qi) continue;
      const double dx = qx - (double)pts[j].x, dy = qy - (double)pts[j].y;
      const double2 md = pd[j];
      const double tx = dx * dx * md.y, ty = dy * dy * md.y;
      double fy[B];
      for (int b = 0; b < B; ++b) fy[b] = exp2_wide(-ty * ky[b]);
      for (int a = 0; a < A; ++a) {
        const double fx = md.x * exp2_wide(-tx * kx[a]);
        for (int b = 0; b < B; ++b) s[a][b] += fx * fy[b];
      }
    }
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    for (int a = 0; a < A; ++a)
      for (int b = 0; b < B; ++b) {
        const double v = wave_sum(s[a][b]);
        if (lane == 0) red[wave][a * B + b] = v;
      }
    __syncthreads();
    if ((int)threadIdx.x < A * B) {
      const int a = threadIdx.x / B, b = threadIdx.x % B;
      double v = 0.0;
      for (int wv = 0; wv < kCvBlock / 64; ++wv) v += red[wv][threadIdx.x];
      if (sx + a < n_hx && sy + b < n_hy) {
        double *e = P + (long long)((sx + a) * n_hy + sy + b) * n_points + qi;
        if (*e < threshold) *e = v;  // the same entries for every tiling
      }
    }
    __syncthreads();
  }
}

__device__ double block_sum(double v, double *sh) {
  v = wave_sum(v);
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  if (lane == 0) sh[wave] = v;
  __syncthreads();
  double r = 0.0;
  for (int k = 0; k < kCvBlock / 64; ++k) r += sh[k];
  __syncthreads();
  return r;
}

// Per (chunk, bandwidth) partial scores of one dataset:
//   lcv  += w_i log(C m_scale S_i)
//   lscv += w_i (Cc (m_scale tf T_i + m_i) - 2 C m_scale S_i)
// where T = S for the Gaussian product kernel (its convolution kernel has
// the same profile) and the second plane for the Epanechnikov product.
__global__ __launch_bounds__(kCvBlock) void cv_score_kernel(const double *__restrict__ P, long long plane_stride,
                                                            long long n_points, int planes, int begin, int n,
                                                            const double *__restrict__ w,
                                                            const double2 *__restrict__ pd, double m_scale,
                                                            double tf, const double *__restrict__ normC,
                                                            const double *__restrict__ convC, int nchunks,
                                                            double *__restrict__ part) {
  __shared__ double sh[kCvBlock / 64];
  const int chunk = blockIdx.x;
  const int gidx = blockIdx.y;
  const double C = normC[gidx], Cc = convC[gidx];
  const double *S = P + (long long)gidx * n_points + begin;
  const double *T = planes > 1 ? S + plane_stride : S;
  double l = 0.0, s = 0.0;
  const int i0 = chunk * kScoreChunk;
  const int i1 = min(n, i0 + kScoreChunk);
  for (int i = i0 + (int)threadIdx.x; i < i1; i += kCvBlock) {
    const double wi = w[begin + i];
    const double Si = m_scale * S[i];
    const double Ti = m_scale * tf * T[i];
    const double mi = m_scale * pd[begin + i].x;
    l += wi * log(C * Si);
    s += wi * (Cc * (Ti + mi) - 2.0 * C * Si);
  }
  l = block_sum(l, sh);
  s = block_sum(s, sh);
  if (threadIdx.x == 0) {
    part[((long long)gidx * nchunks + chunk) * 2] = l;
    part[((long long)gidx * nchunks + chunk) * 2 + 1] = s;
  }
}

__global__ __launch_bounds__(kCvBlock) void cv_score_final_kernel(const double *__restrict__ part, int nchunks,
                                                                  int G, int set, double *__restrict__ out) {
  __shared__ double sh[kCvBlock / 64];
  const int gidx = blockIdx.x;
  double l = 0.0, s = 0.0;
  for (int c = threadIdx.x; c < nchunks; c += kCvBlock) {
    l += part[((long long)gidx * nchunks + c) * 2];
    s += part[((long long)gidx * nchunks + c) * 2 + 1];
  }
  l = block_sum(l, sh);
  s = block_sum(s, sh);
  if (threadIdx.x == 0) {
    out[((long long)set * G + gidx) * 2] = l;
    out[((long long)set * G + gidx) * 2 + 1] = s;
  }
}

// ---- variants -----------------------------------------------------------------
struct Variant {
  int A, B;
};
constexpr Variant kGaussVariants[BK_CVG_VARIANTS] = {{8, 8}, {4, 8}, {4, 4}};
constexpr Variant kEpanVariants[BK_CVG_VARIANTS] = {{4, 8}, {4, 4}, {2, 4}};

// Unroll candidates of the stage loops (index u of a schedule).
constexpr int kGaussUnrolls[3] = {2, 4, 8};
constexpr int kEpanUnrolls[2] = {1, 2};
constexpr int kGaussUnrollDefault = 0;  // 2
constexpr int kEpanUnrollDefault = 1;   // 2

int unroll_count(int fam) { return fam == CV_GAUSS ? 3 : 2; }
static_assert(BK_CVG_VARIANTS * 3 <= BK_CVG_TUNE_MAX, "bk_cv_grid_tune writes at most BK_CVG_TUNE_MAX timings");

using PairKernel = void (*)(CvArgs);

template <int A, int B>
PairKernel gauss_kernel(bool adapt, int u) {
  switch (u) {
    case 0: return adapt ? cv_gauss_kernel<A, B, true, 2> : cv_gauss_kernel<A, B, false, 2>;
    case 1: return adapt ? cv_gauss_kernel<A, B, true, 4> : cv_gauss_kernel<A, B, false, 4>;
    default: return adapt ? cv_gauss_kernel<A, B, true, 8> : cv_gauss_kernel<A, B, false, 8>;
  }
}

template <int A, int B>
PairKernel epan_kernel(bool adapt, int u) {
  if (u == 0) return adapt ? cv_epan_kernel<A, B, true, 1> : cv_epan_kernel<A, B, false, 1>;
  return adapt ? cv_epan_kernel<A, B, true, 2> : cv_epan_kernel<A, B, false, 2>;
}

PairKernel pair_kernel(int fam, int variant, bool adapt, int u) {
  if (fam == CV_GAUSS) {
    if (variant == 0) return gauss_kernel<8, 8>(adapt, u);
    return variant == 1 ? gauss_kernel<4, 8>(adapt, u) : gauss_kernel<4, 4>(adapt, u);
  }
  if (variant == 0) return epan_kernel<4, 8>(adapt, u);
  return variant == 1 ? epan_kernel<4, 4>(adapt, u) : epan_kernel<2, 4>(adapt, u);
}

Variant variant_shape(int fam, int v) { return fam == CV_GAUSS ? kGaussVariants[v] : kEpanVariants[v]; }

// A schedule of the pair kernel: register tiling and unroll index.  Every
// schedule sums in the same order.
struct Schedule {
  int variant = -1, unroll = 0;
  bool operator==(const Schedule &o) const { return variant == o.variant && unroll == o.unroll; }
};

}  // namespace

struct bk_cv_grid {
  bk_ctx *ctx = nullptr;
  int device = 0;
  int kind = 0;
  int fam = CV_GAUSS;
  hipStream_t stream = nullptr;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};

  // datasets
  int n_sets = 0;
  std::vector<int> begin;        // n_sets + 1 offsets
  std::vector<double> m_scale;   // per set
  std::vector<int> splits;       // per set: reference splits
  std::vector<int> split_len;    // per set: references per split (multiple of the LDS tile)
  int max_splits = 1;
  long long n_points = 0;
  bool adapt = false;
  DevBuf pts, pts64, pd, w, set_of_point, set_begin;
  std::vector<double> m_host;    // FP32-rounded masses (as double), device order
  std::vector<int> perm;         // device position -> the caller's point index

  // per grid shape
  int n_hx = 0, n_hy = 0;
  int variant = -1;          // option: -1 autotune
  Schedule used;             // schedule of the last run
  Schedule forced;           // set by bk_cv_grid_tune while it times a candidate
  int tune_source = 0;       // BK_CVG_TUNE_SOURCE of the last run
  bool use_graph = true;
  int shape_hx = -1, shape_hy = -1;
  Schedule shape_sched;      // schedule the items, buffers and graph were made for
  int items_n = 0, nsx = 0, nsy = 0, grid_blocks = 0, rescue_cap = 0;
  DevBuf items, queue, consts, P, list, part, scores;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool have_run = false;
  double last_main_ms = std::nan(""), last_total_ms = 0.0;
  long long last_rescued = 0;
  bool count_units = false;           // option BK_CVG_COUNT_UNITS
  DevBuf counter;                     // units evaluated by the pair kernel
  long long last_units = -1;          // of the last run; -1: not counted
  int pad_hx = 0, pad_hy = 0;

  // device constant table layout (doubles): kx64[pad_hx] ky64[pad_hy]
  // normC[G] convC[G]; floats after: kx32[pad_hx] ky32[pad_hy]
  size_t off_kx64 = 0, off_ky64 = 0, off_norm = 0, off_conv = 0, off_kx32 = 0, off_ky32 = 0;
  std::vector<double> hx_last, hy_last;
};

namespace {

// perm[begin..end) = begin + the points of [begin, end) in Z-curve order:
// Morton code of (x, y) on a 1024 x 1024 grid over the set's box, ties by
// index.
void spatial_order(const double *coords, int begin, int end, std::vector<int> &perm) {
  double x0 = coords[2 * begin], x1 = x0, y0 = coords[2 * begin + 1], y1 = y0;
  for (int i = begin; i < end; ++i) {
    x0 = std::min(x0, coords[2 * i]);
    x1 = std::max(x1, coords[2 * i]);
    y0 = std::min(y0, coords[2 * i + 1]);
    y1 = std::max(y1, coords[2 * i + 1]);
  }
  const double sx = x1 > x0 ? 1023.0 / (x1 - x0) : 0.0, sy = y1 > y0 ? 1023.0 / (y1 - y0) : 0.0;
  auto spread = [](uint64_t v) {
    v &= 0x3FFu;
    v = (v | (v << 8)) & 0x00FF00FFu;
    v = (v | (v << 4)) & 0x0F0F0F0Fu;
    v = (v | (v << 2)) & 0x33333333u;
    v = (v | (v << 1)) & 0x55555555u;
    return v;
  };
  std::vector<uint64_t> key(end - begin);
  for (int i = begin; i < end; ++i) {
    const uint64_t cx = (uint64_t)((coords[2 * i] - x0) * sx), cy = (uint64_t)((coords[2 * i + 1] - y0) * sy);
    key[i - begin] = ((spread(cx) | (spread(cy) << 1)) << 32) | (uint64_t)(i - begin);
  }
  std::sort(key.begin(), key.end());
  for (int k = 0; k < end - begin; ++k) perm[begin + k] = begin + (int)(key[k] & 0xFFFFFFFFu);
}

void drop_graph(bk_cv_grid *g) {
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  g->exec = nullptr;
  g->graph = nullptr;
}

int cu_count() { return bk::device_cu_count(); }

// Reference splits of each dataset: enough items that the longest is at
// most 1/16 of one resident workgroup's fair share.  Depends only on the
// dataset sizes and the CU count (not on the tiling variant), so every
// variant sums in the same order.
void choose_splits(bk_cv_grid *g) {
  const double slots = 2.0 * cu_count();
  double total = 0.0;
  for (int d = 0; d < g->n_sets; ++d) {
    const double n = g->begin[d + 1] - g->begin[d];
    total += n * n;
  }
  const double fair = total / slots;
  g->splits.assign(g->n_sets, 1);
  g->split_len.assign(g->n_sets, 1);
  g->max_splits = 1;
  for (int d = 0; d < g->n_sets; ++d) {
    const long long n = g->begin[d + 1] - g->begin[d];
    const double item = (double)std::min<long long>(n, kCvBlock) * (double)n;
    long long s = (long long)std::ceil(item / (fair / 16.0));
    const long long cap = std::max<long long>(1, n / (4LL * kCvTile));
    s = std::max<long long>(1, std::min<long long>(s, std::min<long long>(cap, 64)));
    // tile-aligned split length; the count follows from it, so no split is empty
    long long len = (n + s - 1) / s;
    len = (len + kCvTile - 1) / kCvTile * kCvTile;
    s = (n + len - 1) / len;
    g->split_len[d] = (int)len;
    g->splits[d] = (int)s;
    g->max_splits = std::max(g->max_splits, (int)s);
  }
}

// Work items of one tiling (sub-grids of A x B), longest first.
std::vector<CvItem> make_items(const bk_cv_grid *g, int A, int B, int nsx, int nsy) {
  std::vector<CvItem> items;
  for (int d = 0; d < g->n_sets; ++d) {
    const int b = g->begin[d], e = g->begin[d + 1];
    const int s = g->splits[d];
    const long long len = g->split_len[d];
    for (int q0 = b; q0 < e; q0 += kCvBlock) {
      const int qc = std::min(kCvBlock, e - q0);
      for (int sp = 0; sp < s; ++sp) {
        const long long r0 = b + sp * len;
        const long long r1 = std::min<long long>(e, r0 + len);
        for (int sx = 0; sx < nsx; ++sx)
          for (int sy = 0; sy < nsy; ++sy) {
            CvItem it{};
            it.q_begin = q0;
            it.q_count = qc;
            it.r_begin = (int)r0;
            it.r_end = (int)r1;
            it.sub_x = sx * A;
            it.sub_y = sy * B;
            it.slot = sp;
            it.set = d;
            items.push_back(it);
          }
      }
    }
  }
  std::stable_sort(items.begin(), items.end(), [](const CvItem &l, const CvItem &r) {
    return (long long)l.q_count * (l.r_end - l.r_begin) > (long long)r.q_count * (r.r_end - r.r_begin);
  });
  return items;
}

// Every split slot of a dataset must be written before the combine: with the
// split rule above each split range is non-empty, which make_items checks.
void check_split_cover(const bk_cv_grid *g) {
  for (int d = 0; d < g->n_sets; ++d) {
    const long long n = g->begin[d + 1] - g->begin[d];
    const int s = g->splits[d];
    const long long len = g->split_len[d];
    if ((long long)(s - 1) * len >= n || (long long)s * len < n) fail(BK_ERR_STATE, "internal: bad reference split");
  }
}

int resident_blocks(PairKernel k) {
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void *>(k), kCvBlock, 0) !=
          hipSuccess ||
      per_cu <= 0)
    per_cu = 1;
  return per_cu * cu_count();
}

CvArgs make_args(const bk_cv_grid *g, const CvItem *items, int n_items, bool counting = false) {
  CvArgs a{};
  a.count = counting ? g->counter.as<unsigned long long>() : nullptr;
  a.pts = g->pts.as<float4>();
  a.pts64 = g->pts64.as<double4>();
  a.items = items;
  a.n_items = n_items;
  a.queue = g->queue.as<int>();
  const char *base = g->consts.as<char>();
  a.kx = reinterpret_cast<const float *>(base + g->off_kx32);
  a.ky = reinterpret_cast<const float *>(base + g->off_ky32);
  a.kx64 = reinterpret_cast<const double *>(base + g->off_kx64);
  a.ky64 = reinterpret_cast<const double *>(base + g->off_ky64);
  a.P = g->P.as<double>();
  const long long G = (long long)g->n_hx * g->n_hy;
  const int planes = g->fam == CV_GAUSS ? 1 : 2;
  a.plane_stride = G * g->n_points;
  a.slot_stride = planes * a.plane_stride;
  a.n_points = g->n_points;
  a.n_hx = g->n_hx;
  a.n_hy = g->n_hy;
  return a;
}

// Size device buffers and the constant table for the current grid shape.
void size_buffers(bk_cv_grid *g) {
  const long long G = (long long)g->n_hx * g->n_hy;
  const int planes = g->fam == CV_GAUSS ? 1 : 2;
  const size_t pbytes = (size_t)g->max_splits * planes * G * g->n_points * sizeof(double);
  g->P.reserve(pbytes);
  g->queue.reserve(4 * sizeof(int));
  g->pad_hx = g->n_hx + 8;
  g->pad_hy = g->n_hy + 8;
  g->off_kx64 = 0;
  g->off_ky64 = g->off_kx64 + g->pad_hx * sizeof(double);
  g->off_norm = g->off_ky64 + g->pad_hy * sizeof(double);
  g->off_conv = g->off_norm + G * sizeof(double);
  g->off_kx32 = g->off_conv + G * sizeof(double);
  g->off_ky32 = g->off_kx32 + g->pad_hx * sizeof(float);
  g->consts.reserve(g->off_ky32 + g->pad_hy * sizeof(float));
  long long max_chunks = 1;
  for (int d = 0; d < g->n_sets; ++d)
    max_chunks = std::max<long long>(max_chunks, (g->begin[d + 1] - g->begin[d] + kScoreChunk - 1) / kScoreChunk);
  g->part.reserve((size_t)G * max_chunks * 2 * sizeof(double));
  g->scores.reserve((size_t)g->n_sets * G * 2 * sizeof(double));
}

// Upload the bandwidth constants of one run (outside any graph).
void upload_consts(bk_cv_grid *g, const double *hx, const double *hy) {
  const long long G = (long long)g->n_hx * g->n_hy;
  std::vector<char> buf(g->consts.bytes, 0);
  double *kx64 = reinterpret_cast<double *>(buf.data() + g->off_kx64);
  double *ky64 = reinterpret_cast<double *>(buf.data() + g->off_ky64);
  double *normC = reinterpret_cast<double *>(buf.data() + g->off_norm);
  double *convC = reinterpret_cast<double *>(buf.data() + g->off_conv);
  float *kx32 = reinterpret_cast<float *>(buf.data() + g->off_kx32);
  float *ky32 = reinterpret_cast<float *>(buf.data() + g->off_ky32);
  auto kconst = [&](double h) { return g->fam == CV_GAUSS ? 0.5 * kLog2e / (h * h) : 1.0 / h; };
  for (int i = 0; i < g->pad_hx; ++i) {
    const double v = kconst(hx[std::min(i, g->n_hx - 1)]);
    kx64[i] = v;
    kx32[i] = (float)v;
  }
  for (int i = 0; i < g->pad_hy; ++i) {
    const double v = kconst(hy[std::min(i, g->n_hy - 1)]);
    ky64[i] = v;
    ky32[i] = (float)v;
  }
  for (int ix = 0; ix < g->n_hx; ++ix)
    for (int iy = 0; iy < g->n_hy; ++iy) {
      const double a = hx[ix], b = hy[iy];
      // GaussianProductKernel2d: 1/(2 pi hx hy); its convolution kernel
      // (reference: exp(-0.5 arg), KernelTraits.h:432-445) 1/(4 pi hx hy).
      // EpanechnikovProductKernel2d: (9/16)/(hx hy); convolution 0.36/(hx hy).
      normC[ix * g->n_hy + iy] = g->fam == CV_GAUSS ? 1.0 / (2.0 * kPi * a * b) : 0.5625 / (a * b);
      convC[ix * g->n_hy + iy] = g->fam == CV_GAUSS ? 1.0 / (4.0 * kPi * a * b) : 0.36 / (a * b);
    }
  (void)G;
  hip_check(hipMemcpyAsync(g->consts.ptr, buf.data(), buf.size(), hipMemcpyHostToDevice, g->stream),
            "upload bandwidth constants");
  hip_check(hipStreamSynchronize(g->stream), "hipStreamSynchronize");
}

// Enqueue the device pipeline after the pair kernel's queue reset.
void enqueue_pipeline(bk_cv_grid *g, bool time_main) {
  const hipStream_t s = g->stream;
  const long long G = (long long)g->n_hx * g->n_hy;
  const int planes = g->fam == CV_GAUSS ? 1 : 2;
  const Variant v = variant_shape(g->fam, g->used.variant);
  const CvArgs args = make_args(g, g->items.as<CvItem>(), g->items_n, g->count_units);
  hipLaunchKernelGGL(cv_reset_kernel, dim3(1), dim3(64), 0, s, g->queue.as<int>());
  hip_check(hipGetLastError(), "queue reset launch");
  if (time_main) hip_check(hipEventRecord(g->ev[2], s), "hipEventRecord");
  const PairKernel k = pair_kernel(g->fam, g->used.variant, g->adapt, g->used.unroll);
  hipLaunchKernelGGL(k, dim3(g->grid_blocks), dim3(kCvBlock), 0, s, args);
  hip_check(hipGetLastError(), "pair kernel launch");
  if (time_main) hip_check(hipEventRecord(g->ev[3], s), "hipEventRecord");
  for (int d = 0; d < g->n_sets; ++d) {
    if (g->splits[d] <= 1) continue;
    const int n = g->begin[d + 1] - g->begin[d];
    const long long total = (long long)planes * G * n;
    hipLaunchKernelGGL(cv_combine_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, g->P.as<double>(),
                       args.slot_stride, g->n_points, planes, (int)G, g->splits[d], g->begin[d], n);
    hip_check(hipGetLastError(), "combine launch");
  }
  if (g->fam == CV_GAUSS) {
    for (int d = 0; d < g->n_sets; ++d) {
      const int n = g->begin[d + 1] - g->begin[d];
      const long long total = (long long)n * g->nsx * g->nsy;
      hipLaunchKernelGGL(cv_flag_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, g->P.as<double>(),
                         g->n_points, g->n_hx, g->n_hy, v.A, v.B, g->nsx, g->nsy, g->begin[d], n,
                         kRescueFraction, g->queue.as<int>(), g->list.as<int2>(), g->rescue_cap);
      hip_check(hipGetLastError(), "flag launch");
    }
    const char *base = g->consts.as<char>();
    const double *kx64 = reinterpret_cast<const double *>(base + g->off_kx64);
    const double *ky64 = reinterpret_cast<const double *>(base + g->off_ky64);
    const unsigned rblocks = (unsigned)(4 * cu_count());
#define BK_RESCUE(AA, BB)                                                                                        \
  hipLaunchKernelGGL((cv_rescue_kernel<AA, BB>), dim3(rblocks), dim3(kCvBlock), 0, s, g->pts.as<float4>(),        \
                     g->pd.as<double2>(), g->set_of_point.as<int>(), g->set_begin.as<int>(), kx64, ky64, g->nsy, \
                     g->n_hx, g->n_hy, g->queue.as<int>(), g->list.as<int2>(), g->rescue_cap, kRescueFraction,   \
                     g->P.as<double>(), g->n_points)
    if (v.A == 8)
      BK_RESCUE(8, 8);
    else if (v.B == 8)
      BK_RESCUE(4, 8);
    else
      BK_RESCUE(4, 4);
#undef BK_RESCUE
    hip_check(hipGetLastError(), "rescue launch");
  }
  const char *base = g->consts.as<char>();
  const double *normC = reinterpret_cast<const double *>(base + g->off_norm);
  const double *convC = reinterpret_cast<const double *>(base + g->off_conv);
  const double tf = g->fam == CV_GAUSS ? 1.0 : 1.0 / 1024.0;
  for (int d = 0; d < g->n_sets; ++d) {
    const int n = g->begin[d + 1] - g->begin[d];
    const int nchunks = (n + kScoreChunk - 1) / kScoreChunk;
    hipLaunchKernelGGL(cv_score_kernel, dim3(nchunks, (unsigned)G), dim3(kCvBlock), 0, s, g->P.as<double>(),
                       args.plane_stride, g->n_points, planes, g->begin[d], n, g->w.as<double>(),
                       g->pd.as<double2>(), g->m_scale[d], tf, normC, convC, nchunks, g->part.as<double>());
    hip_check(hipGetLastError(), "score launch");
    hipLaunchKernelGGL(cv_score_final_kernel, dim3((unsigned)G), dim3(kCvBlock), 0, s, g->part.as<double>(),
                       nchunks, (int)G, d, g->scores.as<double>());
    hip_check(hipGetLastError(), "score final launch");
  }
}

// Workgroups a schedule launches: the kernel's occupancy limit on every CU.
// (Half of it was a candidate once: slower on every full run at the 4-set
// 8 x 8 workload, by 7% to 36%, so it is not searched.)
int schedule_blocks(const bk_cv_grid *g, const Schedule &sc) {
  return resident_blocks(pair_kernel(g->fam, sc.variant, g->adapt, sc.unroll));
}

// Autotuner key of a plan's current run: device, kind, bandwidth model, grid
// shape and dataset sizes.
std::string tune_key(const bk_cv_grid *g) {
  std::string k = "cvgrid/" + bk::device_tune_tag() + "/kind" + std::to_string(g->kind) +
                  (g->adapt ? "/adaptive" : "/fixed") + "/grid" + std::to_string(g->n_hx) + "x" +
                  std::to_string(g->n_hy) + "/sets";
  for (int d = 0; d < g->n_sets; ++d) k += (d ? "," : "") + std::to_string(g->begin[d + 1] - g->begin[d]);
  return k;
}

bool valid_schedule(const bk_cv_grid *g, const std::vector<int> &v) {
  return v.size() == 2 && v[0] >= 0 && v[0] < BK_CVG_VARIANTS && v[1] >= 0 && v[1] < unroll_count(g->fam);
}

// Default schedule of a kind (the fastest at the 4-set 8 x 8 workload).
Schedule default_schedule(int fam) {
  Schedule sc;
  sc.variant = fam == CV_GAUSS ? 2 : 1;
  sc.unroll = fam == CV_GAUSS ? kGaussUnrollDefault : kEpanUnrollDefault;
  return sc;
}

// The schedule of this run: forced (BK_CVG_VARIANT with the default unroll,
// or a candidate under bk_cv_grid_tune), else the autotuner cache's entry for
// the run, else the default.
Schedule choose_schedule(bk_cv_grid *g) {
  if (g->forced.variant >= 0) {
    g->tune_source = 0;
    return g->forced;
  }
  if (g->variant >= 0) {
    g->tune_source = 0;
    Schedule sc = default_schedule(g->fam);
    sc.variant = g->variant;
    return sc;
  }
  std::vector<int> v;
  if (bk::tune_lookup(tune_key(g), v) && valid_schedule(g, v)) {
    g->tune_source = 2;
    Schedule sc;
    sc.variant = v[0];
    sc.unroll = v[1];
    return sc;
  }
  g->tune_source = 1;
  return default_schedule(g->fam);
}

// Prepare items, buffers and (optionally) the graph for the current shape.
void prepare_shape(bk_cv_grid *g) {
  const Schedule sc = choose_schedule(g);
  if (g->shape_hx == g->n_hx && g->shape_hy == g->n_hy && g->shape_sched == sc && g->items_n > 0) {
    g->used = sc;
    return;
  }
  drop_graph(g);
  g->used = sc;
  const Variant sh = variant_shape(g->fam, sc.variant);
  g->nsx = (g->n_hx + sh.A - 1) / sh.A;
  g->nsy = (g->n_hy + sh.B - 1) / sh.B;
  std::vector<CvItem> items = make_items(g, sh.A, sh.B, g->nsx, g->nsy);
  if (items.size() > (size_t)std::numeric_limits<int>::max()) fail(BK_ERR_INVALID_ARG, "too many work items");
  g->items.reserve(items.size() * sizeof(CvItem));
  hip_check(hipMemcpyAsync(g->items.ptr, items.data(), items.size() * sizeof(CvItem), hipMemcpyHostToDevice,
                           g->stream),
            "upload work items");
  hip_check(hipStreamSynchronize(g->stream), "hipStreamSynchronize");
  g->items_n = (int)items.size();
  g->grid_blocks = schedule_blocks(g, sc);
  const long long cap = (long long)g->n_points * g->nsx * g->nsy;
  if (cap > std::numeric_limits<int>::max()) fail(BK_ERR_INVALID_ARG, "grid too large for the rescue list");
  g->rescue_cap = g->fam == CV_GAUSS ? (int)cap : 0;
  if (g->fam == CV_GAUSS) g->list.reserve((size_t)cap * sizeof(int2));
  g->shape_hx = g->n_hx;
  g->shape_hy = g->n_hy;
  g->shape_sched = sc;
}

void run_plan(bk_cv_grid *g, const double *hx, const double *hy) {
  hip_check(hipStreamSynchronize(bk::ctx_stream(g->ctx)), "hipStreamSynchronize");
  size_buffers(g);
  upload_consts(g, hx, hy);
  prepare_shape(g);
  const hipStream_t s = g->stream;
  const bool counting = g->count_units;
  if (counting) {
    g->counter.reserve(sizeof(unsigned long long));
    hip_check(hipMemsetAsync(g->counter.ptr, 0, sizeof(unsigned long long), s), "clear the unit counter");
  }
  if (g->use_graph && !g->count_units) {
    if (!g->exec) {
      hip_check(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
      try {
        enqueue_pipeline(g, false);
      } catch (...) {
        hipGraph_t tmp = nullptr;
        (void)hipStreamEndCapture(s, &tmp);
        if (tmp) (void)hipGraphDestroy(tmp);
        throw;
      }
      hip_check(hipStreamEndCapture(s, &g->graph), "hipStreamEndCapture");
      hip_check(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0), "hipGraphInstantiate");
    }
    hip_check(hipEventRecord(g->ev[0], s), "hipEventRecord");
    hip_check(hipGraphLaunch(g->exec, s), "hipGraphLaunch");
    hip_check(hipEventRecord(g->ev[1], s), "hipEventRecord");
  } else {
    hip_check(hipEventRecord(g->ev[0], s), "hipEventRecord");
    enqueue_pipeline(g, true);
    hip_check(hipEventRecord(g->ev[1], s), "hipEventRecord");
  }
  hip_check(hipEventSynchronize(g->ev[1]), "hipEventSynchronize");
  float total = 0.f;
  hip_check(hipEventElapsedTime(&total, g->ev[0], g->ev[1]), "hipEventElapsedTime");
  g->last_total_ms = total;
  if (!g->use_graph || g->count_units) {
    float main_ms = 0.f;
    hip_check(hipEventElapsedTime(&main_ms, g->ev[2], g->ev[3]), "hipEventElapsedTime");
    g->last_main_ms = main_ms;
  } else {
    g->last_main_ms = std::nan("");
  }
  int counts[4] = {0, 0, 0, 0};
  hip_check(hipMemcpy(counts, g->queue.ptr, sizeof(counts), hipMemcpyDeviceToHost), "read counters");
  g->last_rescued = counts[1];
  if (g->fam == CV_GAUSS && counts[1] > g->rescue_cap) fail(BK_ERR_STATE, "internal: rescue list overflow");
  g->last_units = -1;
  if (counting) {
    unsigned long long units = 0;
    hip_check(hipMemcpy(&units, g->counter.ptr, sizeof(units), hipMemcpyDeviceToHost), "read the unit counter");
    g->last_units = (long long)units;
  }
  g->have_run = true;
}

bk_cv_grid *require_plan(bk_cv_grid *g) {
  if (!g) fail(BK_ERR_INVALID_ARG, "null plan");
  return g;
}

}  // namespace

extern "C" {

int bk_cv_grid_create(bk_ctx *ctx, int kind, size_t n_sets, const size_t *sizes, const double *coords,
                      const double *weights, const double *abw, bk_cv_grid **plan_out) {
  return guarded([&] {
    if (!plan_out) fail(BK_ERR_INVALID_ARG, "null output pointer");
    *plan_out = nullptr;
    if (!ctx) fail(BK_ERR_INVALID_ARG, "null context");
    if (kind != BK_GAUSSIAN_PRODUCT_2D && kind != BK_EPANECHNIKOV_PRODUCT_2D)
      fail(BK_ERR_UNSUPPORTED,
           "the cross-validation grid supports BK_GAUSSIAN_PRODUCT_2D and BK_EPANECHNIKOV_PRODUCT_2D");
    if (n_sets == 0 || !sizes) fail(BK_ERR_INVALID_ARG, "need at least one dataset");
    if (!coords || !weights) fail(BK_ERR_INVALID_ARG, "null coordinate or weight array");
    size_t total = 0;
    for (size_t d = 0; d < n_sets; ++d) {
      if (sizes[d] == 0) fail(BK_ERR_INVALID_ARG, "every dataset needs at least one point");
      total += sizes[d];
      if (total > (size_t)(std::numeric_limits<int>::max() - 2 * kCvTile))
        fail(BK_ERR_INVALID_ARG, "too many points in one plan");
    }
    DeviceGuard dg(bk::ctx_device(ctx));
    std::unique_ptr<bk_cv_grid> g(new bk_cv_grid());
    g->ctx = ctx;
    g->device = bk::ctx_device(ctx);
    g->kind = kind;
    g->fam = kind == BK_GAUSSIAN_PRODUCT_2D ? CV_GAUSS : CV_EPAN;
    g->n_sets = (int)n_sets;
    g->n_points = (long long)total;
    g->begin.assign(n_sets + 1, 0);
    for (size_t d = 0; d < n_sets; ++d) g->begin[d + 1] = g->begin[d] + (int)sizes[d];

    std::vector<float4> pts(g->fam == CV_GAUSS ? total : 0);
    std::vector<double4> pts64(g->fam == CV_GAUSS ? 0 : total);
    std::vector<double2> pd(total);
    std::vector<int> set_of(total);
    std::vector<double> w_dev(total);
    g->m_host.resize(total);
    g->perm.resize(total);
    g->m_scale.assign(n_sets, 1.0);
    g->adapt = false;
    for (size_t d = 0; d < n_sets; ++d) {
      double mmax = 0.0;
      for (int i = g->begin[d]; i < g->begin[d + 1]; ++i) {
        if (!std::isfinite(coords[2 * i]) || !std::isfinite(coords[2 * i + 1]))
          fail(BK_ERR_INVALID_ARG, "coordinates must be finite");
      }
      // each dataset goes to the device in Z-curve order (Morton code on a
      // 1024 x 1024 grid over its box, ties by index): consecutive points
      // are close, so a work item's queries and a stage's references span
      // small boxes (the pair kernels skip stages out of reach)
      spatial_order(coords, g->begin[d], g->begin[d + 1], g->perm);
      for (int j = g->begin[d]; j < g->begin[d + 1]; ++j) {
        const int i = g->perm[j];
        const double a = abw ? abw[i] : 1.0;
        if (!(a > 0.0) || !std::isfinite(a)) fail(BK_ERR_INVALID_ARG, "abw must be positive and finite");
        if (!(weights[i] >= 0.0) || !std::isfinite(weights[i]))
          fail(BK_ERR_INVALID_ARG, "weights must be non-negative and finite");
        const float a32 = (float)a;
        if (a32 != 1.0f) g->adapt = true;
        const float m32 = (float)(weights[i] / (a * a));
        g->m_host[j] = (double)m32;
        w_dev[j] = weights[i];
        mmax = std::max(mmax, (double)m32);
        set_of[j] = (int)d;
      }
      const double ms = mmax > 0.0 ? mmax : 1.0;
      g->m_scale[d] = ms;
      for (int j = g->begin[d]; j < g->begin[d + 1]; ++j) {
        const int i = g->perm[j];
        const double a = (double)(float)(abw ? abw[i] : 1.0);
        const double rel = g->m_host[j] / ms;
        const float x = (float)coords[2 * i], y = (float)coords[2 * i + 1];
        if (g->fam == CV_GAUSS)
          pts[j] = make_float4(x, y, rel > 0.0 ? (float)std::log2(rel) : kMaskedLog2, (float)(1.0 / (a * a)));
        else
          pts64[j] = make_double4(x, y, rel, 1.0 / a);
        pd[j] = make_double2(rel, 1.0 / (a * a));
      }
    }
    hipStream_t s = nullptr;
    hip_check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreateWithFlags");
    g->stream = s;
    for (hipEvent_t &e : g->ev) hip_check(hipEventCreate(&e), "hipEventCreate");
    if (g->fam == CV_GAUSS)
      g->pts.reserve(total * sizeof(float4));
    else
      g->pts64.reserve(total * sizeof(double4));
    g->pd.reserve(total * sizeof(double2));
    g->w.reserve(total * sizeof(double));
    g->set_of_point.reserve(total * sizeof(int));
    g->set_begin.reserve((n_sets + 1) * sizeof(int));
    if (g->fam == CV_GAUSS)
      hip_check(hipMemcpyAsync(g->pts.ptr, pts.data(), total * sizeof(float4), hipMemcpyHostToDevice, s), "upload");
    else
      hip_check(hipMemcpyAsync(g->pts64.ptr, pts64.data(), total * sizeof(double4), hipMemcpyHostToDevice, s),
                "upload");
    hip_check(hipMemcpyAsync(g->pd.ptr, pd.data(), total * sizeof(double2), hipMemcpyHostToDevice, s), "upload");
    hip_check(hipMemcpyAsync(g->w.ptr, w_dev.data(), total * sizeof(double), hipMemcpyHostToDevice, s), "upload");
    hip_check(hipMemcpyAsync(g->set_of_point.ptr, set_of.data(), total * sizeof(int), hipMemcpyHostToDevice, s),
              "upload");
    hip_check(hipMemcpyAsync(g->set_begin.ptr, g->begin.data(), (n_sets + 1) * sizeof(int), hipMemcpyHostToDevice,
                             s),
              "upload");
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    choose_splits(g.get());
    check_split_cover(g.get());
    *plan_out = g.release();
  });
}

int bk_cv_grid_destroy(bk_cv_grid *g) {
  return guarded([&] {
    if (!g) return;
    {
      DeviceGuard dg(g->device);
      if (g->stream) (void)hipStreamSynchronize(g->stream);
      drop_graph(g);
      for (DevBuf *b : {&g->pts, &g->pts64, &g->pd, &g->w, &g->set_of_point, &g->set_begin, &g->items, &g->queue, &g->consts,
                        &g->P, &g->list, &g->part, &g->scores})
        b->release();
      for (hipEvent_t e : g->ev)
        if (e) (void)hipEventDestroy(e);
      if (g->stream) (void)hipStreamDestroy(g->stream);
    }
    delete g;
  });
}

int bk_cv_grid_run(bk_cv_grid *g, size_t n_hx, const double *hx, size_t n_hy, const double *hy, double *lcv_out,
                   double *lscv_out) {
  return guarded([&] {
    require_plan(g);
    if (!hx || !hy) fail(BK_ERR_INVALID_ARG, "null bandwidth array");
    if (n_hx < 1 || n_hy < 1 || n_hx > (size_t)kMaxGridAxis || n_hy > (size_t)kMaxGridAxis)
      fail(BK_ERR_INVALID_ARG, "grid axes must hold between 1 and 64 bandwidths");
    // the FP32 exponent constants log2(e)/(2 h^2) stay finite and normal
    auto bad = [](double h) { return !(h >= 1e-15 && h <= 1e15); };
    for (size_t i = 0; i < n_hx; ++i)
      if (bad(hx[i])) fail(BK_ERR_INVALID_ARG, "bandwidths must lie in [1e-15, 1e15]");
    for (size_t i = 0; i < n_hy; ++i)
      if (bad(hy[i])) fail(BK_ERR_INVALID_ARG, "bandwidths must lie in [1e-15, 1e15]");
    DeviceGuard dg(g->device);
    g->n_hx = (int)n_hx;
    g->n_hy = (int)n_hy;
    g->hx_last.assign(hx, hx + n_hx);
    g->hy_last.assign(hy, hy + n_hy);
    g->have_run = false;
    run_plan(g, hx, hy);
    const size_t G = n_hx * n_hy;
    std::vector<double> sc((size_t)g->n_sets * G * 2);
    hip_check(hipMemcpy(sc.data(), g->scores.ptr, sc.size() * sizeof(double), hipMemcpyDeviceToHost),
              "copy scores");
    for (size_t k = 0; k < (size_t)g->n_sets * G; ++k) {
      if (lcv_out) lcv_out[k] = sc[2 * k];
      if (lscv_out) lscv_out[k] = sc[2 * k + 1];
    }
  });
}

int bk_cv_grid_tune(bk_cv_grid *g, size_t n_hx, const double *hx, size_t n_hy, const double *hy, double *ms_out,
                    int *n_candidates_out) {
  return guarded([&] {
    require_plan(g);
    if (g->variant >= 0) fail(BK_ERR_STATE, "BK_CVG_VARIANT forces a tiling; set it to -1 to tune");
    std::vector<Schedule> cands;
    for (int t = 0; t < BK_CVG_VARIANTS; ++t)
      for (int u = 0; u < unroll_count(g->fam); ++u) {
        Schedule sc;
        sc.variant = t;
        sc.unroll = u;
        cands.push_back(sc);
      }
    if (n_candidates_out) *n_candidates_out = (int)cands.size();
    const bool graph = g->use_graph, counting = g->count_units;
    g->use_graph = false;  // eager runs: the pair kernel is timed alone
    g->count_units = false;
    Schedule best;
    double best_ms = std::numeric_limits<double>::max();
    try {
      for (size_t i = 0; i < cands.size(); ++i) {
        g->forced = cands[i];
        const int st = bk_cv_grid_run(g, n_hx, hx, n_hy, hy, nullptr, nullptr);
        if (st != BK_OK) fail(st, bk_last_error());
        if (ms_out) ms_out[i] = g->last_main_ms;
        if (g->last_main_ms < best_ms) {
          best_ms = g->last_main_ms;
          best = cands[i];
        }
      }
    } catch (...) {
      g->forced = Schedule();
      g->use_graph = graph;
      g->count_units = counting;
      throw;
    }
    g->forced = Schedule();
    g->use_graph = graph;
    g->count_units = counting;
    bk::tune_record(tune_key(g), {best.variant, best.unroll});
  });
}

int bk_cv_grid_loo(bk_cv_grid *g, size_t set, size_t ix, size_t iy, double *out) {
  return guarded([&] {
    require_plan(g);
    if (!out) fail(BK_ERR_INVALID_ARG, "null output array");
    if (!g->have_run) fail(BK_ERR_STATE, "no completed run");
    if (set >= (size_t)g->n_sets || ix >= (size_t)g->n_hx || iy >= (size_t)g->n_hy)
      fail(BK_ERR_OUT_OF_RANGE, "dataset or bandwidth index out of range");
    DeviceGuard dg(g->device);
    const int b = g->begin[set], n = g->begin[set + 1] - b;
    const double *src = g->P.as<double>() + (long long)(ix * g->n_hy + iy) * g->n_points + b;
    std::vector<double> tmp(n);
    hip_check(hipMemcpy(tmp.data(), src, n * sizeof(double), hipMemcpyDeviceToHost), "copy densities");
    const double hxv = g->hx_last[ix], hyv = g->hy_last[iy];
    const double C = g->fam == CV_GAUSS ? 1.0 / (2.0 * kPi * hxv * hyv) : 0.5625 / (hxv * hyv);
    const double f = C * g->m_scale[set];
    for (int j = 0; j < n; ++j) out[g->perm[b + j] - b] = tmp[j] * f;  // back to the caller's order
  });
}

int bk_cv_grid_set_option(bk_cv_grid *g, int option, int value) {
  return guarded([&] {
    require_plan(g);
    switch (option) {
      case BK_CVG_VARIANT:
        if (value < -1 || value >= BK_CVG_VARIANTS)
          fail(BK_ERR_INVALID_ARG, "BK_CVG_VARIANT takes -1 or 0.." + std::to_string(BK_CVG_VARIANTS - 1));
        g->variant = value;
        break;
      case BK_CVG_GRAPH:
        if (value != 0 && value != 1) fail(BK_ERR_INVALID_ARG, "BK_CVG_GRAPH takes 0 or 1");
        g->use_graph = value != 0;
        break;
      case BK_CVG_COUNT_UNITS:
        if (value != 0 && value != 1) fail(BK_ERR_INVALID_ARG, "BK_CVG_COUNT_UNITS takes 0 or 1");
        g->count_units = value != 0;
        break;
      case BK_CVG_TUNE_SOURCE:
      case BK_CVG_UNROLL:
      case BK_CVG_WORKGROUPS:
        fail(BK_ERR_INVALID_ARG, "option " + std::to_string(option) + " is read-only");
      default:
        fail(BK_ERR_INVALID_ARG, "unknown option " + std::to_string(option));
    }
  });
}

int bk_cv_grid_get_option(const bk_cv_grid *g, int option, int *value) {
  return guarded([&] {
    if (!g) fail(BK_ERR_INVALID_ARG, "null plan");
    if (!value) fail(BK_ERR_INVALID_ARG, "null output pointer");
    switch (option) {
      case BK_CVG_VARIANT:
        *value = g->used.variant >= 0 ? g->used.variant : g->variant;
        break;
      case BK_CVG_TUNE_SOURCE:
        if (!g->have_run) fail(BK_ERR_STATE, "no completed run");
        *value = g->tune_source;
        break;
      case BK_CVG_UNROLL:
        if (!g->have_run) fail(BK_ERR_STATE, "no completed run");
        *value = g->fam == CV_GAUSS ? kGaussUnrolls[g->used.unroll] : kEpanUnrolls[g->used.unroll];
        break;
      case BK_CVG_WORKGROUPS:
        if (!g->have_run) fail(BK_ERR_STATE, "no completed run");
        *value = g->grid_blocks;
        break;
      case BK_CVG_GRAPH:
        *value = g->use_graph ? 1 : 0;
        break;
      case BK_CVG_COUNT_UNITS:
        *value = g->count_units ? 1 : 0;
        break;
      default:
        fail(BK_ERR_INVALID_ARG, "unknown option " + std::to_string(option));
    }
  });
}

int bk_cv_grid_last_timing(const bk_cv_grid *g, double *main_ms, double *total_ms, long long *rescued) {
  return guarded([&] {
    if (!g) fail(BK_ERR_INVALID_ARG, "null plan");
    if (!g->have_run) fail(BK_ERR_STATE, "no completed run");
    if (main_ms) *main_ms = g->last_main_ms;
    if (total_ms) *total_ms = g->last_total_ms;
    if (rescued) *rescued = g->last_rescued;
  });
}

int bk_cv_grid_unit_count(const bk_cv_grid *g, unsigned long long *evaluated) {
  return guarded([&] {
    if (!g) fail(BK_ERR_INVALID_ARG, "null plan");
    if (!evaluated) fail(BK_ERR_INVALID_ARG, "null output pointer");
    if (!g->have_run || g->last_units < 0) fail(BK_ERR_STATE, "the last run did not count (BK_CVG_COUNT_UNITS)");
    *evaluated = (unsigned long long)g->last_units;
  });
}

}  // extern "C"
