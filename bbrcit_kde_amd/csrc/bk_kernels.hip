// HIP kernels of the direct kernel-density sum for gfx950 (MI355X).
//
// The reference runs one thread per query and walks block-sized tiles of
// AoS reference points in shared memory, accumulating
// sum += m * U(q, r; h*abw) in FP32 (include/CudaDirectKde.h:126-173).
// Here the same sum is organised for CDNA4:
//   * queries are read once, coalesced, from SoA arrays and kept in
//     registers (one per thread on FP32 points, kQpt in the FP64-context
//     Gaussian kernel);
//   * references are pre-transformed once per kernel descriptor into
//     per-point feature vectors, so the reference's per-pair division by
//     h^2 a^2 becomes a multiply;
//   * a workgroup stages kTile feature vectors in LDS (double-buffered, one
//     barrier per stage, the next stage prefetched into registers; the
//     Epanechnikov kernels stage per wavefront, below); every lane of a
//     wavefront reads the same entry, so each LDS read is a broadcast;
//   * Gaussian kinds (FP32): the term m*exp(-c d2 / (h a)^2) is one
//     v_exp_f32 of a base-2 exponent.  On spatially compact tiles the
//     exponent is formed from tile-centred coordinates as
//         C_r + A_r . (q - c) - W_r |q - c|^2_g        (3 FMA per pair)
//     and on the others from the differences q - r directly (5 VALU per
//     pair).  Per-query sums are FP32 within one stage (<= kTile terms)
//     and then added into an FP64 accumulator (block compensation);
//   * Epanechnikov kinds (FP64 arithmetic on the FP32 points): their
//     compact support makes 1 - d2/(h a)^2 cancel near the support edge,
//     which FP32 cannot resolve to 1e-5; each reference carries its scale
//     s = 1/(h a) and offset -s r per axis, so u = s (q - r) is one FP64 FMA;
//     each wavefront walks the references on its own in sub-tiles of 64
//     (its own LDS slot, no workgroup barrier) and passes over every
//     sub-tile whose terms are exactly zero for all its queries;
//   * large reference ranges are split over workgroups at tile boundaries,
//     with a fixed-order FP64 reduction of the partial sums, so results are
//     deterministic.
#include "bk_internal.h"

#include <hip/hip_runtime.h>

#include <type_traits>

namespace bk {

namespace {

constexpr float kMaskedLog = -1.0e30f;  // base-2 log mass of a masked reference
constexpr float kGramMaxRho2 = 4.0f;      // compactness bound of the Gram form
#ifndef BK_FACTORED_RHO2
#define BK_FACTORED_RHO2 9.0f
#endif
constexpr float kFactoredMaxRho2 = BK_FACTORED_RHO2;  // compactness bound of the factored form
constexpr float kFarExp = 100.0f;         // W Q2 of the workgroup-level culling (terms < 2^-100 of the mass)
// A factored (sub-)tile of radius rho about its centre (sqrt of base-2
// exponent units, rho <= 3) contributes, to a query with
// sqrt(W Q2) > kReachUnits + rho, terms below 2^-56 of their own masses: every
// reference is then more than kReachUnits >= sqrt(56) units away.  The
// terms a query loses this way add up to less than 2^-56 of the total mass.
constexpr double kReachUnits = 7.4833148;
// The same bound for the difference form, on the distance to a sub-tile's
// bounding box: W_min d^2 > 56, with 2^-16 of margin for FP32 rounding.
constexpr float kReachDiff = 56.0f * (1.0f + 0x1p-16f);
// FP64 contexts: W Q2 beyond which every term of a tile underflows to exactly
// zero (exp2_f64 of an exponent below -1100 is below 2^-1074 before rounding)
constexpr double kFarExp64 = 1100.0;
// FP64 Gaussian culling keeps the culled part of every sum below 2^-kCullRelBits
// of the sum (see gauss_f64_cull_exponent)
constexpr int kCullRelBits = 50;

// Evaluation form of an aligned tile (tiles[t].z).
constexpr float kFormDirect = 0.0f;     // differences q - r
constexpr float kFormGram = 1.0f;       // tile-centred, per-reference W
constexpr float kFormFactored = 2.0f;   // tile-centred, one W for the tile (tiles[t].w)
constexpr float kFormFactored4 = 3.0f;  // factored about the centres of the tile's four sub-tiles
constexpr int kTileMeta = kGaussTileMeta;  // float4 per tile: (cx, cy, form, W), 4 sub-tiles, box, 4 sub-boxes
constexpr int kBoxSlot = 5;                 // the tile's bounding box
constexpr int kSubBoxSlot = 6;              // the sub-tiles' bounding boxes
constexpr int kSubTile = kGaussTile / 4;
// float4 of a factored tile's packed features: (A0, A1, C) of its kGaussTile
// references, held in the first slots of the tile's block of the centred features
constexpr int kGaussPackedF4 = kGaussTile * 3 / 4;
static_assert(kGaussTile % 4 == 0 && kGaussPackedF4 <= kGaussTile, "packed features fit the tile's block");

__device__ __forceinline__ void write_zeros(double *dst, int q_count) {
  for (int qi = threadIdx.x; qi < q_count; qi += kBlock) dst[qi] = 0.0;
}

// Reference range of split `sp`: splits start at tile-aligned positions.
template <int TILE>
__device__ __forceinline__ void split_bounds(long long r_i, long long r_end,
                                             long long len, int sp,
                                             long long &rb, long long &re) {
  const long long base = (r_i / TILE) * TILE;
  rb = base + (long long)sp * len;
  re = rb + len;
  if (rb < r_i) rb = r_i;
  if (re > r_end) re = r_end;
}

__device__ __forceinline__ float wave_reduce_f(float v, bool is_max) {
  for (int o = 32; o > 0; o >>= 1) {
    const float u = __shfl_xor(v, o, 64);
    v = is_max ? __builtin_fmaxf(v, u) : __builtin_fminf(v, u);
  }
  return v;
}

// Bounding box (x0, x1, y0, y1) of the queries held by a workgroup of kBlock
// threads, NQ per thread (padding lanes repeat a real query).  Every thread
// gets the same four values.
template <int NQ, typename T>
__device__ __forceinline__ void block_query_box(const T *px, const T *py, float (&sbox)[4][kBlock / 64], float &bx0,
                                                float &bx1, float &by0, float &by1) {
  const int t = threadIdx.x;
  float x0 = (float)px[0], x1 = x0, y0 = (float)py[0], y1 = y0;
#pragma unroll
  for (int k = 1; k < NQ; ++k) {
    x0 = __builtin_fminf(x0, (float)px[k]);
    x1 = __builtin_fmaxf(x1, (float)px[k]);
    y0 = __builtin_fminf(y0, (float)py[k]);
    y1 = __builtin_fmaxf(y1, (float)py[k]);
  }
  x0 = wave_reduce_f(x0, false);
  x1 = wave_reduce_f(x1, true);
  y0 = wave_reduce_f(y0, false);
  y1 = wave_reduce_f(y1, true);
  if ((t & 63) == 0) {
    sbox[0][t >> 6] = x0;
    sbox[1][t >> 6] = x1;
    sbox[2][t >> 6] = y0;
    sbox[3][t >> 6] = y1;
  }
  __syncthreads();
  bx0 = sbox[0][0];
  bx1 = sbox[1][0];
  by0 = sbox[2][0];
  by1 = sbox[3][0];
#pragma unroll
  for (int w = 1; w < kBlock / 64; ++w) {
    bx0 = __builtin_fminf(bx0, sbox[0][w]);
    bx1 = __builtin_fmaxf(bx1, sbox[1][w]);
    by0 = __builtin_fminf(by0, sbox[2][w]);
    by1 = __builtin_fmaxf(by1, sbox[3][w]);
  }
}

__device__ __forceinline__ double wave_reduce_d(double v, bool is_max) {
  for (int o = 32; o > 0; o >>= 1) {
    const double u = __shfl_xor(v, o, 64);
    v = is_max ? __builtin_fmax(v, u) : __builtin_fmin(v, u);
  }
  return v;
}

// The same box in FP64 for FP64 queries (exact minima and maxima).
template <int NQ>
__device__ __forceinline__ void block_query_box_d(const double *px, const double *py,
                                                  double (&sbox)[4][kBlock / 64], double &bx0, double &bx1,
                                                  double &by0, double &by1) {
  const int t = threadIdx.x;
  double x0 = px[0], x1 = x0, y0 = py[0], y1 = y0;
#pragma unroll
  for (int k = 1; k < NQ; ++k) {
    x0 = __builtin_fmin(x0, px[k]);
    x1 = __builtin_fmax(x1, px[k]);
    y0 = __builtin_fmin(y0, py[k]);
    y1 = __builtin_fmax(y1, py[k]);
  }
  x0 = wave_reduce_d(x0, false);
  x1 = wave_reduce_d(x1, true);
  y0 = wave_reduce_d(y0, false);
  y1 = wave_reduce_d(y1, true);
  if ((t & 63) == 0) {
    sbox[0][t >> 6] = x0;
    sbox[1][t >> 6] = x1;
    sbox[2][t >> 6] = y0;
    sbox[3][t >> 6] = y1;
  }
  __syncthreads();
  bx0 = sbox[0][0];
  bx1 = sbox[1][0];
  by0 = sbox[2][0];
  by1 = sbox[3][0];
#pragma unroll
  for (int w = 1; w < kBlock / 64; ++w) {
    bx0 = __builtin_fmin(bx0, sbox[0][w]);
    bx1 = __builtin_fmax(bx1, sbox[1][w]);
    by0 = __builtin_fmin(by0, sbox[2][w]);
    by1 = __builtin_fmax(by1, sbox[3][w]);
  }
}

// ---- bounds of a tile's contribution (bounded evaluation) ---------------------
// U at per-axis distances dx, dy >= 0 with bandwidth factor a, in the form the
// evaluation kernels use (kx, ky: gx, gy of FAM_GAUSS / FAM_EPAN /
// FAM_EPANPROD, ix, iy of FAM_EPANCONV).  U(0) = 1; U decreases with each
// distance and increases with a.
template <int FAM>
__device__ __forceinline__ double kernel_u_bound(double dx, double dy, double a, double kx, double ky) {
  const double ia2 = 1.0 / (a * a);
  if constexpr (FAM == FAM_GAUSS) {
    return exp2(-(kx * dx * dx + ky * dy * dy) * ia2);
  } else if constexpr (FAM == FAM_EPAN) {
    return fmax(1.0 - kx * (dx * dx + dy * dy) * ia2, 0.0);
  } else if constexpr (FAM == FAM_EPANPROD) {
    return fmax(1.0 - kx * dx * dx * ia2, 0.0) * fmax(1.0 - ky * dy * dy * ia2, 0.0);
  } else {
    const double ux = dx * kx / a, uy = dy * ky / a;
    const double cx = fmax(2.0 - ux, 0.0), cy = fmax(2.0 - uy, 0.0);
    return (cx * cx * cx * (ux * ux + 6.0 * ux + 4.0) / 32.0) * (cy * cy * cy * (uy * uy + 6.0 * uy + 4.0) / 32.0);
  }
}

// Mass of tile ti and the bounds du >= U >= dl over every (query in the box,
// reference in the tile) pair.
template <int FAM>
__device__ __forceinline__ void tile_bounds(const double4 *__restrict__ tsum, long long ti, double bx0, double bx1,
                                            double by0, double by1, double kx, double ky, double &mt, double &du,
                                            double &dl) {
  const double4 b = tsum[kTileSummary * ti];
  const double4 w = tsum[kTileSummary * ti + 1];
  const double nx = fmax(0.0, fmax(b.x - bx1, bx0 - b.y));
  const double ny = fmax(0.0, fmax(b.z - by1, by0 - b.w));
  const double fx = fmax(b.y - bx0, bx1 - b.x);
  const double fy = fmax(b.w - by0, by1 - b.z);
  mt = w.x;
  du = kernel_u_bound<FAM>(nx, ny, w.z, kx, ky);
  dl = kernel_u_bound<FAM>(fx, fy, w.y, kx, ky);
}

// Bounded evaluation: first tile at or after `tl` (a multiple of T) whose bit
// is set in the block's tile mask (bit t of word t / 64: tile t is summed), or
// a value >= r_end.
template <int T>
__device__ __forceinline__ long long next_listed_tile(const unsigned long long *__restrict__ mask, long long tl,
                                                      long long r_end) {
  while (tl < r_end) {
    const long long ti = tl / T;
    const unsigned long long w = mask[ti >> 6] >> (ti & 63);
    if (w & 1ULL) break;
    tl = (w == 0ULL) ? ((ti >> 6) + 1) * 64 * (long long)T : tl + (long long)__builtin_ctzll(w) * T;
  }
  return tl;
}

// ---- FP32 Gaussian ----------------------------------------------------------
// References are grouped in aligned tiles of kGaussTile points; prep_gauss
// gives every tile a centre c = (cx, cy), a form and (for the factored
// form) the tile's common W.  Per reference:
//   fd (direct form) = (x, y, w, lm)        w = gx W,  W = 1/a_r^2 (FP32-rounded),
//                                           lm = log2(m_r / m_scale)
//   fg (centred)     = (A0, A1, W, C)       A0 = 2 gx W (x - cx), A1 = 2 gy W (y - cy),
//                                           C = lm - W (gx (x - cx)^2 + gy (y - cy)^2)
// The base-2 exponent of the term m_r U(q - r) is
//   lm - W (gx dx^2 + gy dy^2) = C + A0 ax + A1 ay - W Q2,
//   ax = qx - cx, ay = qy - cy, Q2 = gx ax^2 + gy ay^2.
// Forms (rho^2 = max_r W (gx (x-cx)^2 + gy (y-cy)^2), the tile's radius in
// bandwidth units):
//   direct:   5 VALU per pair from the differences q - r (any tile);
//   Gram:     3 FMA per pair (C + A.a - W Q2), rho^2 <= kGramMaxRho2 (rho with
//             each reference's own W); a wavefront whose queries all lie
//             beyond the tile's reach, W_min Q2 > (sqrt(56) + rho_t)^2, passes
//             over the tile (each term left out is below 2^-56 of its mass);
//   factored: one W for the whole tile and rho^2 <= kFactoredMaxRho2: the
//             factor exp2(-W Q2) leaves the tile's sum, 2 FMA per pair.  It
//             is computed per (query, tile) in FP64 (exact differences, split
//             exp2).  Queries with W Q2 > (sqrt(56) + rho_t)^2, rho_t the
//             tile's own radius (stored with the tile, rounded up), take
//             factor 0: each of their terms in the tile is below 2^-56 of its
//             mass (every reference is more than sqrt(56) units, 56 exponent
//             units, away); a wavefront whose queries all take factor 0 skips
//             the tile.  C + A.a = exponent + W Q2 <= W Q2 < (sqrt(56) + 3)^2
//             < 110, so the partial sums stay finite;
//   factored4: the same about the centres of the tile's four aligned
//             sub-tiles of kSubTile references (one factor per query and
//             sub-tile, each sub-tile with its own reach), for tiles too wide
//             for one centre whose sub-tiles each satisfy
//             rho^2 <= kFactoredMaxRho2.
// The rounding error of a centred exponent grows like (rho + distance)^2 in
// bandwidth units, which bounds rho for the centred forms.
// Workgroup-level culling: a tile whose bounding box lies farther than
// W_min Q2 = kFarExp from the bounding box of the workgroup's queries is
// skipped whole (not even staged in LDS).  For the factored forms this only
// skips tiles whose factors are all zero anyway (the box distance is at most
// the centre distance), so no result changes; for the direct and Gram forms
// it drops terms below 2^-100 of the largest mass.  Queries in random order
// give workgroup boxes that span the data and cull nothing; spatially sorted
// queries (the surrogate's sampled check and fallback) cull most tiles.

__device__ __forceinline__ float gauss_direct_term(float qx, float qy, float4 f, float rho) {
  const float dx = qx - f.x;
  const float dy = qy - f.y;
  const float d2 = __builtin_fmaf(dy * rho, dy, dx * dx);
  return __builtin_amdgcn_exp2f(__builtin_fmaf(-f.z, d2, f.w));
}

__device__ __forceinline__ float gauss_direct_term_iso(float qx, float qy, float4 f) {
  const float dx = qx - f.x;
  const float dy = qy - f.y;
  const float d2 = __builtin_fmaf(dy, dy, dx * dx);
  return __builtin_amdgcn_exp2f(__builtin_fmaf(-f.z, d2, f.w));
}

__device__ __forceinline__ float gauss_factored_term(float ax, float ay, float4 f) {
  return __builtin_amdgcn_exp2f(__builtin_fmaf(f.x, ax, __builtin_fmaf(f.y, ay, f.w)));
}

__device__ __forceinline__ float gauss_gram_term(float ax, float ay, float aq, float4 f) {
  float a = __builtin_fmaf(-f.z, aq, f.w);
  a = __builtin_fmaf(f.y, ay, a);
  a = __builtin_fmaf(f.x, ax, a);
  return __builtin_amdgcn_exp2f(a);
}

// 2^e for an FP64 exponent e <= 0: exp2 of the fractional part in FP32
// (|f| <= 1/2, so its FP32 rounding costs < 2^-25 absolute) and the integer
// part applied exactly by an FP64 ldexp (no FP32 underflow below 2^-126;
// the result is normal down to 2^-1000).
__device__ __forceinline__ double exp2_split(double e) {
  double n = rint(e);
  if (n < -1000.0) n = -1000.0;
  const float f = (float)(e - n);
  return __builtin_ldexp((double)__builtin_amdgcn_exp2f(f), (int)n);
}

// Sum of one tile's terms for kGaussQpt queries (self pair excluded when diag).
// FORM: 0 direct, 1 Gram, 2 factored.
template <int FORM, bool ISO, bool DIAG>
__device__ __forceinline__ void gauss_tile_loop(const float4 *__restrict__ sf, int cnt, int self_j,
                                                const float *px, const float *py,
                                                const float *ax, const float *ay, const float *aq, float rho,
                                                float *s) {
#pragma unroll BK_GAUSS_UNROLL
  for (int j = 0; j < cnt; ++j) {
    const float4 f = sf[j];
#pragma unroll
    for (int k = 0; k < kGaussQpt; ++k) {
      float v;
      if constexpr (FORM == 2) {
        v = gauss_factored_term(ax[k], ay[k], f);
      } else if constexpr (FORM == 1) {
        v = gauss_gram_term(ax[k], ay[k], aq[k], f);
      } else if constexpr (ISO) {
        v = gauss_direct_term_iso(px[k], py[k], f);
      } else {
        v = gauss_direct_term(px[k], py[k], f, rho);
      }
      if constexpr (DIAG) {
        s[k] += (j == self_j + k * kBlock) ? 0.0f : v;
      } else {
        s[k] += v;
      }
    }
  }
}

template <int FORM, bool ISO>
__device__ __forceinline__ void gauss_tile(const float4 *__restrict__ sf, int cnt, bool diag, int self_j,
                                           const float *px, const float *py, const float *ax,
                                           const float *ay, const float *aq, float rho, float *s) {
  if (diag)
    gauss_tile_loop<FORM, ISO, true>(sf, cnt, self_j, px, py, ax, ay, aq, rho, s);
  else
    gauss_tile_loop<FORM, ISO, false>(sf, cnt, self_j, px, py, ax, ay, aq, rho, s);
}

// Sum of one factored tile staged in the packed layout: reference j's
// (A0, A1, C) are floats 3j .. 3j+2 of the stage, four references in three
// float4, so a wavefront reads 12 bytes of LDS per reference instead of 16
// (the LDS reads, not the VALU, set the pace of the one-float4-per-pair
// loop).  The terms are those of gauss_factored_term, added in the same
// order.  MASK: only references j0 <= j < j1 contribute (a tile that the
// range covers in part); DIAG: the self pair is left out.
template <bool MASK, bool DIAG>
__device__ __forceinline__ void gauss_packed_loop(const float4 *__restrict__ sf, int j0, int j1, int self_j,
                                                  const float *ax, const float *ay, float *s) {
  const int g0 = MASK ? (j0 >> 2) : 0;
  const int g1 = MASK ? ((j1 + 3) >> 2) : kGaussTile / 4;
  auto term = [&](float a0, float a1, float c, int j) {
#pragma unroll
    for (int k = 0; k < kGaussQpt; ++k) {
      const float v = __builtin_amdgcn_exp2f(__builtin_fmaf(a0, ax[k], __builtin_fmaf(a1, ay[k], c)));
      if constexpr (MASK || DIAG) {
        bool keep = true;
        if constexpr (MASK) keep = j >= j0 && j < j1;
        if constexpr (DIAG) keep = keep && (j != self_j + k * kBlock);
        s[k] += keep ? v : 0.0f;
      } else {
        s[k] += v;
      }
    }
  };
  if constexpr (!MASK && !DIAG) {
    // the common case: all kGaussTile references, a fixed trip count (the
    // compiler hoists the LDS reads of later groups across the unrolled body)
    constexpr int G = kGaussTile / 4;
#pragma unroll 8
    for (int g = 0; g < G; ++g) {
      const float4 a = sf[3 * g], b = sf[3 * g + 1], c = sf[3 * g + 2];
      term(a.x, a.y, a.z, 0);
      term(a.w, b.x, b.y, 0);
      term(b.z, b.w, c.x, 0);
      term(c.y, c.z, c.w, 0);
    }
  } else {
    for (int g = g0; g < g1; ++g) {
      const float4 a = sf[3 * g], b = sf[3 * g + 1], c = sf[3 * g + 2];
      term(a.x, a.y, a.z, 4 * g);
      term(a.w, b.x, b.y, 4 * g + 1);
      term(b.z, b.w, c.x, 4 * g + 2);
      term(c.y, c.z, c.w, 4 * g + 3);
    }
  }
}

// The factored loop over the kSubTile references of sub-tile sb only (a
// whole tile, no self pair).
__device__ __forceinline__ void gauss_packed_sub(const float4 *__restrict__ sf, int sb, const float *ax,
                                                 const float *ay, float *s) {
  constexpr int G = kSubTile / 4;
  const float4 *f = sf + 3 * G * sb;
  auto term = [&](float a0, float a1, float c) {
#pragma unroll
    for (int k = 0; k < kGaussQpt; ++k)
      s[k] += __builtin_amdgcn_exp2f(__builtin_fmaf(a0, ax[k], __builtin_fmaf(a1, ay[k], c)));
  };
#pragma unroll 4
  for (int g = 0; g < G; ++g) {
    const float4 a = f[3 * g], b = f[3 * g + 1], c = f[3 * g + 2];
    term(a.x, a.y, a.z);
    term(a.w, b.x, b.y);
    term(b.z, b.w, c.x);
    term(c.y, c.z, c.w);
  }
}

__device__ __forceinline__ void gauss_packed_tile(const float4 *__restrict__ sf, int j0, int j1, bool diag,
                                                  int self_j, const float *ax, const float *ay, float *s) {
  const bool full = j0 == 0 && j1 == kGaussTile;
  if (full) {
    if (diag)
      gauss_packed_loop<false, true>(sf, j0, j1, self_j, ax, ay, s);
    else
      gauss_packed_loop<false, false>(sf, j0, j1, self_j, ax, ay, s);
  } else {
    if (diag)
      gauss_packed_loop<true, true>(sf, j0, j1, self_j, ax, ay, s);
    else
      gauss_packed_loop<true, false>(sf, j0, j1, self_j, ax, ay, s);
  }
}

// A Gaussian tile is out of reach of a query box when W_min Q2 > kFarExp
// between the boxes (its terms are below 2^-kFarExp of their masses).
__device__ __forceinline__ bool gauss_tile_far(const float4 *__restrict__ tiles, long long tile_idx, float bx0,
                                               float bx1, float by0, float by1, double gx, double gy) {
  const float4 *m = tiles + tile_idx * kTileMeta;
  const float4 c = m[0], b = m[kBoxSlot];
  const double dx = fmax(0.0, fmax((double)c.x - (double)b.x - (double)bx1, (double)bx0 - (double)c.x - (double)b.x));
  const double dy = fmax(0.0, fmax((double)c.y - (double)b.y - (double)by1, (double)by0 - (double)c.y - (double)b.y));
  return (double)b.z * __builtin_fma(gy * dy, dy, gx * dx * dx) > (double)kFarExp;
}

// Thread t's float4 of the LDS stage of the tile starting at tl: a factored
// tile comes packed (kGaussPackedF4 float4; the loop leaves out references
// outside the range), the others one float4 per reference (references
// outside [r_begin, r_end) masked).
__device__ __forceinline__ float4 gauss_stage_load(const float4 *__restrict__ fd, const float4 *__restrict__ fg,
                                                   long long tl, int t, bool centred, bool factored,
                                                   long long r_begin, long long r_end) {
  float4 v = make_float4(0.f, 0.f, 0.f, kMaskedLog);
  const long long r = tl + t;
  if (centred && factored) {
    if (t < kGaussPackedF4) v = fg[r];
  } else if (t < kGaussTile && r >= r_begin && r < r_end) {
    v = centred ? fg[r] : fd[r];
  }
  return v;
}

// BOUNDED: only the tiles listed in bmask (the block's tile mask) are summed;
// the bounds of the others are added by the caller.  COUNT (an instrumented
// variant, never the default launch): count[0] += the pair terms the
// wavefronts evaluate (64 lanes per reference they loop over), count[1] +=
// those of lanes that hold a real query.
template <bool ISO, bool BOUNDED, bool COUNT = false>
__device__ __forceinline__ void gauss_item(
    const float *__restrict__ qx, const float *__restrict__ qy,
    const float4 *__restrict__ fd, const float4 *__restrict__ fg,
    const float4 *__restrict__ tiles, long long r_begin, long long r_end,
    long long q_begin, int q_count, double *__restrict__ dst, double out_scale,
    int self_excl, float rho, double gx, double gy, int mode, const unsigned long long *__restrict__ bmask,
    unsigned long long *__restrict__ count = nullptr) {
  constexpr int T = kGaussTile;
  static_assert(!BOUNDED || kGaussTile == kTile, "tile masks are per kTile references");
  __shared__ float4 sf[2][T];
  const int t = threadIdx.x;
  if (r_begin >= r_end) {
    write_zeros(dst, q_count);
    return;
  }
  __shared__ float sbox[4][kBlock / 64];
  float px[kGaussQpt], py[kGaussQpt];
  double acc[kGaussQpt];
#pragma unroll
  for (int k = 0; k < kGaussQpt; ++k) {
    const int qi = t + k * kBlock;
    const long long q = q_begin + (qi < q_count ? qi : 0);
    px[k] = qx[q];
    py[k] = qy[q];
    acc[k] = 0.0;
  }
  const long long q_lo = q_begin, q_hi = q_begin + q_count;
  const long long q_self0 = q_begin + t;
  unsigned long long evals = 0;  // COUNT: references this wavefront loops over

  float bx0, bx1, by0, by1;
  block_query_box<kGaussQpt>(px, py, sbox, bx0, bx1, by0, by1);
  // the box is the same for every thread: keep it in scalar registers
  bx0 = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(bx0)));
  bx1 = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(bx1)));
  by0 = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(by0)));
  by1 = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(by1)));
  // first tile at or after `tl` that is not culled (or listed), or >= r_end
  auto live_from = [&](long long tl) {
    if constexpr (BOUNDED) return next_listed_tile<T>(bmask, tl, r_end);
    for (; tl < r_end; tl += T)
      if (!gauss_tile_far(tiles, tl / T, bx0, bx1, by0, by1, gx, gy)) break;
    return tl;
  };

  long long tile = live_from((r_begin / T) * T);
  float4 meta = make_float4(0.f, 0.f, kFormDirect, 1.f);
  bool centred = false;
  float4 nf = make_float4(0.f, 0.f, 0.f, kMaskedLog);
  if (tile < r_end) {
    meta = tiles[(tile / T) * kTileMeta];
    centred = (mode == GAUSS_AUTO) && meta.z != kFormDirect;
    nf = gauss_stage_load(fd, fg, tile, t, centred, meta.z == kFormFactored, r_begin, r_end);
  }
  int buf = 0;
  long long next_tile = r_end;
  for (; tile < r_end; tile = next_tile) {
    if (t < T) sf[buf][t] = nf;
    __syncthreads();
    const float4 cur = meta;
    const int form = centred ? (int)cur.z : 0;
    {
      next_tile = live_from(tile + T);
      const long long nt = next_tile;
      nf = make_float4(0.f, 0.f, 0.f, kMaskedLog);
      if (nt < r_end) {
        meta = tiles[(nt / T) * kTileMeta];
        centred = (mode == GAUSS_AUTO) && meta.z != kFormDirect;
        nf = gauss_stage_load(fd, fg, nt, t, centred, meta.z == kFormFactored, r_begin, r_end);
      }
    }
    const int cnt = (int)((r_end - tile) < T ? (r_end - tile) : T);
    const bool diag = self_excl && tile < q_hi && q_lo < tile + cnt;
    // position of this thread's first query in the tile (the self pair of a
    // diagonal tile; |q_self0 - tile| < kGaussTile + kGaussQueriesPerBlock there)
    const int self_j = diag ? (int)(q_self0 - tile) : 0;
    float s[kGaussQpt], ax[kGaussQpt], ay[kGaussQpt], aq[kGaussQpt];
    double fac[kGaussQpt];
#pragma unroll
    for (int k = 0; k < kGaussQpt; ++k) {
      s[k] = 0.0f;
      fac[k] = 1.0;
      ax[k] = ay[k] = aq[k] = 0.0f;
    }
    if (form == 3) {
      // factored about each sub-tile's centre: one FP64 factor per query and sub-tile
      const float4 *subc = tiles + (tile / T) * kTileMeta + 1;
      for (int sb = 0; sb * kSubTile < cnt; ++sb) {
        const int c0 = sb * kSubTile;
        const int cn = (cnt - c0) < kSubTile ? (cnt - c0) : kSubTile;
        const float4 sc = subc[sb];
        double sfac[kGaussQpt];
        bool any_near = false;
#pragma unroll
        for (int k = 0; k < kGaussQpt; ++k) {
          const double dx = (double)px[k] - (double)sc.x;
          const double dy = (double)py[k] - (double)sc.y;
          const double e = (double)cur.w * __builtin_fma(gy * dy, dy, gx * dx * dx);
          const bool near = e <= (double)sc.z;  // the sub-tile's reach
          any_near |= near;
          sfac[k] = near ? exp2_split(-e) : 0.0;
          ax[k] = near ? (float)dx : 0.0f;
          ay[k] = near ? (float)dy : 0.0f;
          s[k] = 0.0f;
        }
        // a factor of zero for every query of the wavefront: the sub-tile adds nothing
        if (__any(any_near)) {
          gauss_tile<2, ISO>(sf[buf] + c0, cn, diag, self_j - c0, px, py, ax, ay, aq, rho, s);
          if constexpr (COUNT) evals += (unsigned long long)cn;
        }
#pragma unroll
        for (int k = 0; k < kGaussQpt; ++k) acc[k] = __builtin_fma((double)s[k], sfac[k], acc[k]);
      }
      buf ^= 1;
      continue;
    }
    bool any_near = false;
    // the tile's reach (centred forms): kept in the first sub-tile entry; the
    // Gram form, whose references have their own W, tests the tile's smallest W
    const double reach = form != 0 ? (double)tiles[(tile / T) * kTileMeta + 1].w : 0.0;
    const double w_min = form == 1 ? (double)tiles[(tile / T) * kTileMeta + kBoxSlot].z : 0.0;
    if (form != 0) {
#pragma unroll
      for (int k = 0; k < kGaussQpt; ++k) {
        const double dx = (double)px[k] - (double)cur.x;
        const double dy = (double)py[k] - (double)cur.y;
        const double q2 = __builtin_fma(gy * dy, dy, gx * dx * dx);
        ax[k] = (float)dx;
        ay[k] = (float)dy;
        aq[k] = (float)q2;
        if (form == 2) {
          const double e = (double)cur.w * q2;
          const bool near = e <= reach;
          any_near |= near;
          fac[k] = near ? exp2_split(-e) : 0.0;
          ax[k] = near ? ax[k] : 0.0f;
          ay[k] = near ? ay[k] : 0.0f;
        } else {
          // Gram form: W_r |q - r|^2 >= (sqrt(W_min Q2) - rho_t)^2 for every
          // reference, so beyond the reach each term is below 2^-56 of its mass
          any_near |= w_min * q2 <= reach;
        }
      }
    }
    if (form == 2) {
      // a factor of zero for every query of the wavefront: the tile adds nothing
      if (__any(any_near)) {
        const int j0 = tile < r_begin ? (int)(r_begin - tile) : 0;
        unsigned subs = 0xFu;
        if (j0 == 0 && cnt == T && !diag) {
          // the same rule per sub-tile of kSubTile references, about its own
          // centre and with its own reach: a sub-tile none of whose queries
          // is within reach is passed over (the FP32 exponent is compared
          // with a margin of 2^-16 for its rounding)
          const float4 *subc = tiles + (tile / T) * kTileMeta + 1;
          const float gxf = (float)gx, gyf = (float)gy;
          subs = 0u;
#pragma unroll
          for (int sb = 0; sb < kGaussTile / kSubTile; ++sb) {
            const float4 sc = subc[sb];
            bool near_sub = false;
#pragma unroll
            for (int k = 0; k < kGaussQpt; ++k) {
              const float dx = px[k] - sc.x, dy = py[k] - sc.y;
              near_sub |= cur.w * __builtin_fmaf(gyf * dy, dy, gxf * dx * dx) <= sc.z * (1.0f + 0x1p-16f);
            }
            subs |= __any(near_sub) ? 1u << sb : 0u;
          }
        }
        if (subs == 0xFu) {
          gauss_packed_tile(sf[buf], j0, cnt, diag, self_j, ax, ay, s);
          if constexpr (COUNT) evals += (unsigned long long)cnt;
        } else {
          for (int sb = 0; sb < kGaussTile / kSubTile; ++sb)
            if ((subs >> sb) & 1u) {
              gauss_packed_sub(sf[buf], sb, ax, ay, s);
              if constexpr (COUNT) evals += (unsigned long long)kSubTile;
            }
        }
      }
    } else if (form == 1) {
      // no query of the wavefront within the tile's reach: the tile is passed over
      if (__any(any_near)) {
        gauss_tile<1, ISO>(sf[buf], cnt, diag, self_j, px, py, ax, ay, aq, rho, s);
        if constexpr (COUNT) evals += (unsigned long long)cnt;
      }
    } else {
      // difference form, per sub-tile of kSubTile references: a wavefront
      // passes over a sub-tile when all its queries lie beyond the sub-tile's
      // reach, W_min d^2 > 56 with d the distance to its bounding box (every
      // reference's own W is at least W_min, so each term left out is below
      // 2^-56 of its mass; the FP32 distance is compared with a margin of
      // 2^-16 for its rounding)
      const float4 *subc = tiles + (tile / T) * kTileMeta + 1;
      const float4 *subb = tiles + (tile / T) * kTileMeta + kSubBoxSlot;
      const float w_lo = tiles[(tile / T) * kTileMeta + kBoxSlot].z;
      const float gxf = (float)gx, gyf = (float)gy;
      for (int sb = 0; sb * kSubTile < cnt; ++sb) {
        const int c0 = sb * kSubTile;
        const int cn = (cnt - c0) < kSubTile ? (cnt - c0) : kSubTile;
        const float4 sc = subc[sb], hb = subb[sb];
        bool near_any = false;
#pragma unroll
        for (int k = 0; k < kGaussQpt; ++k) {
          const float dx = __builtin_fmaxf(__builtin_fabsf(px[k] - sc.x) - hb.x, 0.0f);
          const float dy = __builtin_fmaxf(__builtin_fabsf(py[k] - sc.y) - hb.y, 0.0f);
          near_any |= w_lo * __builtin_fmaf(gyf * dy, dy, gxf * dx * dx) <= kReachDiff;
        }
        if (__any(near_any)) {
          gauss_tile<0, ISO>(sf[buf] + c0, cn, diag, self_j - c0, px, py, ax, ay, aq, rho, s);
          if constexpr (COUNT) evals += (unsigned long long)cn;
        }
      }
    }
#pragma unroll
    for (int k = 0; k < kGaussQpt; ++k) acc[k] = __builtin_fma((double)s[k], fac[k], acc[k]);
    buf ^= 1;
  }
#pragma unroll
  for (int k = 0; k < kGaussQpt; ++k) {
    const int qi = t + k * kBlock;
    if (qi < q_count) dst[qi] = acc[k] * out_scale;
  }
  if constexpr (COUNT) {
    int real = 0;
#pragma unroll
    for (int k = 0; k < kGaussQpt; ++k) real += __popcll(__ballot(t + k * kBlock < q_count));
    if ((t & 63) == 0 && evals > 0) {
      atomicAdd(count, evals * 64ULL * (unsigned long long)kGaussQpt);
      atomicAdd(count + 1, evals * (unsigned long long)real);
    }
  }
}

// Minimum waves per SIMD of the range kernel: caps its registers at 64 so
// eight workgroups fit on a CU (measured best at the 1M x 1M workload, with
// one query per thread and the reference loop unrolled twice).
#ifndef BK_GAUSS_WAVES
#define BK_GAUSS_WAVES 8
#endif

template <bool ISO, bool BOUNDED, bool COUNT = false>
__global__ __launch_bounds__(kBlock, BK_GAUSS_WAVES) void gauss_range_kernel(
    const float *__restrict__ qx, const float *__restrict__ qy,
    const float4 *__restrict__ fd, const float4 *__restrict__ fg,
    const float4 *__restrict__ tiles, long long r_i, long long r_end,
    long long split_len, long long q_i, long long nq, double *__restrict__ dst,
    double out_scale, int self_excl, float rho, double gx, double gy, int mode, BoundArgs ba) {
  long long bx = blockIdx.x, by = blockIdx.y;
  if (ba.order) {
    const int it = ba.order[blockIdx.x];
    bx = it % ba.nblocks;
    by = it / ba.nblocks;
  }
  const long long qb = bx * kGaussQueriesPerBlock;
  long long rb, re;
  split_bounds<kGaussTile>(r_i, r_end, split_len, (int)by, rb, re);
  const long long left = nq - qb;
  const int q_count = (int)(left < kGaussQueriesPerBlock ? left : kGaussQueriesPerBlock);
  const long long off = by * nq + qb;
  gauss_item<ISO, BOUNDED, COUNT>(qx, qy, fd, fg, tiles, rb, re, q_i + qb, q_count, dst + off, out_scale, self_excl,
                                  rho, gx, gy, mode, BOUNDED ? ba.mask + bx * ba.nwords : nullptr, ba.count);
}

template <bool ISO>
__global__ __launch_bounds__(kBlock) void gauss_list_kernel(
    const float *__restrict__ qx, const float *__restrict__ qy,
    const float4 *__restrict__ fd, const float4 *__restrict__ fg,
    const float4 *__restrict__ tiles, const WorkItem *__restrict__ items,
    double *__restrict__ dst, double out_scale, int self_excl, float rho,
    double gx, double gy, int mode) {
  const WorkItem it = items[blockIdx.x];
  gauss_item<ISO, false>(qx, qy, fd, fg, tiles, it.r_begin, it.r_end, it.q_begin, it.q_count,
                         dst + it.out_index, out_scale, self_excl, rho, gx, gy, mode, nullptr);
}

// Block-wide min / max over N threads (one value per thread).
template <int N, bool MAX>
__device__ float block_reduce_f(float v, float *sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int o = N / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      const float a = sh[threadIdx.x], b = sh[threadIdx.x + o];
      sh[threadIdx.x] = MAX ? __builtin_fmaxf(a, b) : __builtin_fminf(a, b);
    }
    __syncthreads();
  }
  const float r = sh[0];
  __syncthreads();
  return r;
}

// Half width of [lo, hi] about c, rounded up so the box holds every point.
__device__ __forceinline__ float half_width_up(float lo, float hi, float c) {
  const double hw = fmax((double)hi - (double)c, (double)c - (double)lo);
  float f = (float)hw;
  if ((double)f < hw) f = __uint_as_float(__float_as_uint(f) + 1u);  // next float up (f >= 0)
  return f;
}

// W Q2 beyond which every term of a factored (sub-)tile whose radius^2 about
// its centre is rho2 lies below 2^-56 of its mass: (kReachUnits + rho)^2,
// rounded up (rho2 itself may lie up to half an FP32 ulp low).
__device__ __forceinline__ float gauss_reach_exp(float rho2) {
  const double r = sqrt((double)rho2) * (1.0 + 0x1p-20) + 0x1p-20;
  const double t = (kReachUnits + r) * (kReachUnits + r);
  float f = (float)t;
  if ((double)f < t) f = __uint_as_float(__float_as_uint(f) + 1u);  // next float up (t > 0)
  return f;
}

// One workgroup of kGaussTile threads per aligned tile; each wavefront is one
// sub-tile of kSubTile references.
__global__ __launch_bounds__(kGaussTile) void prep_gauss_kernel(
    EvalParams p, const float *__restrict__ x, const float *__restrict__ y,
    const float *__restrict__ m, const float *__restrict__ abw, long long n,
    double inv_m_scale, float4 *__restrict__ fd, float4 *__restrict__ fg,
    float4 *__restrict__ tiles) {
  constexpr int T = kGaussTile;
  static_assert(kSubTile == 64, "one wavefront per sub-tile");
  __shared__ float sh[T];
  __shared__ float s_form;
  const long long i = (long long)blockIdx.x * T + threadIdx.x;
  const bool valid = i < n;
  const float xi = valid ? x[i] : 0.0f;
  const float yi = valid ? y[i] : 0.0f;
  const float big = 3.0e38f;
  const float xmin = block_reduce_f<T, false>(valid ? xi : big, sh);
  const float xmax = block_reduce_f<T, true>(valid ? xi : -big, sh);
  const float ymin = block_reduce_f<T, false>(valid ? yi : big, sh);
  const float ymax = block_reduce_f<T, true>(valid ? yi : -big, sh);
  const float cx = 0.5f * (xmin + xmax);
  const float cy = 0.5f * (ymin + ymax);
  // sub-tile centre (a sub-tile without points gets the tile's centre)
  const float sxmin = wave_reduce_f(valid ? xi : big, false), sxmax = wave_reduce_f(valid ? xi : -big, true);
  const float symin = wave_reduce_f(valid ? yi : big, false), symax = wave_reduce_f(valid ? yi : -big, true);
  const bool sub_empty = sxmin > sxmax;
  const float scx = sub_empty ? cx : 0.5f * (sxmin + sxmax);
  const float scy = sub_empty ? cy : 0.5f * (symin + symax);
  float rho2 = 0.0f, rho2_sub = 0.0f;
  float w_lo = big, w_hi = -big;
  double W = 1.0, lm = 0.0, mass = 0.0;
  float Wf = 1.0f;
  if (valid) {
    const double a = (double)abw[i];
    Wf = (float)(1.0 / (a * a));
    W = (double)Wf;
    mass = (double)m[i] * inv_m_scale;
    lm = mass > 0.0 ? log2(mass) : (double)kMaskedLog;
    fd[i] = make_float4(xi, yi, (float)(p.gx * W), (float)lm);
    const double rx = (double)xi - (double)cx, ry = (double)yi - (double)cy;
    rho2 = (float)(W * (p.gx * rx * rx + p.gy * ry * ry));
    const double sx = (double)xi - (double)scx, sy = (double)yi - (double)scy;
    rho2_sub = (float)(W * (p.gx * sx * sx + p.gy * sy * sy));
    w_lo = w_hi = Wf;
  }
  const float rmax = block_reduce_f<T, true>(rho2, sh);
  const float rsub = block_reduce_f<T, true>(rho2_sub, sh);
  const float rsub_own = wave_reduce_f(rho2_sub, true);  // this wavefront's sub-tile
  const float wmin = block_reduce_f<T, false>(w_lo, sh);
  const float wmax = block_reduce_f<T, true>(w_hi, sh);
  if (threadIdx.x == 0) {
    float form = kFormDirect;
    if (wmin == wmax && rmax <= kFactoredMaxRho2) {
      form = kFormFactored;
    } else if (wmin == wmax && rsub <= kFactoredMaxRho2) {
      form = kFormFactored4;
    } else if (rmax <= kGramMaxRho2) {
      form = kFormGram;
    }
    s_form = form;
    tiles[blockIdx.x * kTileMeta] = make_float4(cx, cy, form, wmax);
  }
  if (threadIdx.x % kSubTile == 0) {
    tiles[blockIdx.x * kTileMeta + 1 + threadIdx.x / kSubTile] =
        make_float4(scx, scy, gauss_reach_exp(rsub_own), gauss_reach_exp(rmax));
    // the sub-tile's box about its centre (a single point for a sub-tile without points)
    tiles[blockIdx.x * kTileMeta + kSubBoxSlot + threadIdx.x / kSubTile] =
        sub_empty ? make_float4(0.0f, 0.0f, 0.0f, 0.0f)
                  : make_float4(half_width_up(sxmin, sxmax, scx), half_width_up(symin, symax, scy), 0.0f, 0.0f);
  }
  if (threadIdx.x == 0) {
    tiles[blockIdx.x * kTileMeta + kBoxSlot] =
        make_float4(half_width_up(xmin, xmax, cx), half_width_up(ymin, ymax, cy), wmin, 0.0f);
  }
  __syncthreads();
  if (s_form == kFormFactored) {
    // packed: (A0, A1, C) at floats 3j .. 3j+2 of the tile's block; the
    // slots of references past n are masked (a zero term)
    float *pk = reinterpret_cast<float *>(fg + (long long)blockIdx.x * T) + 3 * threadIdx.x;
    float a0 = 0.0f, a1 = 0.0f, c = kMaskedLog;
    if (valid) {
      const double rx = (double)xi - (double)cx;
      const double ry = (double)yi - (double)cy;
      const double q = p.gx * rx * rx + p.gy * ry * ry;
      a0 = (float)(2.0 * p.gx * W * rx);
      a1 = (float)(2.0 * p.gy * W * ry);
      c = mass > 0.0 ? (float)(lm - W * q) : kMaskedLog;
    }
    pk[0] = a0;
    pk[1] = a1;
    pk[2] = c;
  } else if (valid) {
    // centred features about the tile's or the sub-tile's centre
    const bool sub = s_form == kFormFactored4;
    const double rx = (double)xi - (double)(sub ? scx : cx);
    const double ry = (double)yi - (double)(sub ? scy : cy);
    const double q = p.gx * rx * rx + p.gy * ry * ry;
    fg[i] = make_float4((float)(2.0 * p.gx * W * rx), (float)(2.0 * p.gy * W * ry), Wf,
                        mass > 0.0 ? (float)(lm - W * q) : kMaskedLog);
  }
}

// ---- FP64 Epanechnikov kinds on FP32 points ----------------------------------
// Per reference the scale s = 1 / (h a) of each axis and the scaled offset
// o = -s r, so that the scaled difference u = s (q - r) is one FMA,
// fma(s, q, o):
//   FAM_EPAN:     f0 = (s, -s x, -s y, m)
//   FAM_EPANPROD: f0 = (sx, -sx x, sy, -sy y), f1.x = m
//   FAM_EPANCONV: f0 = (sx, -sx x, sy, -sy y), f1.x = m * 2^-10
// u carries at most 2^-53 (|s r| + |u|) of rounding, ~1e-16 of the
// coordinates in bandwidth units; the support edge 1 - |u|^2 is then formed
// by FMAs in FP64.
template <int FAM>
struct EpanNf {
  static constexpr int kNf = (FAM == FAM_EPAN) ? 1 : 2;
};

template <int FAM>
__device__ __forceinline__ double epan_term(double qx, double qy, double4 a, double4 b) {
  if constexpr (FAM == FAM_EPAN) {
    const double ux = __builtin_fma(a.x, qx, a.y);
    const double uy = __builtin_fma(a.x, qy, a.z);
    return a.w * clamp_unit(__builtin_fma(-ux, ux, __builtin_fma(-uy, uy, 1.0)));
  } else if constexpr (FAM == FAM_EPANPROD) {
    const double ux = __builtin_fma(a.x, qx, a.y);
    const double uy = __builtin_fma(a.z, qy, a.w);
    const double tx = clamp_unit(__builtin_fma(-ux, ux, 1.0));
    const double ty = clamp_unit(__builtin_fma(-uy, uy, 1.0));
    return b.x * (tx * ty);
  } else {
    const double ux = __builtin_fabs(__builtin_fma(a.x, qx, a.y));
    const double uy = __builtin_fabs(__builtin_fma(a.z, qy, a.w));
    const double cx = __builtin_fmax(2.0 - ux, 0.0);
    const double cy = __builtin_fmax(2.0 - uy, 0.0);
    const double kx = cx * cx * cx * __builtin_fma(ux, ux + 6.0, 4.0);
    const double ky = cy * cy * cy * __builtin_fma(uy, uy + 6.0, 4.0);
    return b.x * (kx * ky);
  }
}

// Per aligned tile of kTile references, two double4: the bounding box
// (x0, x1, y0, y1) of the FP32 coordinates, and per axis the smallest scale
// s (rounded down to FP32) and the largest |o| (rounded up to FP32).
// A tile is culled for a workgroup when every term is exactly zero, as
// computed, for the whole box of its queries.
//
// u_lower bounds |u| = |fma(s, q, o)| of every (query, reference) pair of
// the two boxes on one axis from below, up to its rounding: with G = s_min
// times the gap between the boxes (G <= s |q - r| exactly), u_lower <=
// G - 2^-50 (1 + o_max), while the computed |u| >= G (1 - 2^-53) -
// 2^-52 o_max.  So u_lower >= c (c >= 1) gives |u| > c, and u_lower <= 4
// gives |u| >= u_lower.
__device__ __forceinline__ double u_lower(double r0, double r1, double b0, double b1, double s_min, double o_max) {
  const double gap = fmax(0.0, fmax(r0 - b1, b0 - r1));
  return fmax(0.0, __builtin_fma(s_min * gap, 1.0 - 0x1p-50, -0x1p-50 * (1.0 + o_max)));
}

// Every term is exactly zero once the lower bounds lx, ly of |ux|, |uy|
// reach the support edge of the family.
template <int FAM>
__device__ __forceinline__ bool epan_beyond_support(double lx, double ly) {
  if constexpr (FAM == FAM_EPAN) {
    // t = fma(-ux, ux, fma(-uy, uy, 1)) <= 0 once ux^2 + uy^2 >= 1 + 2^-52;
    // ux^2 + uy^2 >= lx^2 + ly^2 when both bounds are <= 4
    if (lx >= 1.5 || ly >= 1.5) return true;
    return __builtin_fma(lx, lx, ly * ly) >= 1.0 + 0x1p-48;
  } else if constexpr (FAM == FAM_EPANPROD) {
    return lx >= 1.0 || ly >= 1.0;  // 1 - u^2 <= 0 on that axis
  } else {
    return lx >= 2.0 || ly >= 2.0;  // 2 - u <= 0 on that axis
  }
}

template <int FAM>
__device__ __forceinline__ bool epan_tile_far(const double4 *__restrict__ tb, long long tile_idx, float bx0, float bx1,
                                              float by0, float by1) {
  const double4 b = tb[kEpanTileMeta * tile_idx];
  const double4 w = tb[kEpanTileMeta * tile_idx + 1];
  return epan_beyond_support<FAM>(u_lower(b.x, b.y, bx0, bx1, w.x, w.z), u_lower(b.z, b.w, by0, by1, w.y, w.w));
}

// Reach box (L_x, H_x, L_y, H_y) of a tile or sub-tile: a query with
// q_x <= L_x or q_x >= H_x (likewise on y) has, as computed, |u| > c on that
// axis with every reference of it, c being the support edge of |u| on one
// axis (2 for the convolution kernel, else 1), so each of its terms there is
// exactly zero (for the radial kernel too: ux^2 + uy^2 > 1 then).  With s_min
// the smallest scale and o_max the largest |o|, |s q + o| >= s_min |q - r| -
// 2^-52 o_max, so a reach r = c (1 + 2^-20 + 2^-50 o_max) / s_min past the
// box suffices.  The bounds carry their own rounding outward and are then
// rounded outward to FP32, the precision of the query coordinates.
// The box [x0, x1] x [y0, y1] widened by rx and ry, outward-rounded to
// FP32: below(lo, r) <= lo - r and above(hi, r) >= hi + r however the FP64
// operations round.
__device__ __forceinline__ float4 reach_box_f32(double x0, double x1, double y0, double y1, double rx, double ry) {
  // largest FP32 value <= v, smallest >= v (v finite; the float nearest v
  // is then finite or an infinity of v's sign, so the step stays ordered)
  auto below = [](double lo, double r) {
    const double v = (lo - r) - 0x1p-50 * (fabs(lo) + r);
    const float f = (float)v;
    if (!((double)f > v)) return f;
    return f == 0.0f ? -0x1p-149f : __uint_as_float(__float_as_uint(f) + (f > 0.0f ? -1 : 1));
  };
  auto above = [](double hi, double r) {
    const double v = (hi + r) + 0x1p-50 * (fabs(hi) + r);
    const float f = (float)v;
    if (!((double)f < v)) return f;
    return f == 0.0f ? 0x1p-149f : __uint_as_float(__float_as_uint(f) + (f > 0.0f ? 1 : -1));
  };
  return make_float4(below(x0, rx), above(x1, rx), below(y0, ry), above(y1, ry));
}

__device__ __forceinline__ float4 epan_reach_box(float x0, float x1, float y0, float y1, float sx, float sy,
                                                 float ox, float oy, double c) {
  const double rx = c * (1.0 + 0x1p-20 + 0x1p-50 * (double)ox) / (double)sx;
  const double ry = c * (1.0 + 0x1p-20 + 0x1p-50 * (double)oy) / (double)sy;
  return reach_box_f32(x0, x1, y0, y1, rx, ry);
}

// Reach boxes of tile ti: [0] the tile's, [1 + s] that of its sub-tile s
// (stride kEpanTileMeta on FP32 points, kF64TileMeta in FP64 contexts).
template <bool F64>
__device__ __forceinline__ const float4 *epan_reach(const double4 *__restrict__ tb, long long ti) {
  return reinterpret_cast<const float4 *>(tb + (F64 ? kF64TileMeta : kEpanTileMeta) * ti + 2);
}
// FP32 coordinate boxes (x0, x1, y0, y1) of tile ti's sub-tiles (FP32 points).
__device__ __forceinline__ const float4 *epan_sub_boxes(const double4 *__restrict__ tb, long long ti) {
  return reinterpret_cast<const float4 *>(tb + kEpanTileMeta * ti + 2 + (kEpanReachF4 + 1) / 2);
}

// True when no query of the wavefront is inside the reach box e on both
// axes: every term of the (sub-)tile is then exactly zero for all its
// queries, and passing over it leaves the sums bit-identical.  A NaN
// coordinate compares false and keeps the tile.
// (The comparisons are taken as lane masks, so the test is four compares
// per query and scalar mask operations.)
// FP64 queries compare through FP32 values rounded up (xu >= x) and down
// (xd <= x), so a far verdict holds for the FP64 coordinate.
template <int NQ>
__device__ __forceinline__ bool epan_wave_far(float4 e, const float *xu, const float *xd, const float *yu,
                                              const float *yd) {
  unsigned long long far = ~0ull;
#pragma unroll
  for (int k = 0; k < NQ; ++k)
    far &= __ballot(xu[k] <= e.x) | __ballot(xd[k] >= e.y) | __ballot(yu[k] <= e.z) | __ballot(yd[k] >= e.w);
  return (__builtin_amdgcn_read_exec() & ~far) == 0ull;
}

// FP32 values just above and below an FP64 coordinate.
__device__ __forceinline__ float f32_up(double v) {
  const float f = (float)v;
  return (double)f >= v ? f : (f == 0.0f ? 0x1p-149f : __uint_as_float(__float_as_uint(f) + (f > 0.0f ? 1 : -1)));
}
__device__ __forceinline__ float f32_down(double v) {
  const float f = (float)v;
  return (double)f <= v ? f : (f == 0.0f ? -0x1p-149f : __uint_as_float(__float_as_uint(f) + (f > 0.0f ? -1 : 1)));
}

// Wavefront-scope barrier: orders this wavefront's LDS writes before its
// later reads (and the reads before later writes) in the compiled code; the
// hardware keeps one wavefront's LDS operations in order.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int FAM>
__device__ __forceinline__ double pair_term_f64(double qx, double qy, double4 a, double4 b, double rho);

// One workgroup, kEpanQpt queries per thread.  Each wavefront walks the
// reference range on its own, with no workgroup barrier: it tests the reach
// boxes of kEpanScanBatch tiles at a time (and, for a tile in reach, those of
// its sub-tiles), stages each sub-tile of kEpanSub references in reach in its
// own LDS slot, one reference per lane, and prefetches the next into
// registers.  Terms are summed in reference order, so the sums equal those
// over every pair in that order.  UNI: every reference has the scales
// (usx, usy) (no adaptive bandwidths), so only two coordinates (offsets, or
// FP64 points) and the mass are staged.  In a bounded evaluation only the
// tiles listed in bmask are visited.  F64: FP64 points and features
// (pair_term_f64's layout); else FP32 points and epan_term's layout.
template <int FAM, bool BOUNDED, bool UNI, bool F64>
__device__ __forceinline__ void epan_item(
    const std::conditional_t<F64, double, float> *__restrict__ qx,
    const std::conditional_t<F64, double, float> *__restrict__ qy,
    const double4 *__restrict__ f0, const double4 *__restrict__ f1, const double4 *__restrict__ tb,
    long long r_begin, long long r_end, long long q_begin, int q_count,
    double *__restrict__ dst, double out_scale, int self_excl, const unsigned long long *__restrict__ bmask,
    double usx, double usy, unsigned long long *__restrict__ count = nullptr) {
  constexpr int kNf = EpanNf<FAM>::kNf;
  constexpr int kWaves = kBlock / 64;
  constexpr int kSubs = kTile / kEpanSub;
  // the staged features: all of f0 (and the mass of f1), or with UNI two
  // coordinates and the mass
  __shared__ double4 s0[kWaves][UNI ? 1 : kEpanSub];
  __shared__ double2 so[kWaves][UNI ? kEpanSub : 1];
  __shared__ double sm[kWaves][(kNf == 2 || UNI) ? kEpanSub : 1];
  const int t = threadIdx.x, wv = t >> 6, lane = t & 63;
  if (r_begin >= r_end) {
    write_zeros(dst, q_count);
    return;
  }
  double px[kEpanQpt], py[kEpanQpt], acc[kEpanQpt];
  float xu[kEpanQpt], xd[kEpanQpt], yu[kEpanQpt], yd[kEpanQpt];
#pragma unroll
  for (int k = 0; k < kEpanQpt; ++k) {
    const int qi = t + k * kBlock;
    const long long q = q_begin + (qi < q_count ? qi : 0);
    px[k] = (double)qx[q];
    py[k] = (double)qy[q];
    if constexpr (F64) {
      xu[k] = f32_up(px[k]);
      xd[k] = f32_down(px[k]);
      yu[k] = f32_up(py[k]);
      yd[k] = f32_down(py[k]);
    } else {
      xu[k] = xd[k] = qx[q];
      yu[k] = yd[k] = qy[q];
    }
    acc[k] = 0.0;
  }
  long long batch = -1, stile = -1;
  unsigned near_t = 0, near_s = 0;
  // first sub-tile at or after `sub` (a multiple of kEpanSub) that this
  // wavefront evaluates, or r_end
  auto next_needed = [&](long long sub) __attribute__((always_inline)) {
    for (;;) {
      if (sub >= r_end) return r_end;
      long long ti = sub / kTile;
      const long long b = ti - ti % kEpanScanBatch;
      if (b != batch) {
        unsigned listed = (1u << kEpanScanBatch) - 1;
        if constexpr (BOUNDED) {
          listed &= (unsigned)(bmask[b >> 6] >> (b & 63));
          if (listed == 0) {  // no listed tile in this batch: on to the next listed one
            sub = next_listed_tile<kTile>(bmask, (b + kEpanScanBatch) * kTile, r_end);
            continue;
          }
        }
        batch = b;
        float4 e[kEpanScanBatch];
#pragma unroll
        for (int i = 0; i < kEpanScanBatch; ++i) e[i] = epan_reach<F64>(tb, b + i)[0];
        unsigned m = 0;
#pragma unroll
        for (int i = 0; i < kEpanScanBatch; ++i) m |= epan_wave_far<kEpanQpt>(e[i], xu, xd, yu, yd) ? 0u : 1u << i;
        near_t = m & listed;
      }
      const unsigned rest = near_t >> (int)(ti - b);
      if (rest == 0) {
        sub = (b + kEpanScanBatch) * kTile;
        continue;
      }
      const int skip = __builtin_ctz(rest);
      if (skip) {
        ti += skip;
        sub = ti * kTile;
      }
      if (ti != stile) {
        stile = ti;
        const float4 *er = epan_reach<F64>(tb, ti) + 1;
        float4 e[kSubs];
#pragma unroll
        for (int i = 0; i < kSubs; ++i) e[i] = er[i];
        unsigned m = 0;
#pragma unroll
        for (int i = 0; i < kSubs; ++i) m |= epan_wave_far<kEpanQpt>(e[i], xu, xd, yu, yd) ? 0u : 1u << i;
        if constexpr (FAM == FAM_EPAN && !F64) {
          // the radial kernel's own test on the sub-tiles left: its support
          // is the disc inside the reach box (epan_tile_far's bound, per
          // query, with the tile's smallest scale and largest offsets)
          if (m) {
            const double4 sc = tb[kEpanTileMeta * ti + 1];
            const float4 *sb = epan_sub_boxes(tb, ti);
#pragma unroll
            for (int i = 0; i < kSubs; ++i) {
              if (!((m >> i) & 1u)) continue;
              const float4 bx = sb[i];
              bool near = false;
#pragma unroll
              for (int k = 0; k < kEpanQpt; ++k)
                near |= !epan_beyond_support<FAM_EPAN>(u_lower(bx.x, bx.y, px[k], px[k], sc.x, sc.z),
                                                       u_lower(bx.z, bx.w, py[k], py[k], sc.y, sc.w));
              if (!__any(near)) m &= ~(1u << i);
            }
          }
        }
        near_s = m;
      }
      const unsigned srest = near_s >> (int)((sub - ti * kTile) / kEpanSub);
      if (srest == 0) {
        sub = (ti + 1) * kTile;
        continue;
      }
      return sub + (long long)__builtin_ctz(srest) * kEpanSub;
    }
  };
  // this lane's reference of a sub-tile (zero features outside the range
  // give exact zero terms: a zero mass, or with FP64 features zero scales)
  const double4 zero = make_double4(0.0, 0.0, 0.0, 0.0);
  double4 n0 = zero;
  double nm = 0.0;
  auto load_ref = [&](long long sub) __attribute__((always_inline)) {
    const long long r = sub + lane;
    n0 = zero;
    nm = 0.0;
    if (r >= r_begin && r < r_end) {
      n0 = f0[r];
      if constexpr (kNf == 2) nm = f1[r].x;
    }
  };
  auto stage = [&]() __attribute__((always_inline)) {
    if constexpr (UNI) {
      if constexpr (F64)
        so[wv][lane] = make_double2(n0.x, n0.y);
      else
        so[wv][lane] = (FAM == FAM_EPAN) ? make_double2(n0.y, n0.z) : make_double2(n0.y, n0.w);
      sm[wv][lane] = (FAM == FAM_EPAN) ? n0.w : nm;
    } else {
      s0[wv][lane] = n0;
      if constexpr (kNf == 2) sm[wv][lane] = nm;
    }
  };
  // features of reference j of the staged sub-tile, as the term takes them
  // (a masked reference has zero mass; with UNI its scales are usx, usy)
  auto staged = [&](int j, double4 &a, double4 &b) __attribute__((always_inline)) {
    if constexpr (UNI) {
      const double2 o = so[wv][j];
      const double m = sm[wv][j];
      if constexpr (F64)
        a = (FAM == FAM_EPAN) ? make_double4(o.x, o.y, usx, m) : make_double4(o.x, o.y, usx, usy);
      else
        a = (FAM == FAM_EPAN) ? make_double4(usx, o.x, o.y, m) : make_double4(usx, o.x, usy, o.y);
      b = make_double4(m, 0.0, 0.0, 0.0);
    } else {
      a = s0[wv][j];
      b = make_double4(kNf == 2 ? sm[wv][j] : 0.0, 0.0, 0.0, 0.0);
    }
  };
  auto term = [&](int k, double4 a, double4 b) __attribute__((always_inline)) {
    if constexpr (F64)
      return pair_term_f64<FAM>(px[k], py[k], a, b, 0.0);
    else
      return epan_term<FAM>(px[k], py[k], a, b);
  };
  const long long q_lo = q_begin, q_hi = q_begin + q_count;
  unsigned long long evals = 0;  // count: references this wavefront loops over
  long long cur = next_needed((r_begin / kEpanSub) * kEpanSub);
  if (cur < r_end) load_ref(cur);
  while (cur < r_end) {
    wave_sync();
    stage();
    wave_sync();
    const long long nxt = next_needed(cur + kEpanSub);
    if (nxt < r_end) load_ref(nxt);
    const int cnt = (int)((r_end - cur) < kEpanSub ? (r_end - cur) : kEpanSub);
    const bool diag = self_excl && cur < q_hi && q_lo < cur + cnt;
    if (count) evals += (unsigned long long)cnt;
    if (!diag && cnt == kEpanSub) {
      // a whole sub-tile: a fixed trip count
#pragma unroll BK_EPAN_UNROLL
      for (int j = 0; j < kEpanSub; ++j) {
        double4 a, b;
        staged(j, a, b);
#pragma unroll
        for (int k = 0; k < kEpanQpt; ++k) acc[k] += term(k, a, b);
      }
    } else if (!diag) {
      for (int j = 0; j < cnt; ++j) {
        double4 a, b;
        staged(j, a, b);
#pragma unroll
        for (int k = 0; k < kEpanQpt; ++k) acc[k] += term(k, a, b);
      }
    } else {
      for (int j = 0; j < cnt; ++j) {
        double4 a, b;
        staged(j, a, b);
        const long long rj = cur + j;
#pragma unroll
        for (int k = 0; k < kEpanQpt; ++k) {
          const double v = term(k, a, b);
          acc[k] += (rj == q_begin + t + k * kBlock) ? 0.0 : v;
        }
      }
    }
    cur = nxt;
  }
#pragma unroll
  for (int k = 0; k < kEpanQpt; ++k) {
    const int qi = t + k * kBlock;
    if (qi < q_count) dst[qi] = acc[k] * out_scale;
  }
  if (count) {
    // pair terms evaluated: every lane, and the lanes holding a real query
    int real = 0;
#pragma unroll
    for (int k = 0; k < kEpanQpt; ++k) real += __popcll(__ballot(t + k * kBlock < q_count));
    if (lane == 0 && evals > 0) {
      atomicAdd(count, evals * 64ULL * (unsigned long long)kEpanQpt);
      atomicAdd(count + 1, evals * (unsigned long long)real);
    }
  }
}

// kParts consecutive workgroups take the parts of one block of
// kParts * kEpanQueriesPerBlock queries (the listed blocks of a bounded
// evaluation, and the query blocks of the FP64 path).
template <int FAM, bool BOUNDED, bool UNI, bool F64>
__global__ __launch_bounds__(kBlock) void epan_range_kernel(
    const std::conditional_t<F64, double, float> *__restrict__ qx,
    const std::conditional_t<F64, double, float> *__restrict__ qy,
    const double4 *__restrict__ f0, const double4 *__restrict__ f1, const double4 *__restrict__ tb,
    long long r_i, long long r_end, long long split_len, long long q_i,
    long long nq, double *__restrict__ dst, double out_scale, int self_excl, BoundArgs ba, double usx,
    double usy) {
  constexpr int kParts = (BOUNDED || F64) ? kEpanBoundParts : 1;
  long long bx = blockIdx.x / kParts, by = blockIdx.y;
  const int part = (int)(blockIdx.x % kParts);
  if (ba.order) {
    const int it = ba.order[bx];
    bx = it % ba.nblocks;
    by = it / ba.nblocks;
  }
  const long long qb = bx * (kParts * kEpanQueriesPerBlock) + part * kEpanQueriesPerBlock;
  if (qb >= nq) return;
  long long rb, re;
  split_bounds<kTile>(r_i, r_end, split_len, (int)by, rb, re);
  const long long left = nq - qb;
  const int q_count = (int)(left < kEpanQueriesPerBlock ? left : kEpanQueriesPerBlock);
  const long long off = by * nq + qb;
  epan_item<FAM, BOUNDED, UNI, F64>(qx, qy, f0, f1, tb, rb, re, q_i + qb, q_count, dst + off, out_scale,
                                    self_excl, BOUNDED ? ba.mask + bx * ba.nwords : nullptr, usx, usy, ba.count);
}

// Work items of up to kParts * kEpanQueriesPerBlock queries, kParts
// workgroups each.
template <int FAM, bool UNI, bool F64>
__global__ __launch_bounds__(kBlock) void epan_list_kernel(
    const std::conditional_t<F64, double, float> *__restrict__ qx,
    const std::conditional_t<F64, double, float> *__restrict__ qy,
    const double4 *__restrict__ f0, const double4 *__restrict__ f1, const double4 *__restrict__ tb,
    const WorkItem *__restrict__ items, double *__restrict__ dst,
    double out_scale, int self_excl, double usx, double usy) {
  constexpr int kParts = F64 ? kEpanBoundParts : 1;
  const WorkItem it = items[blockIdx.x / kParts];
  const int q0 = (int)(blockIdx.x % kParts) * kEpanQueriesPerBlock;
  if (q0 >= it.q_count) return;
  const int qn = it.q_count - q0 < kEpanQueriesPerBlock ? it.q_count - q0 : kEpanQueriesPerBlock;
  epan_item<FAM, false, UNI, F64>(qx, qy, f0, f1, tb, it.r_begin, it.r_end, it.q_begin + q0, qn,
                                  dst + it.out_index + q0, out_scale, self_excl, nullptr, usx, usy);
}

__global__ void prep_epan_kernel(EvalParams p, const float *__restrict__ x,
                                 const float *__restrict__ y,
                                 const float *__restrict__ m,
                                 const float *__restrict__ abw, long long n,
                                 double4 *__restrict__ f0,
                                 double4 *__restrict__ f1) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double a = (double)abw[i];
  const double mass = (double)m[i];
  const double xi = (double)x[i], yi = (double)y[i];
  const double sx = p.ix / a, sy = p.iy / a;
  switch (p.family) {
    case FAM_EPAN:
      f0[i] = make_double4(sx, -sx * xi, -sx * yi, mass);
      break;
    case FAM_EPANPROD:
      f0[i] = make_double4(sx, -sx * xi, sy, -sy * yi);
      f1[i] = make_double4(mass, 0.0, 0.0, 0.0);
      break;
    default:
      f0[i] = make_double4(sx, -sx * xi, sy, -sy * yi);
      f1[i] = make_double4(mass * p.mass_fold, 0.0, 0.0, 0.0);
      break;
  }
}

// One workgroup of kTile threads per aligned tile: the tile's box, smallest
// scales and largest offsets (see epan_tile_far), and the reach boxes of the
// tile and of each wavefront's sub-tile of kEpanSub references (see
// epan_reach_box); c = the support edge of |u| on one axis.
__global__ __launch_bounds__(kTile) void prep_epan_tiles_kernel(const float *__restrict__ x,
                                                                 const float *__restrict__ y,
                                                                 const double4 *__restrict__ f0, long long n,
                                                                 int iso, double c, double4 *__restrict__ tb) {
  static_assert(kEpanSub == 64, "one sub-tile per wavefront");
  __shared__ float sh[kTile];
  const long long i = (long long)blockIdx.x * kTile + threadIdx.x;
  const bool valid = i < n;
  const double4 a = valid ? f0[i] : make_double4(0.0, 0.0, 0.0, 0.0);
  const float big = 3.0e38f;
  const float xi = valid ? x[i] : 0.0f, yi = valid ? y[i] : 0.0f;
  // coordinates are FP32 values, so the FP32 reductions are exact
  const float x0 = block_reduce_f<kTile, false>(valid ? xi : big, sh);
  const float x1 = block_reduce_f<kTile, true>(valid ? xi : -big, sh);
  const float y0 = block_reduce_f<kTile, false>(valid ? yi : big, sh);
  const float y1 = block_reduce_f<kTile, true>(valid ? yi : -big, sh);
  // smallest scales rounded down, largest offsets rounded up to FP32, so
  // the test never overstates the first or understates the second
  auto down = [](double v) {
    float f = (float)v;
    if ((double)f > v) f = __uint_as_float(__float_as_uint(f) - 1u);  // v > 0
    return f;
  };
  auto up = [](double v) {
    float f = (float)v;
    if ((double)f < v) f = __uint_as_float(__float_as_uint(f) + 1u);  // v >= 0
    return f;
  };
  const double sxv = a.x, syv = iso ? a.x : a.z;
  const double oxv = fabs(a.y), oyv = fabs(iso ? a.z : a.w);
  const float sx = block_reduce_f<kTile, false>(valid ? down(sxv) : big, sh);
  const float sy = block_reduce_f<kTile, false>(valid ? down(syv) : big, sh);
  const float ox = block_reduce_f<kTile, true>(valid ? up(oxv) : 0.0f, sh);
  const float oy = block_reduce_f<kTile, true>(valid ? up(oyv) : 0.0f, sh);
  double4 *m = tb + kEpanTileMeta * blockIdx.x;
  float4 *reach = reinterpret_cast<float4 *>(m + 2);
  if (threadIdx.x == 0) {
    m[0] = make_double4(x0, x1, y0, y1);
    m[1] = make_double4(sx, sy, ox, oy);
    reach[0] = epan_reach_box(x0, x1, y0, y1, sx, sy, ox, oy, c);
  }
  const float w_x0 = wave_reduce_f(valid ? xi : big, false), w_x1 = wave_reduce_f(valid ? xi : -big, true);
  const float w_y0 = wave_reduce_f(valid ? yi : big, false), w_y1 = wave_reduce_f(valid ? yi : -big, true);
  const float w_sx = wave_reduce_f(valid ? down(sxv) : big, false);
  const float w_sy = wave_reduce_f(valid ? down(syv) : big, false);
  const float w_ox = wave_reduce_f(valid ? up(oxv) : 0.0f, true);
  const float w_oy = wave_reduce_f(valid ? up(oyv) : 0.0f, true);
  if ((threadIdx.x & 63) == 0) {
    reach[1 + (threadIdx.x >> 6)] = epan_reach_box(w_x0, w_x1, w_y0, w_y1, w_sx, w_sy, w_ox, w_oy, c);
    reinterpret_cast<float4 *>(m + 2 + (kEpanReachF4 + 1) / 2)[threadIdx.x >> 6] =
        make_float4(w_x0, w_x1, w_y0, w_y1);
  }
}

// ---- FP64 context path ---------------------------------------------------
constexpr int kBlock64 = 256;
constexpr int kTile64 = 256;

// 2^e in FP64 for e <= 0: e = n + f with n = rint(e) and |f| <= 1/2 (for
// every finite e), 2^f by its Chebyshev interpolant of degree 10 on
// [-1/2, 1/2] in Horner form (coefficients from exp2_f64_coef.py, each
// rounded once to FP64: 2.2e-16 relative interpolation error, 3.8e-16 with
// the FMA roundings), scaled by ldexp.  The scale is raised to 2^-2000 at
// least, which underflows to 0 as the true value does; NaN stays NaN (f is
// NaN).
__device__ __forceinline__ double exp2_f64(double e) {
  const double n = __builtin_rint(e);
  const double f = e - n;
  double p = 7.072585949269223e-09;
  p = __builtin_fma(p, f, 1.0208690299958306e-07);
  p = __builtin_fma(p, f, 1.321544258792169e-06);
  p = __builtin_fma(p, f, 1.5252657260200837e-05);
  p = __builtin_fma(p, f, 0.0001540353044173605);
  p = __builtin_fma(p, f, 0.0013333558230164974);
  p = __builtin_fma(p, f, 0.009618129107606888);
  p = __builtin_fma(p, f, 0.05550410866444772);
  p = __builtin_fma(p, f, 0.24022650695910097);
  p = __builtin_fma(p, f, 0.69314718055995);
  p = __builtin_fma(p, f, 1.0);
  return __builtin_ldexp(p, (int)__builtin_fmax(n, -2000.0));
}

// FP64 points: differences of the coordinates themselves.
//   FAM_GAUSS:    a = (x, y, gx / a_r^2, m)
//   FAM_EPAN:     a = (x, y, w, m),     w = 1/(h a)^2
//   FAM_EPANPROD: a = (x, y, wx, wy), b.x = m
//   FAM_EPANCONV: a = (x, y, 1/(hx a), 1/(hy a)), b.x = m * 2^-10
template <int FAM>
__device__ __forceinline__ double pair_term_f64(double qx, double qy,
                                                double4 a, double4 b,
                                                double rho) {
  const double dx = qx - a.x;
  const double dy = qy - a.y;
  if constexpr (FAM == FAM_GAUSS) {
    // exp2 of the base-2 exponent
    return a.w * exp2_f64(-a.z * __builtin_fma(rho * dy, dy, dx * dx));
  } else if constexpr (FAM == FAM_EPAN) {
    const double d2 = __builtin_fma(dy, dy, dx * dx);
    return a.w * clamp_unit(__builtin_fma(-a.z, d2, 1.0));
  } else if constexpr (FAM == FAM_EPANPROD) {
    const double tx = clamp_unit(__builtin_fma(-a.z * dx, dx, 1.0));
    const double ty = clamp_unit(__builtin_fma(-a.w * dy, dy, 1.0));
    return b.x * (tx * ty);
  } else {
    const double ux = __builtin_fabs(dx) * a.z;
    const double uy = __builtin_fabs(dy) * a.w;
    const double cx = __builtin_fmax(2.0 - ux, 0.0);
    const double cy = __builtin_fmax(2.0 - uy, 0.0);
    const double kx = cx * cx * cx * __builtin_fma(ux, ux + 6.0, 4.0);
    const double ky = cy * cy * cy * __builtin_fma(uy, uy + 6.0, 4.0);
    return b.x * (kx * ky);
  }
}

// The FP64 Gaussian term in a compact tile's centred form: with u = q - c
// (c the tile's centre) and the reference's r' = r - c,
//   e = log2 m - W |q - r|^2_rho = C_r + A0 ux + A1 uy - W |u|^2_rho,
// A0 = 2 W r'x, A1 = 2 W rho r'y, C_r = log2 m - W |r'|^2_rho, a = (A0, A1,
// C_r, -W): three FMAs in place of the differences, their squares and the
// mass product.  Every part is at most (sqrt|e| + 10)^2 in a compact tile
// (W |r'|^2_rho <= 100), so the FP64 rounding of e stays below ~2e-16 of
// that, far inside the 1e-12 bar for every term above underflow.
__device__ __forceinline__ double gauss_centred_f64(double ux, double uy, double u2, double4 a) {
  return exp2_f64(__builtin_fma(a.w, u2, __builtin_fma(a.x, ux, __builtin_fma(a.y, uy, a.z))));
}

// Per aligned tile of kTile64 FP64 references, two double4: the box (x0, x1,
// y0, y1) and (min a.z, min a.w, max a.z, sum |a.w|) of pair_term_f64's
// layout (FAM_GAUSS: smallest and largest exponent scale, smallest mass, sum
// of |mass|).  A tile is culled for a workgroup when its terms vanish for
// the whole box of its queries: the Epanechnikov kinds past their support
// (exactly zero), the Gaussian when every term's base-2 exponent is at least
// the workgroup's t_cull (gauss_f64_cull_exponent).  The test runs the
// term's own FP64 operations on the per-axis gaps between the boxes with the
// tile's smallest scales; each is monotone, so every Epanechnikov term of a
// culled tile is zero as computed.
template <int FAM>
__device__ __forceinline__ bool f64_tile_far(const double4 *__restrict__ tb, long long tile_idx, double bx0,
                                             double bx1, double by0, double by1, double rho, double t_cull) {
  const double4 b = tb[kF64TileMeta * tile_idx];
  const double4 w = tb[kF64TileMeta * tile_idx + 1];
  const double gx = fmax(0.0, fmax(b.x - bx1, bx0 - b.y));
  const double gy = fmax(0.0, fmax(b.z - by1, by0 - b.w));
  if constexpr (FAM == FAM_GAUSS) {
    return w.x * __builtin_fma(rho * gy, gy, gx * gx) >= t_cull;
  } else if constexpr (FAM == FAM_EPAN) {
    return __builtin_fma(-w.x, __builtin_fma(gy, gy, gx * gx), 1.0) <= 0.0;
  } else if constexpr (FAM == FAM_EPANPROD) {
    return __builtin_fma(-(w.x * gx), gx, 1.0) <= 0.0 || __builtin_fma(-(w.y * gy), gy, 1.0) <= 0.0;
  } else {
    return gx * w.x >= 2.0 || gy * w.y >= 2.0;
  }
}

// Lower bound of a query box's FP64 Gaussian sums and the mass that culling
// can drop, accumulated over the tiles of a reference range.  A tile held
// wholly in the range (and, under self exclusion, not holding one of the
// box's own queries) whose smallest mass is positive adds at least
// kTile64 m_min 2^-(W_max Q_far) to every query's sum, Q_far taken between
// the far sides of the two boxes; with no negative mass the sum of these is
// a lower bound of every query's sum.  Culled terms sum to at most 2^-T
// times the range's sum of |m| when every culled term's exponent is at least T.
struct CullBound {
  double lower = 0.0;  // sum of the per-tile lower bounds
  double mass = 0.0;   // sum of |m| over the tiles touching the range
  double min_m = 0.0;  // smallest mass seen (only its sign is used)
};

__device__ __forceinline__ void cull_bound_add(CullBound &c, const double4 *__restrict__ tb, long long ti,
                                               long long r_begin, long long r_end, bool own, double bx0, double bx1,
                                               double by0, double by1, double rho) {
  const double4 b = tb[kF64TileMeta * ti];
  const double4 w = tb[kF64TileMeta * ti + 1];
  c.mass += w.w;
  c.min_m = fmin(c.min_m, w.y);
  const long long s = ti * kTile64;
  if (s >= r_begin && s + kTile64 <= r_end && !own && w.y > 0.0) {
    const double fx = fmax(bx1 - b.x, b.y - bx0);
    const double fy = fmax(by1 - b.z, b.w - by0);
    c.lower += ((double)kTile64 * w.y) * exp2_f64(-w.z * __builtin_fma(rho * fy, fy, fx * fx));
  }
}

// Culling exponent T of a box: log2(mass / lower) + kCullRelBits rounded up
// (frexp's exponent exceeds log2 of the ratio), so the culled part of every
// query's sum is below 2^-kCullRelBits of it; kFarExp64 (culled terms
// underflow to zero) when no tile gives a lower bound or a mass is negative.
__device__ __forceinline__ double cull_exponent(const CullBound &c) {
  if (!(c.min_m >= 0.0) || !(c.lower > 0.0)) return kFarExp64;
  const double ratio = c.mass / c.lower;
  if (!(ratio < 1.0e300)) return kFarExp64;
  int e = 0;
  (void)__builtin_frexp(ratio, &e);
  return fmin((double)e + (double)kCullRelBits, kFarExp64);
}

// The culling exponent of a workgroup: its threads share the tiles of
// [r_begin, r_end) and reduce the bound over the workgroup.
__device__ double gauss_f64_cull_exponent(const double4 *__restrict__ tb, long long r_begin, long long r_end,
                                          long long q_lo, long long q_hi, int self_excl, double bx0, double bx1,
                                          double by0, double by1, double rho, double (&sred)[3][kBlock / 64]) {
  CullBound c;
  const long long t1 = (r_end + kTile64 - 1) / kTile64;
  for (long long ti = r_begin / kTile64 + threadIdx.x; ti < t1; ti += kBlock64) {
    const bool own = self_excl && ti * kTile64 < q_hi && q_lo < ti * kTile64 + kTile64;
    cull_bound_add(c, tb, ti, r_begin, r_end, own, bx0, bx1, by0, by1, rho);
  }
  double lower = c.lower, mass = c.mass;
  for (int o = 32; o > 0; o >>= 1) {
    lower += __shfl_xor(lower, o, 64);
    mass += __shfl_xor(mass, o, 64);
  }
  const double min_m = wave_reduce_d(c.min_m, false);
  const int t = threadIdx.x;
  if ((t & 63) == 0) {
    sred[0][t >> 6] = lower;
    sred[1][t >> 6] = mass;
    sred[2][t >> 6] = min_m;
  }
  __syncthreads();
  CullBound all;
  all.lower = sred[0][0];
  all.mass = sred[1][0];
  all.min_m = sred[2][0];
#pragma unroll
  for (int w = 1; w < kBlock / 64; ++w) {
    all.lower += sred[0][w];
    all.mass += sred[1][w];
    all.min_m = fmin(all.min_m, sred[2][w]);
  }
  return cull_exponent(all);
}

// kQpt queries per thread (kBlock64 * kQpt per workgroup), each reference
// read once from LDS for all of them; the self pair is excluded only on
// tiles that hold the workgroup's own queries.  Reference tiles are aligned
// to kTile64 (entries before r_begin are loaded as zero masses), and tiles out
// of reach of the workgroup's query box are skipped (tb) or, when BOUNDED,
// only the tiles listed in bmask are summed.  [c_begin, c_end) is the whole
// range whose sum [r_begin, r_end) is a part of: Gaussian culling bounds the
// dropped terms against that sum.
template <int FAM, bool BOUNDED>
__device__ __forceinline__ void direct_f64_item(
    const double *__restrict__ qx, const double *__restrict__ qy,
    const double4 *__restrict__ f0, const double4 *__restrict__ f1, const double4 *__restrict__ tb,
    long long r_begin, long long r_end, long long c_begin, long long c_end, long long q_begin, int q_count,
    double *__restrict__ dst, double out_scale, int self_excl, double rho,
    const unsigned long long *__restrict__ bmask) {
  constexpr int kNf = (FAM == FAM_EPANPROD || FAM == FAM_EPANCONV) ? 2 : 1;
  static_assert(kBlock64 == kBlock && kTile64 == kTile, "workgroup geometry of the tile summaries and masks");
  __shared__ double4 s0[kTile64];
  __shared__ double4 s1[kNf == 2 ? kTile64 : 1];
  __shared__ double sbox[4][kBlock / 64];
  __shared__ double sred[3][kBlock / 64];
  const int t = threadIdx.x;
  if (r_begin >= r_end) {
    write_zeros(dst, q_count);
    return;
  }
  double px[kQpt], py[kQpt], acc[kQpt];
#pragma unroll
  for (int k = 0; k < kQpt; ++k) {
    const int qi = t + k * kBlock64;
    const long long q = q_begin + (qi < q_count ? qi : 0);
    px[k] = qx[q];
    py[k] = qy[q];
    acc[k] = 0.0;
  }
  double bx0, bx1, by0, by1;
  block_query_box_d<kQpt>(px, py, sbox, bx0, bx1, by0, by1);
  const long long q_lo = q_begin, q_hi = q_begin + q_count;
  double t_cull = kFarExp64;
  if constexpr (FAM == FAM_GAUSS && !BOUNDED)
    t_cull = gauss_f64_cull_exponent(tb, c_begin, c_end, q_lo, q_hi, self_excl, bx0, bx1, by0, by1, rho, sred);
  auto live_from = [&](long long tl) {
    if constexpr (BOUNDED) return next_listed_tile<kTile64>(bmask, tl, r_end);
    for (; tl < r_end; tl += kTile64)
      if (!f64_tile_far<FAM>(tb, tl / kTile64, bx0, bx1, by0, by1, rho, t_cull)) break;
    return tl;
  };
  const double4 zero = make_double4(0.0, 0.0, 0.0, 0.0);
  for (long long tile = live_from((r_begin / kTile64) * kTile64); tile < r_end;
       tile = live_from(tile + kTile64)) {
    const long long r = tile + t;
    const bool in = r >= r_begin && r < r_end;
    if constexpr (FAM == FAM_GAUSS) {
      // a compact tile is summed in its centred form (f1; a masked slot
      // has C = -1e300, a zero term)
      const double4 cm = tb[kF64TileMeta * (tile / kTile64) + 2];
      if (cm.z != 0.0) {
        s0[t] = in ? f1[r] : make_double4(0.0, 0.0, -1.0e300, 0.0);
        __syncthreads();
        const int cnt = (int)((r_end - tile) < kTile64 ? (r_end - tile) : kTile64);
        const bool diag = self_excl && tile < q_hi && q_lo < tile + cnt;
        double ux[kQpt], uy[kQpt], u2[kQpt];
#pragma unroll
        for (int k = 0; k < kQpt; ++k) {
          ux[k] = px[k] - cm.x;
          uy[k] = py[k] - cm.y;
          u2[k] = __builtin_fma(rho * uy[k], uy[k], ux[k] * ux[k]);
        }
        if (!diag) {
#pragma unroll 2
          for (int j = 0; j < cnt; ++j) {
            const double4 a = s0[j];
#pragma unroll
            for (int k = 0; k < kQpt; ++k) acc[k] += gauss_centred_f64(ux[k], uy[k], u2[k], a);
          }
        } else {
          for (int j = 0; j < cnt; ++j) {
            const double4 a = s0[j];
            const long long rj = tile + j;
#pragma unroll
            for (int k = 0; k < kQpt; ++k) {
              const double v = gauss_centred_f64(ux[k], uy[k], u2[k], a);
              acc[k] += (rj == q_begin + t + k * kBlock64) ? 0.0 : v;
            }
          }
        }
        __syncthreads();
        continue;
      }
    }
    s0[t] = in ? f0[r] : zero;
    if constexpr (kNf == 2) s1[t] = in ? f1[r] : zero;
    __syncthreads();
    const int cnt = (int)((r_end - tile) < kTile64 ? (r_end - tile) : kTile64);
    const bool diag = self_excl && tile < q_hi && q_lo < tile + cnt;
    if (!diag) {
#pragma unroll 2
      for (int j = 0; j < cnt; ++j) {
        const double4 a = s0[j];
        const double4 b = (kNf == 2) ? s1[j] : a;
#pragma unroll
        for (int k = 0; k < kQpt; ++k) acc[k] += pair_term_f64<FAM>(px[k], py[k], a, b, rho);
      }
    } else {
      for (int j = 0; j < cnt; ++j) {
        const double4 a = s0[j];
        const double4 b = (kNf == 2) ? s1[j] : a;
        const long long rj = tile + j;
#pragma unroll
        for (int k = 0; k < kQpt; ++k) {
          const double v = pair_term_f64<FAM>(px[k], py[k], a, b, rho);
          acc[k] += (rj == q_begin + t + k * kBlock64) ? 0.0 : v;
        }
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int k = 0; k < kQpt; ++k) {
    const int qi = t + k * kBlock64;
    if (qi < q_count) dst[qi] = acc[k] * out_scale;
  }
}

template <int FAM, bool BOUNDED>
__global__ __launch_bounds__(kBlock64) void direct_f64_range_kernel(
    const double *__restrict__ qx, const double *__restrict__ qy,
    const double4 *__restrict__ f0, const double4 *__restrict__ f1, const double4 *__restrict__ tb,
    long long r_i, long long r_end, long long split_len, long long q_i,
    long long nq, double *__restrict__ dst, double out_scale, int self_excl,
    double rho, BoundArgs ba) {
  constexpr int kQb = kBlock64 * kQpt;
  long long bx = blockIdx.x, by = blockIdx.y;
  if (ba.order) {
    const int it = ba.order[blockIdx.x];
    bx = it % ba.nblocks;
    by = it / ba.nblocks;
  }
  const long long qb = bx * kQb;
  const long long rb = r_i + by * split_len;
  long long re = rb + split_len;
  if (re > r_end) re = r_end;
  const long long left = nq - qb;
  const int q_count = (int)(left < kQb ? left : kQb);
  const long long off = by * nq + qb;
  direct_f64_item<FAM, BOUNDED>(qx, qy, f0, f1, tb, rb, re, r_i, r_end, q_i + qb, q_count, dst + off, out_scale,
                                self_excl, rho, BOUNDED ? ba.mask + bx * ba.nwords : nullptr);
}

template <int FAM>
__global__ __launch_bounds__(kBlock64) void direct_f64_list_kernel(
    const double *__restrict__ qx, const double *__restrict__ qy,
    const double4 *__restrict__ f0, const double4 *__restrict__ f1, const double4 *__restrict__ tb,
    const WorkItem *__restrict__ items, double *__restrict__ dst,
    double out_scale, int self_excl, double rho) {
  const WorkItem it = items[blockIdx.x];
  direct_f64_item<FAM, false>(qx, qy, f0, f1, tb, it.r_begin, it.r_end, it.r_begin, it.r_end, it.q_begin,
                              it.q_count, dst + it.out_index, out_scale, self_excl, rho, nullptr);
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
      f0[i] = make_double4(x[i], y[i], p.gx * ia2, m[i]);
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

// One workgroup of kTile64 threads per aligned tile: the tile's FP64 box and
// scale and mass summary (see f64_tile_far).
//
// For the Epanechnikov families also the FP32 reach boxes of the tile and of
// its sub-tiles (see reach_box_f32): a query whose coordinate on one axis is
// at least r from every reference there has a zero term with each of them,
// as computed -- r = (1 + 2^-20) / sqrt(s_min) with s_min the smallest
// squared scale (a.z, or a.w on y for the product kernel), or
// 2 (1 + 2^-20) / s_min with s_min the smallest scale (convolution kernel):
// the FP64 products and sums that form the term lose at most a few 2^-53 of
// its parts.
// For the Gaussian: [2] = (cx, cy, compact, 0), the box centre and whether
// every reference has W |r - c|^2_rho <= 100, and for a compact tile the
// centred features of gauss_centred_f64 in f1.
__global__ __launch_bounds__(kTile64) void prep_f64_tiles_kernel(const double *__restrict__ x,
                                                                  const double *__restrict__ y,
                                                                  const double *__restrict__ m,
                                                                  const double4 *__restrict__ f0, long long n,
                                                                  int family, double rho, double4 *__restrict__ tb,
                                                                  double4 *__restrict__ f1) {
  static_assert(kTile64 == kTile && kEpanSub == 64, "one sub-tile per wavefront");
  __shared__ double sh[kTile64];
  const long long i = (long long)blockIdx.x * kTile64 + threadIdx.x;
  const bool valid = i < n;
  const double big = 1.0e308;
  const double4 a = valid ? f0[i] : make_double4(0.0, 0.0, big, big);
  const double xi = valid ? x[i] : 0.0, yi = valid ? y[i] : 0.0;
  auto reduce = [&](double v, bool is_max) {
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int o = kTile64 / 2; o > 0; o >>= 1) {
      if ((int)threadIdx.x < o) {
        const double u = sh[threadIdx.x + o], w = sh[threadIdx.x];
        sh[threadIdx.x] = is_max ? __builtin_fmax(w, u) : __builtin_fmin(w, u);
      }
      __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
  };
  const double x0 = reduce(valid ? xi : big, false);
  const double x1 = reduce(valid ? xi : -big, true);
  const double y0 = reduce(valid ? yi : big, false);
  const double y1 = reduce(valid ? yi : -big, true);
  const double wz = reduce(a.z, false);
  const double ww = reduce(a.w, false);
  const double wz_max = reduce(valid ? a.z : 0.0, true);
  // sum of |a.w| by the same tree (an upper bound of the tile's |mass|)
  sh[threadIdx.x] = valid ? __builtin_fabs(a.w) : 0.0;
  __syncthreads();
  for (int o = kTile64 / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  double4 *mt = tb + kF64TileMeta * blockIdx.x;
  if (threadIdx.x == 0) {
    mt[0] = make_double4(x0, x1, y0, y1);
    mt[1] = make_double4(wz, ww, wz_max, sh[0]);
  }
  if (family == FAM_GAUSS) {
    const double cx = 0.5 * (x0 + x1), cy = 0.5 * (y0 + y1);
    const double rx = xi - cx, ry = yi - cy;
    const double W = a.z;
    const double r2 = valid ? W * __builtin_fma(rho * ry, ry, rx * rx) : 0.0;
    const bool compact = !(reduce(r2, true) > 100.0);
    if (threadIdx.x == 0) mt[2] = make_double4(cx, cy, compact ? 1.0 : 0.0, 0.0);
    if (compact && valid) {
      const double mi = m[i];
      const double c = mi > 0.0 ? log2(mi) - W * __builtin_fma(rho * ry, ry, rx * rx) : -1.0e300;
      f1[i] = make_double4(2.0 * W * rx, 2.0 * W * rho * ry, c, -W);
    }
    return;
  }
  auto reach = [&](double s_min) {
    return family == FAM_EPANCONV ? 2.0 * (1.0 + 0x1p-20) / s_min : (1.0 + 0x1p-20) / sqrt(s_min);
  };
  float4 *rb = reinterpret_cast<float4 *>(mt + 2);
  const double sy_t = family == FAM_EPAN ? wz : ww;
  if (threadIdx.x == 0) rb[0] = reach_box_f32(x0, x1, y0, y1, reach(wz), reach(sy_t));
  const double w_x0 = wave_reduce_d(valid ? xi : big, false), w_x1 = wave_reduce_d(valid ? xi : -big, true);
  const double w_y0 = wave_reduce_d(valid ? yi : big, false), w_y1 = wave_reduce_d(valid ? yi : -big, true);
  const double w_sx = wave_reduce_d(a.z, false);
  const double w_sy = family == FAM_EPAN ? w_sx : wave_reduce_d(a.w, false);
  if ((threadIdx.x & 63) == 0)
    rb[1 + (threadIdx.x >> 6)] = reach_box_f32(w_x0, w_x1, w_y0, w_y1, reach(w_sx), reach(w_sy));
}

// ---- general dimension (3 <= D <= 16), FP64 distances -----------------------
constexpr int kBlockG = 256;
constexpr int kTileG = 128;

template <int D, int FAM>
__device__ __forceinline__ void direct_general_item(
    const float *__restrict__ qc, const float *__restrict__ rc,
    const double2 *__restrict__ wm, long long r_begin, long long r_end,
    long long q_begin, int q_count, double *__restrict__ dst,
    double out_scale, int self_excl) {
  __shared__ double sc[kTileG * D];
  __shared__ double2 sw[kTileG];
  const int t = threadIdx.x;
  const bool valid = t < q_count;
  const long long q = q_begin + (valid ? t : 0);
  double pq[D];
#pragma unroll
  for (int d = 0; d < D; ++d) pq[d] = (double)qc[q * D + d];
  double acc = 0.0;
  for (long long tile = r_begin; tile < r_end; tile += kTileG) {
    const int cnt = (int)((r_end - tile) < kTileG ? (r_end - tile) : kTileG);
    for (int e = t; e < cnt * D; e += kBlockG) sc[e] = (double)rc[tile * D + e];
    for (int e = t; e < cnt; e += kBlockG) sw[e] = wm[tile + e];
    __syncthreads();
    for (int j = 0; j < cnt; ++j) {
      double d2 = 0.0;
#pragma unroll
      for (int d = 0; d < D; ++d) {
        const double df = pq[d] - sc[j * D + d];
        d2 = __builtin_fma(df, df, d2);
      }
      const double2 w = sw[j];
      double v;
      if constexpr (FAM == FAM_GAUSS) {
        // w = (gx / a^2 in base-2 exponent units, mass)
        v = w.y * exp2_split(-w.x * d2);
      } else {
        v = w.y * clamp_unit(__builtin_fma(-w.x, d2, 1.0));
      }
      if (!(self_excl && tile + j == q)) acc += v;
    }
    __syncthreads();
  }
  if (valid) dst[t] = acc * out_scale;
}

template <int D, int FAM>
__global__ __launch_bounds__(kBlockG) void direct_general_range_kernel(
    const float *__restrict__ qc, const float *__restrict__ rc,
    const double2 *__restrict__ wm, long long r_i, long long r_end,
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
    const double2 *__restrict__ wm, const WorkItem *__restrict__ items,
    double *__restrict__ dst, double out_scale, int self_excl) {
  const WorkItem it = items[blockIdx.x];
  direct_general_item<D, FAM>(qc, rc, wm, it.r_begin, it.r_end, it.q_begin,
                              it.q_count, dst + it.out_index, out_scale,
                              self_excl);
}

__global__ void prep_general_kernel(EvalParams p, const float *__restrict__ abw,
                                    const float *__restrict__ m, long long n,
                                    double2 *__restrict__ wm) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double a = (double)abw[i];
  wm[i] = make_double2(p.gx / (a * a), (double)m[i]);
}

// ---- Gram form for 8 <= D <= 16: FP64 MFMA dot products ----------------------
// The exponent of a Gaussian term, log2 m - g |q - r|^2, is the bilinear form
//   C_r + (2 g r) . q - g |q|^2,   C_r = log2 m - g |r|^2,
// and the Epanechnikov factor 1 - g |q - r|^2 is the same form with
// C_r = 1 - g |r|^2.  A chain of v_mfma_f64_16x16x4f64 (K = padded D in steps
// of 4) computes C_r + (2 g r) . q for 16 references x 16 queries, with C_r as
// the accumulator's initial value, so the VALU is left with one FP64 FMA
// (- g |q|^2), the conversion to FP32 and the v_exp_f32 of a Gaussian term (or
// a max and an FMA for Epanechnikov).  Everything up to that point is FP64, so
// the cancellation between the three parts costs ~1e-16 of their size (an FP32
// Gram matrix leaves ~1e-5 relative error in the terms at d = 16).  Layout of
// the FP64 16x16x4 MFMA (gfx950): lane l gives A[m = l%16][k = l/16] and
// B[k = l/16][n = l%16] and receives D[l/16 + 4i][l%16], i = 0..3 (the FP32
// and BF16 16x16 forms return rows 4(l/16) + i instead).  References sit on
// M, queries on N, so a lane's four results belong to its own query.
typedef __attribute__((__vector_size__(4 * sizeof(double)))) double v4d;

constexpr int kGramGroups = 4;  // groups of 16 queries per wavefront (64 queries)

// Terms of one 16 x 16 block of (reference, query) pairs from d = C_r +
// (2 g r) . q (lane layout above); g4 holds g of the lane's four reference
// rows, m4 their masses (Epanechnikov only).  Gaussian: exp2 of the exponent
// formed in FP64 and rounded to FP32 once (|error| <= 2^-24 |e|: 2.6e-6
// relative at e = -128, below which a term is < 2^-126 m and far under any
// density floor), summed in FP32 within a tile of 128 references.
// Epanechnikov: m max(e, 0) in FP64.
template <int FAM, bool DIAG>
__device__ __forceinline__ void gram_terms(const v4d &d, const v4d &g4, const v4d &m4, int jb, double qn,
                                           long long tile, long long q, float &acc32, double &acc) {
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int j = jb + 4 * i;
    const double e = __builtin_fma(-g4[i], qn, d[i]);
    if constexpr (FAM == FAM_GAUSS) {
      float v = __builtin_amdgcn_exp2f((float)e);
      if (DIAG) v = (tile + j == q) ? 0.0f : v;
      acc32 += v;
    } else {
      double v = __builtin_fmax(e, 0.0);
      if (DIAG) v = (tile + j == q) ? 0.0 : v;
      acc = __builtin_fma(m4[i], v, acc);
    }
  }
}

// rw: per reference (C_r, g, m, |r|^2) from prep_gram_kernel.
template <int DP, int FAM>
__device__ __forceinline__ void gram_item(const float *__restrict__ qc, const float *__restrict__ rc,
                                          const double4 *__restrict__ rw, int D, long long r_begin,
                                          long long r_end, long long q_begin, int q_count,
                                          double *__restrict__ dst, double out_scale, int self_excl) {
  constexpr int KS = DP / 4;
  __shared__ double sr[kTileG][DP + 1];
  __shared__ double4 sw[kTileG];
  const int t = threadIdx.x, wave = t / 64, lane = t % 64, g16 = lane / 16, n16 = lane % 16;
  double bq[kGramGroups][KS], qn[kGramGroups], acc[kGramGroups];
  float acc32[kGramGroups];
  long long qidx[kGramGroups];
  bool qok[kGramGroups];
#pragma unroll
  for (int c = 0; c < kGramGroups; ++c) {
    const int qi = wave * 64 + c * 16 + n16;
    qok[c] = qi < q_count;
    const long long q = q_begin + (qok[c] ? qi : 0);
    qidx[c] = q;
    double nrm = 0.0;
    for (int k = 0; k < D; ++k) {
      const double v = (double)qc[q * D + k];
      nrm = __builtin_fma(v, v, nrm);
    }
    qn[c] = nrm;
#pragma unroll
    for (int s = 0; s < KS; ++s) {
      const int k = 4 * s + g16;
      bq[c][s] = k < D ? (double)qc[q * D + k] : 0.0;
    }
    acc[c] = 0.0;
    acc32[c] = 0.0f;
  }
  // padding rows: A = 0 and g = 0; C = -inf gives a zero Gaussian term, and
  // m = 0 a zero Epanechnikov term
  const double4 pad = FAM == FAM_GAUSS ? make_double4(-__builtin_inf(), 0.0, 0.0, 0.0)
                                       : make_double4(0.0, 0.0, 0.0, 0.0);
  const long long q_end = q_begin + q_count;
  for (long long tile = r_begin; tile < r_end; tile += kTileG) {
    const int cnt = (int)((r_end - tile) < kTileG ? (r_end - tile) : kTileG);
    for (int e = t; e < kTileG * DP; e += kBlockG) {
      const int j = e / DP, k = e % DP;
      // A = 2 g r, rounded once in FP64
      sr[j][k] = (j < cnt && k < D) ? 2.0 * rw[tile + j].y * (double)rc[(tile + j) * D + k] : 0.0;
    }
    for (int e = t; e < kTileG; e += kBlockG) sw[e] = e < cnt ? rw[tile + e] : pad;
    __syncthreads();
    const bool diag = self_excl && tile < q_end && q_begin < tile + kTileG;
#pragma unroll 1
    for (int sub = 0; sub < kTileG; sub += 16) {
      double a[KS];
#pragma unroll
      for (int s = 0; s < KS; ++s) a[s] = sr[sub + n16][4 * s + g16];
      // the four references of this lane's result rows, shared by the query groups
      v4d c4, g4, m4;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const double4 w = sw[sub + g16 + 4 * i];
        c4[i] = w.x;
        g4[i] = w.y;
        m4[i] = w.z;
      }
      v4d d[kGramGroups];
#pragma unroll
      for (int c = 0; c < kGramGroups; ++c) d[c] = c4;
      // independent chains side by side: consecutive MFMAs never wait on each other
#pragma unroll
      for (int s = 0; s < KS; ++s)
#pragma unroll
        for (int c = 0; c < kGramGroups; ++c) d[c] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[s], bq[c][s], d[c], 0, 0, 0);
#pragma unroll
      for (int c = 0; c < kGramGroups; ++c) {
        if (diag)
          gram_terms<FAM, true>(d[c], g4, m4, sub + g16, qn[c], tile, qidx[c], acc32[c], acc[c]);
        else
          gram_terms<FAM, false>(d[c], g4, m4, sub + g16, qn[c], tile, qidx[c], acc32[c], acc[c]);
      }
    }
    if constexpr (FAM == FAM_GAUSS) {
#pragma unroll
      for (int c = 0; c < kGramGroups; ++c) {
        acc[c] += (double)acc32[c];
        acc32[c] = 0.0f;
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int c = 0; c < kGramGroups; ++c) {
    double v = acc[c];
    v += __shfl_xor(v, 16, 64);
    v += __shfl_xor(v, 32, 64);
    if (g16 == 0 && qok[c]) dst[wave * 64 + c * 16 + n16] = v * out_scale;
  }
}

template <int DP, int FAM>
__global__ __launch_bounds__(kBlockG) void gram_range_kernel(
    const float *__restrict__ qc, const float *__restrict__ rc, const double4 *__restrict__ rw, int D,
    long long r_i, long long r_end, long long split_len, long long q_i, long long nq, double *__restrict__ dst,
    double out_scale, int self_excl) {
  const long long qb = (long long)blockIdx.x * kBlockG;
  const long long rb = r_i + (long long)blockIdx.y * split_len;
  long long re = rb + split_len;
  if (re > r_end) re = r_end;
  const long long left = nq - qb;
  const int q_count = (int)(left < kBlockG ? left : kBlockG);
  gram_item<DP, FAM>(qc, rc, rw, D, rb, re, q_i + qb, q_count, dst + (long long)blockIdx.y * nq + qb, out_scale,
                     self_excl);
}

template <int DP, int FAM>
__global__ __launch_bounds__(kBlockG) void gram_list_kernel(
    const float *__restrict__ qc, const float *__restrict__ rc, const double4 *__restrict__ rw, int D,
    const WorkItem *__restrict__ items, double *__restrict__ dst, double out_scale, int self_excl) {
  const WorkItem it = items[blockIdx.x];
  gram_item<DP, FAM>(qc, rc, rw, D, it.r_begin, it.r_end, it.q_begin, it.q_count, dst + it.out_index, out_scale,
                     self_excl);
}

// Per reference: (C_r, g, m, |r|^2) with g = gx / a^2 (in base-2 exponent
// units for the Gaussian), |r|^2 in FP64 of the FP32 coordinates and
// C_r = log2 m - g |r|^2 (Gaussian; -inf for m = 0) or 1 - g |r|^2
// (Epanechnikov).
__global__ void prep_gram_kernel(EvalParams p, const float *__restrict__ rc, const float *__restrict__ abw,
                                 const float *__restrict__ m, long long n, int D, double4 *__restrict__ rw) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double nrm = 0.0;
  for (int k = 0; k < D; ++k) {
    const double v = (double)rc[i * D + k];
    nrm = __builtin_fma(v, v, nrm);
  }
  const double a = (double)abw[i];
  const double mi = (double)m[i];
  const double g = p.gx / (a * a);
  const double c = p.family == FAM_GAUSS ? (mi > 0.0 ? __builtin_fma(-g, nrm, log2(mi)) : -__builtin_inf())
                                         : __builtin_fma(-g, nrm, 1.0);
  rw[i] = make_double4(c, g, mi, nrm);
}

// ---- FP64 rescue of underflowed Gaussian sums ---------------------------------
__global__ void zero_int_kernel(int *p) { *p = 0; }

__global__ void flag_small_kernel(const double *__restrict__ S, long long nq,
                                  double threshold, int *__restrict__ scratch) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nq) return;
  if (S[k] < threshold) {
    const int slot = atomicAdd(scratch, 1);
    scratch[1 + slot] = (int)k;
  }
}

constexpr int kRescueThreads = 256;

__device__ double block_sum_d(double v, double *sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int o = kRescueThreads / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}

__global__ __launch_bounds__(kRescueThreads) void rescue_gauss_kernel(
    EvalParams p, const float *__restrict__ x, const float *__restrict__ y,
    const float *__restrict__ m, const float *__restrict__ abw, long long n_ref,
    const float *__restrict__ qx, const float *__restrict__ qy, long long q_i,
    double *__restrict__ S, int self_excl, const int *__restrict__ scratch) {
  __shared__ double sh[kRescueThreads];
  const int count = scratch[0];
  for (int k = blockIdx.x; k < count; k += gridDim.x) {
    const long long qk = scratch[1 + k];
    const long long qa = q_i + qk;
    const double px = (double)qx[qa], py = (double)qy[qa];
    double s = 0.0;
    for (long long r = threadIdx.x; r < n_ref; r += kRescueThreads) {
      if (self_excl && r == qa) continue;
      const double dx = px - (double)x[r];
      const double dy = py - (double)y[r];
      const double a = (double)abw[r];
      s += (double)m[r] * exp2(-(p.gx * dx * dx + p.gy * dy * dy) / (a * a));
    }
    const double tot = block_sum_d(s, sh);
    if (threadIdx.x == 0) S[qk] = tot;
  }
}

// ---- reductions ---------------------------------------------------------------
__global__ void scale_kernel(double *__restrict__ v, long long n, double a) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] *= a;
}

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

__global__ __launch_bounds__(kReduceThreads) void weighted_reduce_kernel(
    int mode, const double *__restrict__ w, const double *__restrict__ u,
    const double *__restrict__ v, double a, double b, long long n,
    double *__restrict__ partial) {
  __shared__ double sh[kReduceThreads];
  double s = 0.0;
  for (long long i = (long long)blockIdx.x * kReduceThreads + threadIdx.x; i < n;
       i += (long long)gridDim.x * kReduceThreads) {
    if (mode == 0) {
      s += w[i] * log(a * u[i]);
    } else {
      s += w[i] * (a * u[i] - b * v[i]);
    }
  }
  const double tot = block_sum_d(s, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = tot;
}

__global__ __launch_bounds__(kReduceThreads) void final_reduce_kernel(
    const double *__restrict__ partial, int n, double *__restrict__ result) {
  __shared__ double sh[kReduceThreads];
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += kReduceThreads) s += partial[i];
  const double tot = block_sum_d(s, sh);
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

// Split length: a multiple of `tile`, counted from the tile boundary at or
// below r_i.

// ---- bounded evaluation: tile summaries, first pass, bounds planes ----------------
// Block-wide (kTile threads) FP64 reduction: min, or sum.
template <bool SUM>
__device__ double block_reduce_d(double v, double *sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int o = kTile / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] = SUM ? sh[threadIdx.x] + sh[threadIdx.x + o]
                                                    : fmin(sh[threadIdx.x], sh[threadIdx.x + o]);
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}

template <typename T>
__global__ __launch_bounds__(kTile) void tile_summary_kernel(const T *__restrict__ x, const T *__restrict__ y,
                                                             const T *__restrict__ m, const T *__restrict__ abw,
                                                             long long n, double4 *__restrict__ tsum) {
  __shared__ double sh[kTile];
  const long long i = (long long)blockIdx.x * kTile + threadIdx.x;
  const bool valid = i < n;
  const double inf = __builtin_inf();
  const double xi = valid ? (double)x[i] : 0.0, yi = valid ? (double)y[i] : 0.0;
  const double ai = valid ? (double)abw[i] : 0.0;
  const double x0 = block_reduce_d<false>(valid ? xi : inf, sh);
  const double x1 = -block_reduce_d<false>(valid ? -xi : inf, sh);
  const double y0 = block_reduce_d<false>(valid ? yi : inf, sh);
  const double y1 = -block_reduce_d<false>(valid ? -yi : inf, sh);
  const double a0 = block_reduce_d<false>(valid ? ai : inf, sh);
  const double a1 = -block_reduce_d<false>(valid ? -ai : inf, sh);
  const double mt = block_reduce_d<true>(valid ? (double)m[i] : 0.0, sh);
  if (threadIdx.x == 0) {
    tsum[kTileSummary * blockIdx.x] = make_double4(x0, x1, y0, y1);
    tsum[kTileSummary * blockIdx.x + 1] = make_double4(mt, a0, a1, 0.0);
  }
}

// One workgroup of kTile threads per block of qpb queries: the block's box,
// L = sum_t M_t dl, thr = max(rel L, abs_u) / m_total; then per tile the
// decision (summed: bit set in the block's mask; otherwise its mass times dl
// and times du - dl enter lo_b and gap_b), and the pruning counts.
template <int FAM, typename T>
__global__ __launch_bounds__(kTile) void bound_pass_kernel(const T *__restrict__ qx, const T *__restrict__ qy,
                                                           long long nq, int qpb, const double4 *__restrict__ tsum,
                                                           long long ntiles, double kx, double ky, double rel,
                                                           double abs_u, double m_total, long long nwords,
                                                           unsigned long long *__restrict__ mask,
                                                           double *__restrict__ lo_b, double *__restrict__ gap_b,
                                                           unsigned long long *__restrict__ counts) {
  __shared__ double sh[kTile];
  const long long q0 = (long long)blockIdx.x * qpb;
  const long long q1 = (q0 + qpb < nq) ? q0 + qpb : nq;
  const double inf = __builtin_inf();
  double x0 = inf, nx1 = inf, y0 = inf, ny1 = inf;
  for (long long q = q0 + threadIdx.x; q < q1; q += kTile) {
    const double xq = (double)qx[q], yq = (double)qy[q];
    x0 = fmin(x0, xq);
    nx1 = fmin(nx1, -xq);
    y0 = fmin(y0, yq);
    ny1 = fmin(ny1, -yq);
  }
  const double bx0 = block_reduce_d<false>(x0, sh), bx1 = -block_reduce_d<false>(nx1, sh);
  const double by0 = block_reduce_d<false>(y0, sh), by1 = -block_reduce_d<false>(ny1, sh);
  double L = 0.0;
  for (long long ti = threadIdx.x; ti < ntiles; ti += kTile) {
    double mt, du, dl;
    tile_bounds<FAM>(tsum, ti, bx0, bx1, by0, by1, kx, ky, mt, du, dl);
    L += mt * dl;
  }
  L = block_reduce_d<true>(L, sh);
  const double th = fmax(rel * L, abs_u) / m_total;
  double lo = 0.0, gp = 0.0, c_ex = 0.0, c_ap = 0.0, c_ev = 0.0;
  unsigned long long *bm = mask + (long long)blockIdx.x * nwords;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (long long base = 0; base < ntiles; base += kTile) {
    const long long ti = base + threadIdx.x;
    bool summed = false;
    if (ti < ntiles) {
      double mt, du, dl;
      tile_bounds<FAM>(tsum, ti, bx0, bx1, by0, by1, kx, ky, mt, du, dl);
      if (du - dl <= th) {
        lo += mt * dl;
        gp += mt * (du - dl);
        if (du == 0.0)
          c_ex += 1.0;
        else
          c_ap += 1.0;
      } else {
        summed = true;
        c_ev += 1.0;
      }
    }
    const unsigned long long bits = __ballot(summed);
    const long long word = (base >> 6) + wave;
    if (lane == 0 && word < nwords) bm[word] = bits;
  }
  lo = block_reduce_d<true>(lo, sh);
  gp = block_reduce_d<true>(gp, sh);
  c_ex = block_reduce_d<true>(c_ex, sh);
  c_ap = block_reduce_d<true>(c_ap, sh);
  c_ev = block_reduce_d<true>(c_ev, sh);
  if (threadIdx.x == 0) {
    lo_b[blockIdx.x] = lo;
    gap_b[blockIdx.x] = gp;
    atomicAdd(counts + 0, (unsigned long long)c_ex);
    atomicAdd(counts + 1, (unsigned long long)c_ap);
    atomicAdd(counts + 2, (unsigned long long)c_ev);
  }
}

// lower[k] = exact[k] + factor lo_b[k / qpb], upper[k] = lower[k] + factor gap_b[k / qpb].
__global__ void bounds_combine_kernel(const double *__restrict__ exact, const double *__restrict__ lo_b,
                                      const double *__restrict__ gap_b, long long n, int qpb, double factor,
                                      double *__restrict__ lower, double *__restrict__ upper) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const long long b = k / qpb;
  const double lo = exact[k] + factor * lo_b[b];
  lower[k] = lo;
  upper[k] = lo + factor * gap_b[b];
}

// ---- dispatch order of culled direct launches --------------------------------
// Box of each block's queries as doubles (exact for FP32 and FP64
// coordinates); padding slots repeat the block's first query, as the
// evaluation kernels do.
template <typename T>
__global__ __launch_bounds__(256) void block_box_kernel(const T *__restrict__ qx, const T *__restrict__ qy,
                                                        long long nq, int qpb, double4 *__restrict__ boxes) {
  __shared__ double sh[4][256];
  const long long q0 = (long long)blockIdx.x * qpb;
  const double inf = __builtin_inf();
  double x0 = inf, x1 = -inf, y0 = inf, y1 = -inf;
  for (int k = threadIdx.x; k < qpb; k += 256) {
    const long long q = q0 + k < nq ? q0 + k : q0;
    const double x = (double)qx[q], y = (double)qy[q];
    x0 = fmin(x0, x);
    x1 = fmax(x1, x);
    y0 = fmin(y0, y);
    y1 = fmax(y1, y);
  }
  sh[0][threadIdx.x] = x0;
  sh[1][threadIdx.x] = x1;
  sh[2][threadIdx.x] = y0;
  sh[3][threadIdx.x] = y1;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      sh[0][threadIdx.x] = fmin(sh[0][threadIdx.x], sh[0][threadIdx.x + o]);
      sh[1][threadIdx.x] = fmax(sh[1][threadIdx.x], sh[1][threadIdx.x + o]);
      sh[2][threadIdx.x] = fmin(sh[2][threadIdx.x], sh[2][threadIdx.x + o]);
      sh[3][threadIdx.x] = fmax(sh[3][threadIdx.x], sh[3][threadIdx.x + o]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) boxes[blockIdx.x] = make_double4(sh[0][0], sh[1][0], sh[2][0], sh[3][0]);
}

// keys[i] = ~(tiles of pair i = block + split * nblocks that the kernel's
// own culling test keeps), vals[i] = i.  KIND: 0 FP32 Gaussian (tile meta
// float4), 1 Epanechnikov on FP32 points, 2 FP64 context (splits start at
// r_i instead of the tile below it).
template <int KIND, int FAM>
__global__ void live_count_kernel(const double4 *__restrict__ boxes, const void *__restrict__ meta, long long r_i,
                                  long long r_end, long long len, int nblocks, int splits, double gx, double gy,
                                  double rho, unsigned int *__restrict__ keys, int *__restrict__ vals) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)nblocks * splits) return;
  const long long b = i % nblocks, sp = i / nblocks;
  const double4 box = boxes[b];
  constexpr int T = KIND == 0 ? kGaussTile : kTile;
  long long rb, re;
  if (KIND == 2) {
    rb = r_i + sp * len;
    re = rb + len < r_end ? rb + len : r_end;
  } else {
    const long long base = (r_i / T) * T;
    rb = base + sp * len;
    re = rb + len < r_end ? rb + len : r_end;
    if (rb < r_i) rb = r_i;
  }
  double t_cull = kFarExp64;
  if constexpr (KIND == 2 && FAM == FAM_GAUSS) {
    // the kernel's culling exponent for this box, without the exclusion of
    // the box's own tiles (an estimate for the launch order only)
    CullBound cb;
    for (long long ti = r_i / T; ti < (r_end + T - 1) / T; ++ti)
      cull_bound_add(cb, (const double4 *)meta, ti, r_i, r_end, false, box.x, box.y, box.z, box.w, rho);
    t_cull = cull_exponent(cb);
  }
  unsigned int c = 0;
  for (long long tl = (rb / T) * T; tl < re; tl += T) {
    bool far;
    if constexpr (KIND == 0)
      far = gauss_tile_far((const float4 *)meta, tl / T, (float)box.x, (float)box.y, (float)box.z, (float)box.w, gx,
                           gy);
    else if constexpr (KIND == 1)
      far = epan_tile_far<FAM>((const double4 *)meta, tl / T, (float)box.x, (float)box.y, (float)box.z,
                               (float)box.w);
    else
      far = f64_tile_far<FAM>((const double4 *)meta, tl / T, box.x, box.y, box.z, box.w, rho, t_cull);
    c += far ? 0u : 1u;
  }
  keys[i] = ~c;
  vals[i] = (int)i;
}

}  // namespace

long long split_length(long long r_i, long long r_end, int splits, int tile) {
  const long long base = (r_i / tile) * tile;
  long long len = (r_end - base + splits - 1) / splits;
  len = (len + tile - 1) / tile * tile;
  return len;
}

hipError_t launch_live_counts(const EvalParams &p, int kind, const float *qx32, const float *qy32, const double *qx64,
                              const double *qy64, long long nq, int qpb, const void *meta, long long r_i,
                              long long r_end, int splits, double4 *boxes, unsigned int *keys, int *vals,
                              hipStream_t s) {
  const long long nblocks = (nq + qpb - 1) / qpb;
  if (nblocks <= 0) return hipSuccess;
  if (kind == 2)
    block_box_kernel<double><<<(unsigned)nblocks, 256, 0, s>>>(qx64, qy64, nq, qpb, boxes);
  else
    block_box_kernel<float><<<(unsigned)nblocks, 256, 0, s>>>(qx32, qy32, nq, qpb, boxes);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  long long len;
  if (kind == 2) {
    len = (r_end - r_i + splits - 1) / splits;
    len = (len + kTile64 - 1) / kTile64 * kTile64;
  } else {
    len = split_length(r_i, r_end, splits, kind == 0 ? kGaussTile : kTile);
  }
  const long long n = nblocks * splits;
  const unsigned grid = (unsigned)((n + 255) / 256);
  const double rho = p.gy / p.gx;
#define BK_LC(K, F)                                                                                            \
  live_count_kernel<K, F><<<grid, 256, 0, s>>>(boxes, meta, r_i, r_end, len, (int)nblocks, splits, p.gx, p.gy, \
                                               rho, keys, vals)
  if (kind == 0) {
    BK_LC(0, FAM_GAUSS);
  } else {
    switch (p.family) {
      case FAM_GAUSS: BK_LC(2, FAM_GAUSS); break;
      case FAM_EPAN: if (kind == 1) BK_LC(1, FAM_EPAN); else BK_LC(2, FAM_EPAN); break;
      case FAM_EPANPROD: if (kind == 1) BK_LC(1, FAM_EPANPROD); else BK_LC(2, FAM_EPANPROD); break;
      default: if (kind == 1) BK_LC(1, FAM_EPANCONV); else BK_LC(2, FAM_EPANCONV); break;
    }
  }
#undef BK_LC
  return hipGetLastError();
}


// ---- launchers -----------------------------------------------------------------

hipError_t launch_tile_summary_f32(const float *x, const float *y, const float *m, const float *abw, size_t n,
                                   double4 *tsum, hipStream_t s) {
  if (n == 0) return hipSuccess;
  tile_summary_kernel<float><<<(unsigned)((n + kTile - 1) / kTile), kTile, 0, s>>>(x, y, m, abw, (long long)n, tsum);
  return hipGetLastError();
}

hipError_t launch_tile_summary_f64(const double *x, const double *y, const double *m, const double *abw, size_t n,
                                   double4 *tsum, hipStream_t s) {
  if (n == 0) return hipSuccess;
  tile_summary_kernel<double><<<(unsigned)((n + kTile - 1) / kTile), kTile, 0, s>>>(x, y, m, abw, (long long)n,
                                                                                     tsum);
  return hipGetLastError();
}

hipError_t launch_bound_pass(const EvalParams &p, const float *qx32, const float *qy32, const double *qx64,
                             const double *qy64, long long nq, int qpb, const double4 *tsum, long long ntiles,
                             double rel, double abs_u, double m_total, long long nwords, unsigned long long *mask,
                             double *lo_b, double *gap_b, unsigned long long *counts, hipStream_t s) {
  if (nq <= 0) return hipSuccess;
  const unsigned grid = (unsigned)((nq + qpb - 1) / qpb);
  const bool conv = p.family == FAM_EPANCONV;
  const double kx = conv ? p.ix : p.gx, ky = conv ? p.iy : p.gy;
#define BK_BP(F)                                                                                                  \
  if (qx64)                                                                                                       \
    bound_pass_kernel<F, double><<<grid, kTile, 0, s>>>(qx64, qy64, nq, qpb, tsum, ntiles, kx, ky, rel, abs_u,    \
                                                        m_total, nwords, mask, lo_b, gap_b, counts);              \
  else                                                                                                            \
    bound_pass_kernel<F, float><<<grid, kTile, 0, s>>>(qx32, qy32, nq, qpb, tsum, ntiles, kx, ky, rel, abs_u,     \
                                                       m_total, nwords, mask, lo_b, gap_b, counts)
  switch (p.family) {
    case FAM_GAUSS: BK_BP(FAM_GAUSS); break;
    case FAM_EPAN: BK_BP(FAM_EPAN); break;
    case FAM_EPANPROD: BK_BP(FAM_EPANPROD); break;
    default: BK_BP(FAM_EPANCONV); break;
  }
#undef BK_BP
  return hipGetLastError();
}

hipError_t launch_bounds_combine(const double *exact, const double *lo_b, const double *gap_b, long long n, int qpb,
                                 double factor, double *lower, double *upper, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  bounds_combine_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(exact, lo_b, gap_b, n, qpb, factor, lower, upper);
  return hipGetLastError();
}

hipError_t launch_prep_gauss(const EvalParams &p, const float *x, const float *y,
                             const float *m, const float *abw, size_t n,
                             double m_scale, float4 *fd, float4 *fg,
                             float4 *tiles, hipStream_t s) {
  if (n == 0) return hipSuccess;
  const unsigned ntiles = (unsigned)((n + kGaussTile - 1) / kGaussTile);
  prep_gauss_kernel<<<ntiles, kGaussTile, 0, s>>>(p, x, y, m, abw, (long long)n, 1.0 / m_scale, fd, fg, tiles);
  return hipGetLastError();
}

hipError_t launch_gauss(const EvalParams &p, const float *qx, const float *qy,
                        const float4 *fd, const float4 *fg, const float4 *tiles,
                        const WorkItem *items, int n_items, long long r_i,
                        long long r_end, long long q_i, long long nq,
                        int splits, double *dst, double out_scale,
                        int self_excl, int mode, hipStream_t s, const BoundArgs *ba) {
  const float rho = (float)(p.gy / p.gx);
  const double gx = p.gx, gy = p.gy;
  const bool iso = p.iso != 0;
  if (items) {
    if (n_items <= 0) return hipSuccess;
    if (iso)
      gauss_list_kernel<true><<<n_items, kBlock, 0, s>>>(qx, qy, fd, fg, tiles, items, dst, out_scale, self_excl,
                                                         rho, gx, gy, mode);
    else
      gauss_list_kernel<false><<<n_items, kBlock, 0, s>>>(qx, qy, fd, fg, tiles, items, dst, out_scale, self_excl,
                                                          rho, gx, gy, mode);
  } else {
    if (nq <= 0) return hipSuccess;
    const long long len = split_length(r_i, r_end, splits, kGaussTile);
    dim3 grid((unsigned)((nq + kGaussQueriesPerBlock - 1) / kGaussQueriesPerBlock), (unsigned)splits);
    const BoundArgs b = ba ? *ba : BoundArgs{};
    if (b.order) grid = dim3(grid.x * grid.y, 1);
#define BK_GR(I, B)                                                                                              \
  gauss_range_kernel<I, B><<<grid, kBlock, 0, s>>>(qx, qy, fd, fg, tiles, r_i, r_end, len, q_i, nq, dst, out_scale, \
                                                   self_excl, rho, gx, gy, mode, b)
    if (ba && ba->mask) {
      if (iso)
        BK_GR(true, true);
      else
        BK_GR(false, true);
    } else if (ba && ba->count) {
      if (iso)
        gauss_range_kernel<true, false, true><<<grid, kBlock, 0, s>>>(qx, qy, fd, fg, tiles, r_i, r_end, len, q_i, nq,
                                                                      dst, out_scale, self_excl, rho, gx, gy, mode, b);
      else
        gauss_range_kernel<false, false, true><<<grid, kBlock, 0, s>>>(qx, qy, fd, fg, tiles, r_i, r_end, len, q_i,
                                                                       nq, dst, out_scale, self_excl, rho, gx, gy,
                                                                       mode, b);
    } else {
      if (iso)
        BK_GR(true, false);
      else
        BK_GR(false, false);
    }
#undef BK_GR
  }
  return hipGetLastError();
}

hipError_t launch_prep_epan(const EvalParams &p, const float *x, const float *y,
                            const float *m, const float *abw, size_t n,
                            double4 *f0, double4 *f1, double4 *tb, hipStream_t s) {
  if (n == 0) return hipSuccess;
  const int th = 256;
  prep_epan_kernel<<<(unsigned)((n + th - 1) / th), th, 0, s>>>(p, x, y, m, abw, (long long)n, f0, f1);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  prep_epan_tiles_kernel<<<(unsigned)epan_meta_tiles(n), kTile, 0, s>>>(
      x, y, f0, (long long)n, p.family == FAM_EPAN ? 1 : 0, p.family == FAM_EPANCONV ? 2.0 : 1.0, tb);
  return hipGetLastError();
}

hipError_t launch_epan(const EvalParams &p, const float *qx, const float *qy,
                       const double4 *f0, const double4 *f1, const double4 *tb,
                       const WorkItem *items, int n_items, long long r_i,
                       long long r_end, long long q_i, long long nq, int splits,
                       double *dst, double out_scale, int self_excl, bool uniform_scales,
                       hipStream_t s, const BoundArgs *ba) {
  // the scales of every reference when uniform_scales: 1/h_d (a = 1)
  const double usx = p.ix, usy = p.iy;
  if (items) {
    if (n_items <= 0) return hipSuccess;
#define BK_EL(F, U) \
  epan_list_kernel<F, U, false><<<n_items, kBlock, 0, s>>>(qx, qy, f0, f1, tb, items, dst, out_scale, self_excl, usx, usy)
#define BK_EL2(F)     \
  if (uniform_scales) \
    BK_EL(F, true);   \
  else                \
    BK_EL(F, false)
    switch (p.family) {
      case FAM_EPAN: BK_EL2(FAM_EPAN); break;
      case FAM_EPANPROD: BK_EL2(FAM_EPANPROD); break;
      default: BK_EL2(FAM_EPANCONV); break;
    }
#undef BK_EL2
#undef BK_EL
  } else {
    if (nq <= 0) return hipSuccess;
    const long long len = split_length(r_i, r_end, splits, kTile);
    const BoundArgs b = ba ? *ba : BoundArgs{};
    const bool bounded = ba && ba->mask;
    // a bounded launch covers its listed blocks of kQueriesPerBlock queries
    // with kEpanBoundParts workgroups each
    dim3 grid(bounded ? (unsigned)((nq + kQueriesPerBlock - 1) / kQueriesPerBlock * kEpanBoundParts)
                      : (unsigned)((nq + kEpanQueriesPerBlock - 1) / kEpanQueriesPerBlock),
              (unsigned)splits);
    if (b.order) grid = dim3(grid.x * grid.y, 1);
#define BK_ER(F, B, U)                                                                                          \
  epan_range_kernel<F, B, U, false><<<grid, kBlock, 0, s>>>(qx, qy, f0, f1, tb, r_i, r_end, len, q_i, nq, dst, \
                                                            out_scale, self_excl, b, usx, usy)
#define BK_ER2(F)                 \
  if (bounded) {                  \
    if (uniform_scales)           \
      BK_ER(F, true, true);       \
    else                          \
      BK_ER(F, true, false);      \
  } else {                        \
    if (uniform_scales)           \
      BK_ER(F, false, true);      \
    else                          \
      BK_ER(F, false, false);     \
  }
    switch (p.family) {
      case FAM_EPAN: BK_ER2(FAM_EPAN); break;
      case FAM_EPANPROD: BK_ER2(FAM_EPANPROD); break;
      default: BK_ER2(FAM_EPANCONV); break;
    }
#undef BK_ER2
#undef BK_ER
  }
  return hipGetLastError();
}

hipError_t launch_prep_f64(const EvalParams &p, const double *x,
                           const double *y, const double *m,
                           const double *abw, size_t n, double4 *f0,
                           double4 *f1, double4 *tb, hipStream_t s) {
  if (n == 0) return hipSuccess;
  const int th = 256;
  prep_f64_kernel<<<(unsigned)((n + th - 1) / th), th, 0, s>>>(
      p, x, y, m, abw, (long long)n, f0, f1);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  prep_f64_tiles_kernel<<<(unsigned)epan_meta_tiles(n), kTile64, 0, s>>>(x, y, m, f0, (long long)n, p.family,
                                                                          p.gy / p.gx, tb, f1);
  return hipGetLastError();
}

hipError_t launch_direct_f64(const EvalParams &p, const double *qx,
                             const double *qy, const double4 *f0,
                             const double4 *f1, const double4 *tb, const WorkItem *items,
                             int n_items, long long r_i, long long r_end,
                             long long q_i, long long nq, int splits,
                             double *dst, double out_scale, int self_excl, bool uniform_scales,
                             hipStream_t s, const BoundArgs *ba) {
  if (p.family != FAM_GAUSS) {
    // scales of every reference when uniform_scales (a = 1): 1/h_d^2, or
    // 1/h_d for the convolution kernel
    const double usx = p.family == FAM_EPANCONV ? p.ix : p.gx;
    const double usy = p.family == FAM_EPANCONV ? p.iy : p.gy;
    if (items) {
      if (n_items <= 0) return hipSuccess;
      const unsigned g = (unsigned)n_items * kEpanBoundParts;
#define BK_EL(F, U) \
  epan_list_kernel<F, U, true><<<g, kBlock, 0, s>>>(qx, qy, f0, f1, tb, items, dst, out_scale, self_excl, usx, usy)
#define BK_EL2(F)     \
  if (uniform_scales) \
    BK_EL(F, true);   \
  else                \
    BK_EL(F, false)
      switch (p.family) {
        case FAM_EPAN: BK_EL2(FAM_EPAN); break;
        case FAM_EPANPROD: BK_EL2(FAM_EPANPROD); break;
        default: BK_EL2(FAM_EPANCONV); break;
      }
#undef BK_EL2
#undef BK_EL
    } else {
      if (nq <= 0) return hipSuccess;
      const long long len = split_length(r_i, r_end, splits, kTile);
      const BoundArgs b = ba ? *ba : BoundArgs{};
      const bool bounded = ba && ba->mask;
      dim3 grid((unsigned)((nq + kQueriesPerBlock - 1) / kQueriesPerBlock * kEpanBoundParts), (unsigned)splits);
      if (b.order) grid = dim3(grid.x * grid.y, 1);
#define BK_ER(F, B, U)                                                                                      \
  epan_range_kernel<F, B, U, true><<<grid, kBlock, 0, s>>>(qx, qy, f0, f1, tb, r_i, r_end, len, q_i, nq, dst, \
                                                           out_scale, self_excl, b, usx, usy)
#define BK_ER2(F)                 \
  if (bounded) {                  \
    if (uniform_scales)           \
      BK_ER(F, true, true);       \
    else                          \
      BK_ER(F, true, false);      \
  } else {                        \
    if (uniform_scales)           \
      BK_ER(F, false, true);      \
    else                          \
      BK_ER(F, false, false);     \
  }
      switch (p.family) {
        case FAM_EPAN: BK_ER2(FAM_EPAN); break;
        case FAM_EPANPROD: BK_ER2(FAM_EPANPROD); break;
        default: BK_ER2(FAM_EPANCONV); break;
      }
#undef BK_ER2
#undef BK_ER
    }
    return hipGetLastError();
  }
  const double rho = p.gy / p.gx;
  if (items) {
    if (n_items <= 0) return hipSuccess;
    direct_f64_list_kernel<FAM_GAUSS><<<n_items, kBlock64, 0, s>>>(qx, qy, f0, f1, tb, items, dst, out_scale,
                                                                   self_excl, rho);
  } else {
    if (nq <= 0) return hipSuccess;
    long long len = (r_end - r_i + splits - 1) / splits;
    len = (len + kTile64 - 1) / kTile64 * kTile64;
    dim3 grid((unsigned)((nq + kBlock64 * kQpt - 1) / (kBlock64 * kQpt)), (unsigned)splits);
    const BoundArgs b = ba ? *ba : BoundArgs{};
    if (b.order) grid = dim3(grid.x * grid.y, 1);
#define BK_R64(F, B)                                                          \
  direct_f64_range_kernel<F, B><<<grid, kBlock64, 0, s>>>(                    \
      qx, qy, f0, f1, tb, r_i, r_end, len, q_i, nq, dst, out_scale, self_excl, rho, b)
    if (ba && ba->mask)
      BK_R64(FAM_GAUSS, true);
    else
      BK_R64(FAM_GAUSS, false);
#undef BK_R64
  }
  return hipGetLastError();
}

hipError_t launch_prep_general(const EvalParams &p, const float *abw,
                               const float *m, size_t n, double2 *wm,
                               hipStream_t s) {
  if (n == 0) return hipSuccess;
  const int th = 256;
  prep_general_kernel<<<(unsigned)((n + th - 1) / th), th, 0, s>>>(p, abw, m, (long long)n, wm);
  return hipGetLastError();
}

namespace {
template <int D, int FAM>
void general_dispatch(const float *qc, const float *rc, const double2 *wm,
                      const WorkItem *items, int n_items, long long r_i,
                      long long r_end, long long q_i, long long nq, int splits,
                      double *dst, double out_scale, int self_excl,
                      hipStream_t s) {
  if (items) {
    direct_general_list_kernel<D, FAM><<<n_items, kBlockG, 0, s>>>(
        qc, rc, wm, items, dst, out_scale, self_excl);
  } else {
    long long len = (r_end - r_i + splits - 1) / splits;
    len = (len + kTileG - 1) / kTileG * kTileG;
    const dim3 grid((unsigned)((nq + kBlockG - 1) / kBlockG), (unsigned)splits);
    direct_general_range_kernel<D, FAM><<<grid, kBlockG, 0, s>>>(
        qc, rc, wm, r_i, r_end, len, q_i, nq, dst, out_scale, self_excl);
  }
}

template <int D>
void general_dispatch_fam(int fam, const float *qc, const float *rc,
                          const double2 *wm, const WorkItem *items, int n_items,
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
                                 const float *rc, const double2 *wm,
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

hipError_t launch_prep_gram(const EvalParams &p, const float *rc, const float *abw, const float *m, size_t n,
                            double4 *rw, hipStream_t s) {
  if (n == 0) return hipSuccess;
  const int th = 256;
  prep_gram_kernel<<<(unsigned)((n + th - 1) / th), th, 0, s>>>(p, rc, abw, m, (long long)n, p.dim, rw);
  return hipGetLastError();
}

namespace {
template <int DP, int FAM>
void gram_dispatch(const EvalParams &p, const float *qc, const float *rc, const double4 *rw, const WorkItem *items,
                   int n_items, long long r_i, long long r_end, long long q_i, long long nq, int splits, double *dst,
                   double out_scale, int self_excl, hipStream_t s) {
  if (items) {
    gram_list_kernel<DP, FAM><<<n_items, kBlockG, 0, s>>>(qc, rc, rw, p.dim, items, dst, out_scale, self_excl);
  } else {
    const long long len = split_length(r_i, r_end, splits, kTileG);
    const dim3 grid((unsigned)((nq + kBlockG - 1) / kBlockG), (unsigned)splits);
    gram_range_kernel<DP, FAM><<<grid, kBlockG, 0, s>>>(qc, rc, rw, p.dim, r_i, r_end, len, q_i, nq, dst, out_scale,
                                                        self_excl);
  }
}
}  // namespace

hipError_t launch_gram(const EvalParams &p, const float *qc, const float *rc, const double4 *rw,
                       const WorkItem *items, int n_items, long long r_i, long long r_end, long long q_i,
                       long long nq, int splits, double *dst, double out_scale, int self_excl, hipStream_t s) {
  if ((items && n_items <= 0) || (!items && nq <= 0)) return hipSuccess;
  if (p.dim < 8 || p.dim > 16) return hipErrorInvalidValue;
  const int dp = (p.dim + 3) / 4 * 4;
  const bool gauss = p.family == FAM_GAUSS;
#define BK_GR(DPV)                                                                                              \
  case DPV:                                                                                                     \
    if (gauss)                                                                                                  \
      gram_dispatch<DPV, FAM_GAUSS>(p, qc, rc, rw, items, n_items, r_i, r_end, q_i, nq, splits, dst, out_scale, \
                                    self_excl, s);                                                              \
    else                                                                                                        \
      gram_dispatch<DPV, FAM_EPAN>(p, qc, rc, rw, items, n_items, r_i, r_end, q_i, nq, splits, dst, out_scale,  \
                                   self_excl, s);                                                               \
    break
  switch (dp) {
    BK_GR(8);
    BK_GR(12);
    BK_GR(16);
    default:
      return hipErrorInvalidValue;
  }
#undef BK_GR
  return hipGetLastError();
}

hipError_t launch_rescue_gauss(const EvalParams &p, const float *x,
                               const float *y, const float *m, const float *abw,
                               long long n_ref, const float *qx, const float *qy,
                               long long q_i, long long nq, double *S,
                               double threshold, int self_excl, int *scratch,
                               hipStream_t s) {
  if (nq <= 0) return hipSuccess;
  zero_int_kernel<<<1, 1, 0, s>>>(scratch);
  const int th = 256;
  flag_small_kernel<<<(unsigned)((nq + th - 1) / th), th, 0, s>>>(S, nq, threshold, scratch);
  rescue_gauss_kernel<<<1024, kRescueThreads, 0, s>>>(p, x, y, m, abw, n_ref, qx, qy, q_i, S, self_excl,
                                                      scratch);
  return hipGetLastError();
}

hipError_t launch_scale(double *v, long long n, double a, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const int th = 256;
  scale_kernel<<<(unsigned)((n + th - 1) / th), th, 0, s>>>(v, n, a);
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

int gauss_resident_blocks() {
  static int cached[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 1;
  if (cached[dev] == 0) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void *>(gauss_range_kernel<true, false>),
                                                     kBlock, 0) != hipSuccess ||
        per_cu <= 0)
      per_cu = 1;
    cached[dev] = per_cu;
  }
  return cached[dev];
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
