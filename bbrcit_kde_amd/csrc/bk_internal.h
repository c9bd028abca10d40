// Internal declarations shared by the HIP kernels (bk_kernels.hip) and the
// C-ABI implementation (bk_capi.hip).  Not part of the public interface.
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace bk {

// Device evaluation families.  Every reference kernel class maps onto one
// of them (see params_of() in bk_capi.hip):
//   FAM_GAUSS     m * exp(-c * sum_d (dx_d / (h_d a))^2)       Gaussian kinds
//   FAM_EPAN      m * max(1 - sum_d (dx_d / (h a))^2, 0)       EpanechnikovKernel<D>
//   FAM_EPANPROD  m * prod_d max(1 - (dx_d / (h_d a))^2, 0)    EpanechnikovProductKernel2d
//   FAM_EPANCONV  m * prod_d k(dx_d / (h_d a)),                 Epanechnikov convolution kinds
//                 k(u) = (2-|u|)^3 (u^2 + 6|u| + 4) / 32 for |u| < 2
enum Family : int { FAM_GAUSS = 0, FAM_EPAN = 1, FAM_EPANPROD = 2, FAM_EPANCONV = 3 };

// max(v, 0) of an Epanechnikov factor v = 1 - u^2 formed by an FMA (so
// v <= 1, and v is never -0): written as a clamp to [0, 1], which equals
// max(v, 0) bit for bit there and which the compiler folds into the FMA's
// clamp output modifier (no separate v_max_f64).
__device__ __forceinline__ double clamp_unit(double v) { return __builtin_fmin(__builtin_fmax(v, 0.0), 1.0); }

// Host-derived constants of one kernel descriptor.
struct EvalParams {
  int family;
  int dim;          // point dimension
  int iso;          // 1 if the quadratic form weights both axes equally
  double gx, gy;    // FAM_GAUSS: c*log2(e)/h_d^2; FAM_EPAN/EPANPROD: 1/h_d^2
  double ix, iy;    // FAM_EPAN, FAM_EPANPROD, FAM_EPANCONV: 1/h_d
  double c;         // FAM_GAUSS: exponent constant (0.5 or 0.25)
  double mass_fold; // constant folded into every mass (FAM_EPANCONV: 2^-5 per axis)
  double norm;      // normalisation C(h), in double
};

// One unit of work of the direct kernels: the references [r_begin, r_end)
// against the queries [q_begin, q_begin + q_count), written to
// dst[out_index + k] for the k-th query.
struct WorkItem {
  long long r_begin, r_end;
  long long q_begin;
  long long out_index;
  int q_count;
  int pad;
};

// Geometry of the direct kernels (D <= 2).
constexpr int kBlock = 256;        // threads per workgroup (4 wavefronts)
#ifndef BK_QPT
#define BK_QPT 4
#endif
#ifndef BK_GAUSS_QPT
#define BK_GAUSS_QPT 1
#endif
#ifndef BK_GAUSS_UNROLL
#define BK_GAUSS_UNROLL 2
#endif
#ifndef BK_EPAN_QPT
#define BK_EPAN_QPT 1
#endif
#ifndef BK_EPAN_UNROLL
#define BK_EPAN_UNROLL 4
#endif
constexpr int kQpt = BK_QPT;       // queries held in registers per thread (FP64-context kernels)
// queries per thread of the Epanechnikov kernels on FP32 points: one, so a
// wavefront's 64 queries pass over every sub-tile none of them reaches
constexpr int kEpanQpt = BK_EPAN_QPT;
constexpr int kEpanQueriesPerBlock = kBlock * kEpanQpt;
constexpr int kGaussQpt = BK_GAUSS_QPT;  // queries per thread of the FP32 Gaussian kernel
constexpr int kQueriesPerBlock = kBlock * kQpt;
// A bounded evaluation lists tiles per block of kQueriesPerBlock queries on
// the Epanechnikov path too; that many workgroups share one block's list.
constexpr int kEpanBoundParts = kQueriesPerBlock / kEpanQueriesPerBlock;
static_assert(kEpanBoundParts * kEpanQueriesPerBlock == kQueriesPerBlock, "whole workgroups per listed block");
constexpr int kGaussQueriesPerBlock = kBlock * kGaussQpt;
constexpr int kTile = 256;         // references per LDS stage (tiles are aligned to kTile)
#ifndef BK_GAUSS_TILE
#define BK_GAUSS_TILE 256
#endif
#ifndef BK_SPLIT_TARGET
#define BK_SPLIT_TARGET 512
#endif
constexpr int kGaussTile = BK_GAUSS_TILE;  // references per LDS stage of the FP32 Gaussian kernel
constexpr int kSplitTarget = BK_SPLIT_TARGET;  // workgroups per CU aimed at by reference splitting
static_assert(kGaussTile <= kBlock, "a Gaussian LDS stage is loaded one reference per thread");

// Gaussian evaluation form selection.
enum GaussMode : int {
  GAUSS_AUTO = 0,    // tile-centred Gram form on compact tiles, direct form elsewhere
  GAUSS_DIRECT = 1,  // direct form everywhere
};

// Bounded (dual-tree) evaluation, D <= 2 -------------------------------------
// Per aligned tile of kTile references, kTileSummary double4: the box
// (x0, x1, y0, y1) and (mass sum, smallest abw, largest abw, 0).  For a block
// of queries and a tile, du / dl bound the kernel U between any query of the
// block and any reference of the tile (U at the boxes' nearest per-axis
// distances with the largest abw, at the farthest with the smallest).  A
// block's tiles with du - dl <= thr are not summed: they add mass * dl to the
// lower bound and mass * (du - dl) to the gap between the bounds.  With
// thr = max(rel * L, abs) / (total mass), L = sum over tiles of mass * dl (a
// lower bound of every query's sum), the gap stays <= max(rel * lower, abs).
constexpr int kTileSummary = 2;
struct BoundArgs {
  const unsigned long long *mask;  // per query block, nwords words: bit t set = tile t is summed
  long long nwords;
  // optional dispatch order of the (query block, reference split) pairs:
  // workgroup k takes pair order[k] = block + split * nblocks (from
  // listed_order: most listed tiles first); nullptr = the 2-D grid's order
  const int *order = nullptr;
  int nblocks = 0;
  // FP32 Gaussian and Epanechnikov range launches (not BOUNDED): when set,
  // the pair terms their wavefronts evaluate are added to count[0] (every
  // lane) and count[1] (lanes holding a real query); the Gaussian runs an
  // instrumented variant
  unsigned long long *count = nullptr;
};
// Reference split length of the range kernels (a multiple of `tile`).
long long split_length(long long r_i, long long r_end, int splits, int tile);
// For a culled direct launch over nq queries (blocks of qpb) and references
// [r_i, r_end) in `splits` splits: keys[i] = ~(tiles of pair i = block +
// split * nblocks that survive the kernel's own culling test), vals[i] = i.
// kind: 0 FP32 Gaussian (meta = tile meta), 1 Epanechnikov on FP32 points,
// 2 FP64 context (meta = tb); boxes receives one double4 per block.
hipError_t launch_live_counts(const EvalParams &p, int kind, const float *qx32, const float *qy32, const double *qx64,
                              const double *qy64, long long nq, int qpb, const void *meta, long long r_i,
                              long long r_end, int splits, double4 *boxes, unsigned int *keys, int *vals,
                              hipStream_t s);

hipError_t launch_tile_summary_f32(const float *x, const float *y, const float *m, const float *abw, size_t n,
                                   double4 *tsum, hipStream_t s);
hipError_t launch_tile_summary_f64(const double *x, const double *y, const double *m, const double *abw, size_t n,
                                   double4 *tsum, hipStream_t s);
// First pass, one workgroup per block of qpb queries: the block's tile mask
// (nwords = ceil(ntiles / 64) words per block), lo_b[b] / gap_b[b] (mass units)
// over the tiles not summed, and counts[0..2] += tiles passed over with
// du == 0, passed over with du > 0, summed.  Queries are FP32 (qx32, qy32) or
// FP64 (qx64, qy64; the other pair null).
hipError_t launch_bound_pass(const EvalParams &p, const float *qx32, const float *qy32, const double *qx64,
                             const double *qy64, long long nq, int qpb, const double4 *tsum, long long ntiles,
                             double rel, double abs_u, double m_total, long long nwords, unsigned long long *mask,
                             double *lo_b, double *gap_b, unsigned long long *counts, hipStream_t s);
// lower[k] = exact[k] + factor * lo_b[k / qpb], upper[k] = lower[k] + factor * gap_b[k / qpb].
hipError_t launch_bounds_combine(const double *exact, const double *lo_b, const double *gap_b, long long n, int qpb,
                                 double factor, double *lower, double *upper, hipStream_t s);

// FP32 Gaussian path, D <= 2 ---------------------------------------------
// Per reference: fd = direct-form features, fg = centred-form features; per
// aligned tile of kGaussTile references: tiles holds kGaussTileMeta float4
// per tile: [0] (centre x, centre y, form, W); [1..4] per sub-tile of 64
// (centre x, centre y, reach of the sub-tile, reach of the tile) with reach
// the W Q2 beyond which a centred (sub-)tile's terms are all below 2^-56 of
// their masses; [5] (half width x, half width y, smallest W, 0) of its
// bounding box about the centre; [6..9] per sub-tile (half width x, half
// width y, 0, 0) of its bounding box about its centre.
constexpr int kGaussTileMeta = 10;
hipError_t launch_prep_gauss(const EvalParams &p, const float *x, const float *y,
                             const float *m, const float *abw, size_t n,
                             double m_scale, float4 *fd, float4 *fg,
                             float4 *tiles, hipStream_t s);

// items == nullptr: range mode, grid (query blocks) x (reference splits),
// dst holds `splits` consecutive arrays of nq partial sums.
// items != nullptr: one workgroup per work item.
// ba != nullptr (range mode only, r_i aligned to kTile): only the tiles
// listed in each block's mask are summed.
hipError_t launch_gauss(const EvalParams &p, const float *qx, const float *qy,
                        const float4 *fd, const float4 *fg, const float4 *tiles,
                        const WorkItem *items, int n_items, long long r_i,
                        long long r_end, long long q_i, long long nq,
                        int splits, double *dst, double out_scale,
                        int self_excl, int mode, hipStream_t s, const BoundArgs *ba = nullptr);

// FP64 Epanechnikov-family path, D <= 2, on FP32 point data ----------------
// tb: kEpanTileMeta double4 per aligned tile of kTile references: [0] the
// FP32 coordinate box (x0, x1, y0, y1), [1] the smallest scales and largest
// |offsets| (sx, sy, ox, oy) -- both for the culling of a query box -- then,
// from [2] on, 1 + kTile / kEpanSub float4: the FP32 reach boxes
// (L_x, H_x, L_y, H_y) of the tile and of its sub-tiles, for the
// wavefront-level tests; then the FP32 coordinate boxes (x0, x1, y0, y1) of
// the sub-tiles, for the radial kernel's own test.
// The tiles carry metadata up to a multiple of kEpanScanBatch (tiles past the
// points hold empty boxes, out of reach of every query).
constexpr int kEpanSub = 64;  // references per wavefront stage
constexpr int kEpanReachF4 = 1 + kTile / kEpanSub;  // reach boxes per tile
constexpr int kEpanTileMeta = 2 + (kEpanReachF4 + 1) / 2 + kTile / kEpanSub / 2;
constexpr int kEpanScanBatch = 4;  // tile-level reach tests whose loads are issued together
// FP64 path: the same layout up to the reach boxes ([1] holds that path's
// scale summary; the reach boxes are written for the Epanechnikov families).
constexpr int kF64TileMeta = 2 + (kEpanReachF4 + 1) / 2;
inline long long epan_meta_tiles(long long n) {
  return (n + (long long)kTile * kEpanScanBatch - 1) / ((long long)kTile * kEpanScanBatch) * kEpanScanBatch;
}
hipError_t launch_prep_epan(const EvalParams &p, const float *x, const float *y,
                            const float *m, const float *abw, size_t n,
                            double4 *f0, double4 *f1, double4 *tb, hipStream_t s);
hipError_t launch_epan(const EvalParams &p, const float *qx, const float *qy,
                       const double4 *f0, const double4 *f1, const double4 *tb,
                       const WorkItem *items, int n_items, long long r_i,
                       long long r_end, long long q_i, long long nq, int splits,
                       double *dst, double out_scale, int self_excl, bool uniform_scales,
                       hipStream_t s, const BoundArgs *ba = nullptr);

// FP64 context path, D <= 2 ------------------------------------------------
hipError_t launch_prep_f64(const EvalParams &p, const double *x,
                           const double *y, const double *m,
                           const double *abw, size_t n, double4 *f0,
                           double4 *f1, double4 *tb, hipStream_t s);
// tb: kF64TileMeta double4 per aligned tile of kTile references (up to
// epan_meta_tiles(n) tiles), written by launch_prep_f64.  The Epanechnikov
// families run the wavefront-level kernel of the FP32 path on FP64 data
// (uniform_scales: every abw is 1).
hipError_t launch_direct_f64(const EvalParams &p, const double *qx,
                             const double *qy, const double4 *f0,
                             const double4 *f1, const double4 *tb, const WorkItem *items,
                             int n_items, long long r_i, long long r_end,
                             long long q_i, long long nq, int splits,
                             double *dst, double out_scale, int self_excl, bool uniform_scales,
                             hipStream_t s, const BoundArgs *ba = nullptr);

// General dimension, 3 <= D <= 16 (FAM_GAUSS and FAM_EPAN) ---------------
// FP32 point data, FP64 distances.  wm = (w = g / abw^2, mass) per point.
hipError_t launch_prep_general(const EvalParams &p, const float *abw,
                               const float *m, size_t n, double2 *wm,
                               hipStream_t s);
// Gram form for 8 <= D <= 16 (FP64 MFMA dot products); rw per reference:
// (|r|^2, g / a^2, m, 0).
hipError_t launch_prep_gram(const EvalParams &p, const float *rc, const float *abw, const float *m, size_t n,
                            double4 *rw, hipStream_t s);
hipError_t launch_gram(const EvalParams &p, const float *qc, const float *rc, const double4 *rw,
                       const WorkItem *items, int n_items, long long r_i, long long r_end, long long q_i,
                       long long nq, int splits, double *dst, double out_scale, int self_excl, hipStream_t s);

hipError_t launch_direct_general(const EvalParams &p, const float *qc,
                                 const float *rc, const double2 *wm,
                                 const WorkItem *items, int n_items,
                                 long long r_i, long long r_end, long long q_i,
                                 long long nq, int splits, double *dst,
                                 double out_scale, int self_excl,
                                 hipStream_t s);

// Exact FP64 recomputation of FP32 Gaussian sums that fell below
// `threshold` (they may have lost accuracy to FP32 underflow).  S holds nq
// unnormalised sums of queries q_i.. (query index k -> S[k - q_i]) over the
// references [0, n_ref); entries below threshold are recomputed with the
// same masses.  scratch needs room for nq + 1 ints.
hipError_t launch_rescue_gauss(const EvalParams &p, const float *x,
                               const float *y, const float *m, const float *abw,
                               long long n_ref, const float *qx, const float *qy,
                               long long q_i, long long nq, double *S,
                               double threshold, int self_excl, int *scratch,
                               hipStream_t s);

// v[i] *= a for i < n.
hipError_t launch_scale(double *v, long long n, double a, hipStream_t s);

// Reductions -----------------------------------------------------------
// out[q] = sum_s partial[s * nq + q], fixed order.
hipError_t launch_reduce_splits(const double *partial, int splits,
                                long long nq, double *out, hipStream_t s);

// Deterministic FP64 reductions into *result (device):
//   mode 0: sum_i w_i * log(a * u_i)
//   mode 1: sum_i w_i * (a * u_i - b * v_i)
// scratch must hold kReduceBlocks doubles.
constexpr int kReduceBlocks = 256;
hipError_t launch_weighted_reduce(int mode, const double *w, const double *u,
                                  const double *v, double a, double b,
                                  long long n, double *scratch,
                                  double *result, hipStream_t s);

// lambda_i = (norm * S_i / g)^(-alpha), with log(g) read from device memory.
hipError_t launch_adapt_lambda(const double *S, long long n, double norm,
                               const double *log_g_dev, double alpha,
                               double *lambda, hipStream_t s);

// Spatial order ------------------------------------------------------------
struct DevBuf;  // bk_common.h
// Sorts list[0, n) (indices into qx, qy, each below index_limit <= 2^32;
// list == nullptr means the identity 0..n-1, written into list_out) by the
// Morton code of the point on a 1024 x 1024 grid over [lo, hi]^2 (points
// outside are clamped onto the border), ties broken by the index, so the
// order is deterministic.  The result goes to list_out (may equal list).
// keys and tmp are workspaces the function grows.
void spatial_sort(const float *qx, const float *qy, const long long *list, long long n, long long index_limit,
                  float lo, float hi, long long *list_out, DevBuf &keys_a, DevBuf &keys_b, DevBuf &tmp,
                  hipStream_t s);
void spatial_sort(const double *qx, const double *qy, const long long *list, long long n, long long index_limit,
                  float lo, float hi, long long *list_out, DevBuf &keys_a, DevBuf &keys_b, DevBuf &tmp,
                  hipStream_t s);
// order_out[k]: the (block, split) pairs, block + split * nblocks, sorted by
// the number of tiles listed in `mask` within the split, most first (ties in
// pair order), so a bounded launch starts its longest workgroups first.
void listed_order(const unsigned long long *mask, long long nwords, int nblocks, int splits, long long split_tiles,
                  int *order_out, DevBuf &keys_a, DevBuf &keys_b, DevBuf &vals, DevBuf &tmp, hipStream_t s);
// order_out = vals sorted by keys (ascending, stable): with keys ~count, the
// pairs with the most work first.
void order_by_keys(unsigned int *keys, int *vals, long long n, int *order_out, DevBuf &keys_b, DevBuf &tmp,
                   hipStream_t s);
// Stable lexicographic order (ExactLexicoLess, -0 equal to +0, ties in index
// order) of n device points (row-major, dim per point): a radix sort per
// coordinate, least significant first.  The order is left in va; ka, kb, vb
// and tmp are workspaces the function grows.
void lexico_sort(const double *coords, int dim, long long n, DevBuf &ka, DevBuf &kb, DevBuf &va, DevBuf &vb,
                 DevBuf &tmp, hipStream_t s);
// gx[k] = qx[list[k]], gy[k] = qy[list[k]] for k < n.
hipError_t launch_gather_xy(const float *qx, const float *qy, const long long *list, long long n, float *gx,
                            float *gy, hipStream_t s);
hipError_t launch_gather_xy(const double *qx, const double *qy, const long long *list, long long n, double *gx,
                            double *gy, hipStream_t s);
// out[list[k]] = vals[k] for k < n.
hipError_t launch_scatter_f64(const long long *list, long long n, const double *vals, double *out, hipStream_t s);

// Number of compute units of the current device (cached per device).
int device_cu_count();
// Workgroups of the FP32 Gaussian range kernel resident per CU.
int gauss_resident_blocks();

// Autotuner cache (bk_tune.hip).  kTuneVersion names the meaning of the
// tuned parameters; a cache file of another version is ignored.
constexpr int kTuneVersion = 1;
// "<architecture>/<compute units>cu" of the current device
std::string device_tune_tag();
// the configuration recorded for key, if any
bool tune_lookup(const std::string &key, std::vector<int> &value);
// record a configuration (and save the cache file when one is named)
void tune_record(const std::string &key, const std::vector<int> &value);

}  // namespace bk
