/*
 * bbrcit_hip.h -- C ABI of the MI355X (gfx950) direct kernel-density library.
 *
 * This is the drop-in boundary that replaces the reference's device backend,
 * the class template `CudaDirectKde<D,FloatT,KernelT>`
 * (include/CudaDirectKde.h:28-123 in dchao34/bbrcit_kde).  Everything here is
 * plain C: POD structs, plain pointers and sizes, no C++ or torch types and
 * no exceptions.  Every function returns an int status: BK_OK (0) on
 * success, a negative BK_ERR_* code on failure; bk_last_error() then returns
 * a message describing the failure (thread-local, valid until the next call
 * on the same thread).
 *
 * Data model.  A context owns two point sets on one GPU, the references
 * ("data") and the queries, uploaded once and addressed afterwards by index
 * ranges, exactly as CudaDirectKde addresses the points it uploaded in its
 * constructor.  Each point carries D coordinates, a mass and an adaptive
 * bandwidth correction `abw`.  A kernel evaluation computes, for each query
 * q in a closed range [q_i, q_j],
 *
 *     S_q = sum_{r = r_i .. r_j} mass_r * U_K(x_q - x_r ; h * abw_r)
 *
 * (the unnormalised sum), optionally multiplied by the kernel's normalisation
 * C(h).  U_K and C are those of the reference's kernel families
 * (include/Kernels/KernelTraits.h), selected by bk_kernel_desc.
 *
 * Precision.  BK_PREC_F32 contexts round coordinates, masses and abw to FP32
 * on upload (as Point2d<float>::operator= does, include/Point2d.h:31-38),
 * evaluate every pair term in FP32 and accumulate per query with block
 * compensation into FP64.  BK_PREC_F64 contexts keep everything in FP64.
 * Results are always returned as double.
 *
 * Threading.  A context is not thread-safe; use one context per host thread.
 * All work of a context runs in order on its HIP stream (bk_set_stream).
 */
#ifndef BBRCIT_HIP_H_
#define BBRCIT_HIP_H_

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------- */
#define BK_OK                0
#define BK_ERR_INVALID_ARG  -1  /* reference: std::invalid_argument (i > j, ...) */
#define BK_ERR_OUT_OF_RANGE -2  /* reference: std::out_of_range (j >= n)         */
#define BK_ERR_HIP          -3  /* a HIP runtime call failed                      */
#define BK_ERR_STATE        -4  /* call not valid in the context's current state  */
#define BK_ERR_UNSUPPORTED  -5  /* kernel family / dimension not supported        */
#define BK_ERR_ALLOC        -6  /* device or host allocation failed               */

/* ---- enums ------------------------------------------------------------ */
/* Point sets of a context. */
#define BK_SET_REF   0
#define BK_SET_QUERY 1

/* Arithmetic precision of a context. */
#define BK_PREC_F32  0
#define BK_PREC_F64  1

/* Kernel families: one per kernel class template of the reference
 * (the headers of include/Kernels/).  `dim` of the descriptor must be 1 or 2, except for
 * BK_GAUSSIAN and BK_EPANECHNIKOV, which accept any D from 1 to 16. */
#define BK_EPANECHNIKOV                 1 /* EpanechnikovKernel<D,T>          */
#define BK_GAUSSIAN                     2 /* GaussianKernel<D,T>              */
#define BK_EPANECHNIKOV_PRODUCT_2D      3 /* EpanechnikovProductKernel2d<T>   */
#define BK_GAUSSIAN_PRODUCT_2D          4 /* GaussianProductKernel2d<T>       */
#define BK_EPANECHNIKOV_CONV_1D         5 /* EpanechnikovConvKernel1d<T>      */
#define BK_EPANECHNIKOV_PRODUCT_CONV_2D 6 /* EpanechnikovProductConvKernel2d<T> */
#define BK_GAUSSIAN_CONV_1D             7 /* GaussianConvKernel1d<T>          */
#define BK_GAUSSIAN_PRODUCT_CONV_2D     8 /* GaussianProductConvKernel2d<T>   */

/* Options of bk_set_option. */
#define BK_OPT_GAUSS_FORM      1  /* how FP32 Gaussian exponents are formed:         */
#define BK_GAUSS_FORM_AUTO     0  /*   tile-centred form on compact tiles (default)  */
#define BK_GAUSS_FORM_DIRECT   1  /*   differences q - r on every tile               */
#define BK_OPT_HIGH_DIM_FORM   2  /* distances for 8 <= D <= 16 (FP32 contexts):      */
#define BK_HIGH_DIM_AUTO       0  /*   Gram form, dot products on FP64 MFMA (default) */
#define BK_HIGH_DIM_DIRECT     1  /*   differences q - r on the FP64 VALU             */
#define BK_OPT_QUERY_ORDER     3  /* order in which bk_eval_range(_device) works     */
                                  /* through the queries (results always come back  */
                                  /* in the caller's order):                        */
#define BK_QUERY_ORDER_GIVEN   0  /*   as stored (default)                          */
#define BK_QUERY_ORDER_SPATIAL 1  /*   sorted along a Z curve on the device first   */
                                  /*   (D <= 2), so each workgroup                  */
                                  /*   holds nearby queries and skips the reference */
                                  /*   tiles beyond the kernel's reach              */
#define BK_OPT_COUNT_TERMS     4  /* 1: FP32-point Gaussian and Epanechnikov range */
                                  /* evaluations (D <= 2) count the pair terms    */
                                  /* their kernels evaluate (read with            */
                                  /* bk_term_count; the Gaussian runs an          */
                                  /* instrumented variant, slower); for           */
                                  /* measurement only.  0 (default): off          */

/* ---- POD types -------------------------------------------------------- */
typedef struct bk_ctx bk_ctx;

/* Kernel descriptor: the value of a reference kernel object.  hx is the
 * bandwidth of the isotropic and 1-D kernels; (hx, hy) are the two
 * bandwidths of the product kernels.  hy is ignored where it has no
 * meaning. */
typedef struct bk_kernel_desc {
  int    kind;  /* BK_EPANECHNIKOV ... BK_GAUSSIAN_PRODUCT_CONV_2D */
  int    dim;   /* point dimension D */
  double hx;
  double hy;
} bk_kernel_desc;

/* A closed pair of index ranges: references [r_i, r_j], queries [q_i, q_j]. */
typedef struct bk_range_pair {
  size_t r_i, r_j;
  size_t q_i, q_j;
} bk_range_pair;

/* ---- context ----------------------------------------------------------- */

/* Create a context on HIP device `device` with arithmetic precision
 * BK_PREC_F32 or BK_PREC_F64.  Replaces CudaDirectKde's default
 * construction (include/CudaDirectKde.h:42, 313-320). */
int bk_create(bk_ctx **ctx_out, int device, int precision);

/* Release every device buffer of the context and the context itself.
 * Replaces ~CudaDirectKde (include/CudaDirectKde.h:376-381). */
int bk_destroy(bk_ctx *ctx);

/* Message for the last failure on this thread ("" if none). */
const char *bk_last_error(void);

/* Make the context run its work on `stream` (a hipStream_t); NULL selects
 * the context's own stream.  The reference uses the default stream only. */
int bk_set_stream(bk_ctx *ctx, void *stream);

/* Set a context option (BK_OPT_*).  The reference has no counterpart: it
 * lets callers choose between evaluation forms of equal accuracy class. */
int bk_set_option(bk_ctx *ctx, int option, int value);

/* Upload point set `set` (BK_SET_REF or BK_SET_QUERY): n points of dimension
 * dim, coords row-major (coords[i*dim + d]), one mass and one abw per point.
 * mass and abw may be NULL (mass 1, abw 1).  The library keeps its own
 * device copy; the host buffers may be freed on return.  Replaces the
 * constructor CudaDirectKde(ref_pts, query_pts) and its helper
 * alloc_copy_host_dev_points (include/CudaDirectKde.h:57-59, 330-374). */
int bk_upload(bk_ctx *ctx, int set, int dim, size_t n, const double *coords,
              const double *mass, const double *abw);

/* Make the query set the reference set itself (no second copy).  Required
 * by the leave-one-out entry points, which exclude the pair i == i by index.
 * The reference reaches the same state by copying its data tree into a
 * query tree (include/KernelDensityImpl.h:40, 156, 228, 319). */
int bk_alias_queries_to_refs(bk_ctx *ctx);

/* Number of points in `set`.  Replaces reference_size()/query_size()
 * (include/CudaDirectKde.h:64-67). */
int bk_size(const bk_ctx *ctx, int set, size_t *n_out);

/* Update the masses and abw of the reference set in place (same n).  Used
 * after an adaptive pass; NULL leaves that array unchanged. */
int bk_update_refs(bk_ctx *ctx, const double *mass, const double *abw);

/* order_out[k] = index of the k-th of n points (row-major coords, dim per
 * point, 1 <= dim <= 16, n < 2^31) in lexicographic order: the order of
 * ExactLexicoLess (include/FloatUtils.h:72-75; -0 equals +0), ties (exact
 * duplicates) in index order.  A stable radix sort on HIP device `device`,
 * least significant coordinate first; needs no context.  Replaces the
 * std::sort that starts Kdtree construction (include/Kdtree.h:369-371). */
int bk_lexico_order(int device, int dim, size_t n, const double *coords, long long *order_out);

/* Pair terms evaluated since counting was switched on or last read
 * (BK_OPT_COUNT_TERMS): terms_out = wavefront lanes x references they loop
 * over (the vector work), query_terms_out = the same for lanes that hold a
 * real query.  (Sub-)tiles that a whole wavefront or workgroup passes over
 * are not counted: Gaussian, all its terms below 2^-56 (workgroup: 2^-100)
 * of their mass; Epanechnikov, all its terms exactly zero.  Resets both
 * counters.  No counterpart in the reference, whose kernel evaluates
 * every pair (include/CudaDirectKde.h:147-164). */
int bk_term_count(bk_ctx *ctx, unsigned long long *terms_out,
                  unsigned long long *query_terms_out);

/* The first steps of Kdtree construction on HIP device `device`
 * (include/Kdtree.h:364-390 merge_duplicates, 396-420 compute_bounding_box):
 * order_out (n values) receives the order of bk_lexico_order; *n_runs_out
 * the number of distinct points; run_start_out[u] (n rows) the position in
 * that order of the first point of the u-th run of exactly equal points
 * (ExactEqual: every coordinate ==), so run u holds positions
 * run_start_out[u] .. run_start_out[u + 1] - 1 (the last run ends at n), its
 * points in input order; box_out[2d], box_out[2d + 1]: the least and greatest
 * coordinate d, grown from DBL_MAX and DBL_MIN as the reference grows its box
 * (so the upper edge is at least DBL_MIN). */
int bk_lexico_runs(int device, int dim, size_t n, const double *coords,
                   long long *order_out, size_t *n_runs_out,
                   long long *run_start_out, double *box_out);

/* The attribute merge of runs of duplicates (merge_duplicates,
 * include/Kdtree.h:364-390) on HIP device `device`.  Attribute values come
 * packed in three groups per row (0 to 8 values each, row-major, NULL when
 * empty): n_sum merged by summation, n_min by std::min, n_max by std::max.
 * Run u holds rows run_offsets[u] .. run_offsets[u + 1] - 1 in merge order;
 * its result (sums_out / mins_out / maxs_out[u * n_group + j]) is its first
 * row merged with each next one, left to right: the host merge's exact
 * operations, so bit-identical to it. */
int bk_run_merge(int device, size_t n_runs, const long long *run_offsets,
                 int n_sum, const double *sums, int n_min, const double *mins,
                 int n_max, const double *maxs, double *sums_out,
                 double *mins_out, double *maxs_out);

/* ---- direct sums ----------------------------------------------------- */

/* out_host[k - q_i] = S_k (times C(h) if normalize != 0) for k in [q_i, q_j],
 * summing references r_i..r_j.  Errors as the reference: BK_ERR_OUT_OF_RANGE
 * if r_j >= n_ref or q_j >= n_query, BK_ERR_INVALID_ARG if r_i > r_j or
 * q_i > q_j.  Blocks until the results are on the host.  Replaces
 * CudaDirectKde::unnormalized_eval and ::eval
 * (include/CudaDirectKde.h:176-229). */
int bk_eval_range(bk_ctx *ctx, const bk_kernel_desc *desc,
                  size_t r_i, size_t r_j, size_t q_i, size_t q_j,
                  double *out_host, int normalize);

/* As bk_eval_range, but writes to device memory out_dev and returns without
 * waiting (work is queued on the context's stream).  For callers that keep
 * results on the GPU and for timing with events on that stream. */
int bk_eval_range_device(bk_ctx *ctx, const bk_kernel_desc *desc,
                         size_t r_i, size_t r_j, size_t q_i, size_t q_j,
                         double *out_dev, int normalize);

/* As bk_eval_range_device, allocating nothing: no hipMalloc, hipFree or
 * allocating library call, on any call.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid allocation).
 * The reference features of desc and the split partial sums live in one
 * 64 MiB device array per GPU that the library holds until the process
 * ends (its address is looked up in bk_create).  Limit: the features of the
 * context's whole reference set must fit in it (FP32 Gaussian kinds
 * 32.4 bytes per reference, about 2.07M references; Epanechnikov kinds
 * 64.9 bytes, about 1.03M; FP64 contexts 64.6 bytes, about 1.04M; D >= 3
 * 16 or 32 bytes); larger
 * sets fail with BK_ERR_INVALID_ARG before any work is queued.  The split
 * count is lowered until the partial sums fit beside the features, so
 * results agree with bk_eval_range_device up to the FP64 order of the split
 * combination.  Queries are always taken in their stored order
 * (BK_OPT_QUERY_ORDER is ignored).  Every call on one GPU shares that GPU's
 * array, which keeps the last call's contents: calls that overlap on one
 * GPU (other streams, other host threads, other contexts) race; order them
 * on one stream or synchronise between them. */
int bk_eval_range_device_noalloc(bk_ctx *ctx, const bk_kernel_desc *desc,
                                 size_t r_i, size_t r_j, size_t q_i, size_t q_j,
                                 double *out_dev, int normalize);

/* Batched leaf-pair evaluation: all n_pairs (reference range, query range)
 * pairs in one launch.  The results of pair p are written contiguously,
 * pair after pair: out_host[off_p + (k - q_i)] with off_p the sum of the
 * query-range lengths of the pairs before p.  Unnormalised unless
 * normalize != 0.  Replaces the per-leaf-pair synchronous calls
 * cu_kde.unnormalized_eval(...) of the dual-tree base case
 * (include/KernelDensityImpl.h:921-983, call at 939-942). */
int bk_eval_leaf_pairs(bk_ctx *ctx, const bk_kernel_desc *desc,
                       const bk_range_pair *pairs, size_t n_pairs,
                       double *out_host, int normalize);

/* Counts of (query block, reference tile) pairs of one bounded
 * evaluation: tiles passed over because the kernel is zero across the
 * pair of boxes, passed over with their bounds, and summed exactly. */
typedef struct bk_bound_stats {
  unsigned long long excluded;
  unsigned long long approximated;
  unsigned long long evaluated;
} bk_bound_stats;

/* Dual-tree evaluation with error bounds over every reference, for the
 * queries q_i..q_j: lower_host / upper_host[k - q_i] bracket the density
 * of query k (times C(h) if normalize != 0), with
 * upper - lower <= max(abs_err, rel_err * lower).  Queries are taken in
 * blocks of consecutive indices (put them in tree or Z-curve order, or set
 * BK_OPT_QUERY_ORDER) and references in aligned tiles of 256 in their
 * stored (tree) order.  A tile whose kernel bounds over the two bounding
 * boxes differ by at most max(rel_err * L, abs_err) / (total mass), L being
 * a lower bound of the block's sums, enters both bounds through its mass;
 * every other tile is summed exactly.  FP32 and FP64 contexts, D <= 2
 * (larger D: exact sums, lower == upper).  Replaces
 * KernelDensity::eval(query_tree, kernel, rel_err, abs_err)
 * (include/KernelDensityImpl.h:681-737) with its traversal and pruning rules
 * (747-983, 1043-1110): the same bounds contract, the pruning decided per
 * (query block, reference tile) on the GPU.  stats may be NULL. */
int bk_eval_bounded(bk_ctx *ctx, const bk_kernel_desc *desc, size_t q_i, size_t q_j,
                    double rel_err, double abs_err, double *lower_host,
                    double *upper_host, int normalize, bk_bound_stats *stats);

/* As bk_eval_bounded, with the bounds in device memory; returns after the
 * work is queued except for reading the pruning counts when stats is not
 * NULL. */
int bk_eval_bounded_device(bk_ctx *ctx, const bk_kernel_desc *desc, size_t q_i,
                           size_t q_j, double rel_err, double abs_err,
                           double *lower_dev, double *upper_dev, int normalize,
                           bk_bound_stats *stats);

/* ---- self evaluations and cross-validation ---------------------------- */

/* Requires bk_alias_queries_to_refs.  out_host[i] = S_i over all
 * references; with exclude_self != 0 the pair (i, i) is left out of the sum
 * (by index, so no cancellation).  Normalised if normalize != 0. */
int bk_self_eval(bk_ctx *ctx, const bk_kernel_desc *desc, int exclude_self,
                 double *out_host, int normalize);

/* Leave-one-out log-likelihood score
 *     LCV = sum_i w_i * log( C * sum_{j != i} m_j U(x_i - x_j) )
 * with weights w (host array, n_ref values), reduced in FP64 on the GPU.
 * Same value definition as KernelDensity::likelihood_cross_validate
 * (include/KernelDensityImpl.h:215-252), whose value minus self term is
 * computed here by excluding j == i.  Requires bk_alias_queries_to_refs. */
int bk_likelihood_cv(bk_ctx *ctx, const bk_kernel_desc *desc,
                     const double *weights, double *score_out);

/* Least-squares cross-validation score with the convolution kernel K* of
 * desc (the kind's convolution partner, see ConvKernelAssociator):
 *     LSCV = sum_i w_i C* sum_j m_j U*(x_i - x_j)
 *            - 2 sum_i w_i C sum_{j != i} m_j U(x_i - x_j)
 * reduced in FP64 on the GPU.  Same value definition as
 * KernelDensity::lsq_convolution_cross_validate
 * (include/KernelDensityImpl.h:142-209), including the reference's
 * Gaussian-product convolution kernel (KernelTraits.h:423-447).  desc.kind
 * must be one of BK_EPANECHNIKOV (D=1), BK_GAUSSIAN (D=1),
 * BK_EPANECHNIKOV_PRODUCT_2D, BK_GAUSSIAN_PRODUCT_2D.  Requires
 * bk_alias_queries_to_refs. */
int bk_lsq_conv_cv(bk_ctx *ctx, const bk_kernel_desc *desc,
                   const double *weights, double *score_out);

/* Adaptive-bandwidth pass (KernelDensity::adapt_density,
 * include/KernelDensityImpl.h:291-359): pilot f_i = C * S_i (all pairs,
 * self included) with the non-adaptive masses w_i and abw 1, geometric mean
 * g = exp(sum_i w_i log f_i) reduced in FP64 on the GPU, local corrections
 * lambda_i = (f_i / g)^(-alpha).  On return the context's reference set
 * holds abw_i = lambda_i and mass_i = w_i * lambda_i^(-D); lambda_out
 * (n_ref values, may be NULL) and g_out (may be NULL) receive the results.
 * alpha must lie in [0, 1] (BK_ERR_INVALID_ARG otherwise); alpha == 0
 * resets to the non-adaptive state.  Requires bk_alias_queries_to_refs. */
int bk_adapt(bk_ctx *ctx, const bk_kernel_desc *desc, double alpha,
             const double *weights, double *g_out, double *lambda_out);

/* ---- batched cross-validation grid ------------------------------------ */
/* Replaces the bandwidth loop of the reference's grid_search tool
 * (grid_search/grid_search.cc:366-402), which sets each (hx, hy) of a grid
 * in turn and runs one likelihood or least-squares CV per bandwidth
 * (KernelDensity::likelihood_cross_validate, KernelDensityImpl.h:215-252;
 * lsq_convolution_cross_validate, KernelDensityImpl.h:142-209).  Here one
 * persistent launch evaluates every dataset of a plan against every
 * bandwidth pair of a cross-product grid hx[ix] x hy[iy], and both scores
 * come out of the same leave-one-out sums:
 *     LCV_d(ix,iy)  = sum_i w_i log(C S_i)
 *     LSCV_d(ix,iy) = sum_i w_i (C* (T_i + m_i) - 2 C S_i)
 * with S_i = sum_{j != i} m_j U(x_i - x_j; (hx,hy) abw_j) and T_i the same
 * sum with the convolution kernel U* (for BK_GAUSSIAN_PRODUCT_2D the
 * reference's U* = U and C* = C/2, KernelTraits.h:423-447).  Pair terms are
 * FP32 with per-stage FP32 sums and FP64 accumulation across stages; the
 * score reductions are FP64 on the GPU; FP32 Gaussian sums below 2^-40 of
 * the dataset's largest mass are recomputed in FP64.
 *
 * Supported kinds: BK_GAUSSIAN_PRODUCT_2D (the default of this path) and
 * BK_EPANECHNIKOV_PRODUCT_2D (the kernel of the reference's grid_search).
 * A plan belongs to the context it was created on (device and error
 * reporting) and runs on a stream of its own; it is not thread-safe. */
typedef struct bk_cv_grid bk_cv_grid;

/* Options of a plan (bk_cv_grid_set_option / bk_cv_grid_get_option). */
#define BK_CVG_VARIANT    1  /* -1 (default): autotune on the first run of a
                                grid shape; 0..BK_CVG_VARIANTS-1: force one
                                register tiling of the bandwidth grid */
#define BK_CVG_GRAPH      2  /* 1 (default): replay a captured hipGraph of the
                                whole pipeline; 0: launch eagerly */
#define BK_CVG_VARIANTS   3  /* number of tilings: 0 = 8x8, 1 = 4x8, 2 = 4x4
                                (Gaussian), 0 = 4x8, 1 = 4x4, 2 = 2x4
                                (Epanechnikov) bandwidth pairs per thread */
#define BK_CVG_TUNE_MAX   9  /* most candidate schedules of bk_cv_grid_tune:
                                BK_CVG_VARIANTS x at most 3 unrolls */
#define BK_CVG_COUNT_UNITS 3 /* 0 (default); 1: count the (query, reference,
                                bandwidth pair) units the pair kernel
                                evaluates (runs launch eagerly while set;
                                read with bk_cv_grid_unit_count) */
/* Read-only options describing the last run's schedule: */
#define BK_CVG_TUNE_SOURCE 4 /* 0: forced by BK_CVG_VARIANT; 1: the default
                                schedule (no cache entry for the run); 2:
                                taken from the autotuner cache */
#define BK_CVG_UNROLL      5 /* unroll of the pair kernel's reference loop */
#define BK_CVG_WORKGROUPS  6 /* workgroups the persistent pair kernel ran */

/* Create a plan over n_sets datasets of 2-D points stored back to back:
 * set d holds sizes[d] >= 1 points; coords is row-major (x, y) over all
 * points, weights the per-point weights w_i (used as given; the reference
 * normalises them to sum 1 per dataset, KernelDensityImpl.h:444-455) and
 * abw the adaptive bandwidth corrections (NULL = all 1).  Masses are
 * m_i = w_i * abw_i^-2 (KernelDensityImpl.h:430-442).  Points, masses and
 * abw are rounded to FP32 once, here. */
int bk_cv_grid_create(bk_ctx *ctx, int kind, size_t n_sets,
                      const size_t *sizes, const double *coords,
                      const double *weights, const double *abw,
                      bk_cv_grid **plan_out);
int bk_cv_grid_destroy(bk_cv_grid *plan);

/* Evaluate the grid: lcv_out and lscv_out (either may be NULL) receive
 * n_sets * n_hx * n_hy scores, index (d * n_hx + ix) * n_hy + iy.
 * 1 <= n_hx, n_hy <= 64; bandwidths must lie in [1e-15, 1e15]. */
int bk_cv_grid_run(bk_cv_grid *plan, size_t n_hx, const double *hx,
                   size_t n_hy, const double *hy, double *lcv_out,
                   double *lscv_out);

/* Normalised leave-one-out densities C S_i of dataset `set` at bandwidth
 * (hx[ix], hy[iy]) of the last run (sizes[set] values). */
int bk_cv_grid_loo(bk_cv_grid *plan, size_t set, size_t ix, size_t iy,
                   double *out);

int bk_cv_grid_set_option(bk_cv_grid *plan, int option, int value);
/* BK_CVG_VARIANT reports the tiling the last run used. */
int bk_cv_grid_get_option(const bk_cv_grid *plan, int option, int *value);

/* Timing of the last run: main_ms = the persistent pair kernel alone,
 * total_ms = the whole device pipeline (both from HIP events on the plan's
 * stream); rescued = number of (point, bandwidth tile) entries recomputed
 * in FP64.  Any pointer may be NULL. */
int bk_cv_grid_last_timing(const bk_cv_grid *plan, double *main_ms,
                           double *total_ms, long long *rescued);

/* Units of the last run evaluated by the pair kernel, when it ran with
 * BK_CVG_COUNT_UNITS set: per wavefront (64 queries of a work item), the
 * units of its lanes holding a query over the 64-reference chunks it did not
 * pass over.  A wavefront passes over an Epanechnikov chunk beyond the
 * support of every bandwidth of its sub-grid for all its queries (its terms
 * exactly zero), a Gaussian chunk when every product of its x and y factors
 * is below 2^-152 for all its queries (each would add nothing to its FP32
 * stage sum).  Self pairs are counted as evaluated.  No
 * counterpart in the reference, whose cross-validation evaluates every pair
 * a tree traversal does not prune. */
int bk_cv_grid_unit_count(const bk_cv_grid *plan, unsigned long long *evaluated);

/* ---- autotuner cache ----------------------------------------------------- */
/* With BK_CVG_VARIANT at -1 (the default) a plan runs the schedule of its
 * pair kernel (register tiling x unroll of the reference loop; all sum in
 * the same order, so results do not depend on the choice) that the
 * autotuner cache holds for the run's (device, kind, bandwidth model, grid
 * shape, dataset sizes), else a default.  bk_cv_grid_tune fills the cache:
 * it runs the plan's whole grid once per candidate schedule (eagerly, the
 * pair kernel timed with HIP events), writes each candidate's pair-kernel
 * milliseconds to ms_out (BK_CVG_VARIANTS x unroll candidates, tiling-major;
 * the count goes to n_candidates_out; either may be NULL; at most
 * BK_CVG_TUNE_MAX) and
 * records the fastest.  It fails when BK_CVG_VARIANT forces a tiling.
 * bk_tune_set_cache names a file
 * that the cache is loaded from now and saved to after every new choice
 * (path NULL: memory only, and the entries in memory are dropped); without
 * a call the file named by the environment variable BBRCIT_HIP_TUNE_CACHE,
 * if any, is used.  The file is text with a versioned header; a file of
 * another version is ignored and replaced on the next save.  Saves go
 * through a temporary file renamed over the cache.  No counterpart in the
 * reference, which has no autotuner. */
int bk_tune_set_cache(const char *path);
/* Number of entries in the in-memory cache. */
int bk_tune_entry_count(size_t *count_out);
int bk_cv_grid_tune(bk_cv_grid *plan, size_t n_hx, const double *hx,
                    size_t n_hy, const double *hy, double *ms_out,
                    int *n_candidates_out);

/* ---- sampling ------------------------------------------------------------ */
/* The device part of KernelDensity::simulate (include/KernelDensityImpl.h:
 * 99-138), for batches of samples.  A sampler holds, on HIP device `device`,
 * the n points of a density (row-major coords, dim 1 or 2, FP64), their
 * cumulative weights (non-decreasing; the reference's cum_weights_, last
 * entry 1) and their bandwidth factors abw.  Sample k, with the caller's
 * uniform draw u[k] in [0, 1) and the kernel's unit draws z[k*dim + d] (the
 * standard-normal or Epanechnikov draws the reference's kernel::simulate
 * multiplies by a * h), is
 *     i = the first index with cum_weights[i] > u[k] (n - 1 if none),
 *     out[k*dim + d] = coords[i*dim + d] + (a_i * h[d]) * z[k*dim + d],
 * the product rounded in the kernel's float type (kernel_float = 1: FP32,
 * as the reference's float kernels compute it; 0: FP64), the sum in FP64:
 * bit-identical to the reference's simulate for the same draws.  index_out
 * (n_samples values) may be NULL.  No counterpart exists for the batch in the
 * reference, which samples one point per call on the host. */
typedef struct bk_sampler bk_sampler;
int bk_sampler_create(int device, int dim, size_t n, const double *coords,
                      const double *cum_weights, const double *abw,
                      bk_sampler **out);
int bk_sampler_destroy(bk_sampler *sampler);
int bk_sampler_draw(bk_sampler *sampler, int kernel_float, const double *h,
                    size_t n_samples, const double *u, const double *z,
                    double *out, long long *index_out);

/* ---- learned density surrogate ------------------------------------------ */
/* No counterpart in the reference.  An MLP 2 -> 128 -> 128 -> 128 -> 1 with
 * SiLU activations maps a 2-D query q to y, and the surrogate density is
 *     f(q) = max(exp(mu + sigma y(q)) - eps, 0)
 * with inputs mapped onto [-1, 1]^2 from the box [lo, hi]^2.  It is trained on
 * (query, density) pairs produced by the exact direct sum (bk_eval_range),
 * with the loss mean (y - t)^2, t = (ln(f + eps) - mu) / sigma.  Weights are
 * BF16 (FP32 master copies for Adam), biases FP32, accumulation FP32.
 * Parameter order (floats): W1[128][2], b1[128], W2[128][128], b2[128],
 * W3[128][128], b3[128], W4[128], b4 (W[out][in]).  The exact direct sum is
 * both the fallback and the arbiter of the surrogate (bk_mlp_serve_device). */
#define BK_MLP_NUM_PARAMS 33537

typedef struct bk_mlp bk_mlp;

/* Statistics of one bk_mlp_serve_device call. */
typedef struct bk_serve_stats {
  size_t n;              /* queries served                                  */
  size_t n_samples;      /* queries checked against the exact sum           */
  size_t n_fallback;     /* queries re-evaluated exactly (the unsampled     
                            queries of failing cells)                       */
  size_t cells;          /* cells of the check grid                         */
  size_t cells_failed;   /* cells whose sampled error exceeded rel_tol or
                            that held no sample                             */
  double max_sample_err; /* largest sampled error of the surrogate          */
  double ms_infer, ms_check, ms_fallback, ms_total;  /* HIP-event times     */
} bk_serve_stats;

/* Create a surrogate on the context's device with Glorot-uniform weights
 * drawn from `seed` and zero biases. */
int bk_mlp_create(bk_ctx *ctx, unsigned long long seed, bk_mlp **out);
int bk_mlp_destroy(bk_mlp *mlp);

/* Load / save the FP32 master parameters (BK_MLP_NUM_PARAMS floats) and the
 * input box and output normalisation.  Any output pointer may be NULL. */
int bk_mlp_set_params(bk_mlp *mlp, const float *theta, double lo, double hi,
                      double mu, double sigma, double eps);
int bk_mlp_get_params(const bk_mlp *mlp, float *theta, double *lo,
                      double *hi, double *mu, double *sigma, double *eps);

/* Checkpoint and resume: bk_mlp_save writes a versioned binary file (magic
 * "BKMLPCKP", format version 1, the network shape, the Adam step count, the
 * input box and output normalisation, the FP32 master parameters, their
 * BF16-rounded working copy, the Adam moments m and v, and an FNV-1a 64
 * checksum), through a temporary file renamed over `path`.  bk_mlp_load
 * restores all of it, so bk_mlp_train continues exactly as it would have
 * without the save; it fails (leaving the surrogate unchanged) on another
 * magic, version or shape, a checksum mismatch, or a working copy that does
 * not follow from the master weights.  The training set is not part of a
 * checkpoint.  No counterpart in the reference, which has no learned part. */
int bk_mlp_save(const bk_mlp *mlp, const char *path);
int bk_mlp_load(bk_mlp *mlp, const char *path);

/* Training set: n queries q (row-major x, y) with exact densities f >= 0;
 * sets lo, hi, eps and mu, sigma (mean and deviation of ln(f + eps)). */
int bk_mlp_set_training_data(bk_mlp *mlp, size_t n, const double *q,
                             const double *f, double lo, double hi,
                             double eps);

/* `steps` Adam steps (beta 0.9 / 0.999) on batches of `batch` samples drawn
 * from the training set by a seeded hash, learning rate cosine-decayed from
 * lr_max to lr_min.  mean_loss_out: mean loss of the last 16 batches;
 * ms_out: device time.  Either may be NULL. */
int bk_mlp_train(bk_mlp *mlp, size_t steps, size_t batch, double lr_max,
                 double lr_min, unsigned long long seed,
                 double *mean_loss_out, double *ms_out);

/* Loss and its gradient with respect to the parameters (straight through
 * the BF16 rounding) for n samples x (row-major, already on the network's
 * [-1, 1] input scale) with targets t: the training kernels' own forward and
 * backward pass.  n must be a multiple of 256 (at most 65536). */
int bk_mlp_loss_grad(bk_mlp *mlp, size_t n, const double *x,
                     const double *t, double *loss_out, float *grad_out);

/* Network outputs y for n inputs x on the [-1, 1] scale (serving kernel). */
int bk_mlp_forward(bk_mlp *mlp, size_t n, const double *x, double *y_out);

/* Surrogate densities for n device-resident FP32 queries into device
 * memory out (FP64). */
int bk_mlp_infer_device(bk_mlp *mlp, size_t n, const float *qx,
                        const float *qy, double *out);

/* Serve n device-resident queries: surrogate densities for all, then a
 * deterministic sample of about sample_fraction of them is evaluated
 * exactly by the direct sum of `ctx` (its FP32 2-D reference set, kernel
 * desc) and replaces the surrogate there; the error of a sample is
 * |s - e| / max(e, floor_density).  The box [lo, hi]^2 is cut into
 * cells_per_axis^2 cells; every query of a cell whose largest sampled error
 * exceeds rel_tol, or which holds no sample, is evaluated exactly. */
int bk_mlp_serve_device(bk_mlp *mlp, bk_ctx *ctx, const bk_kernel_desc *desc,
                        size_t n, const float *qx, const float *qy,
                        double *out, double sample_fraction, double rel_tol,
                        double floor_density, int cells_per_axis,
                        unsigned long long seed, bk_serve_stats *stats);

#ifdef __cplusplus
}
#endif

#endif /* BBRCIT_HIP_H_ */
