/*
 * kde_oracle.c -- FP64 CPU restatement of the reference's direct
 * kernel-density sums.  TEST INFRASTRUCTURE ONLY: the checker that the HIP
 * library is compared against, and the CPU baseline timed beside it.
 * Nothing in the product (bbrcit_kde_amd/, include/) may call it.
 *
 * Every function restates dchao34/bbrcit_kde behaviour in double precision;
 * the reference file:line each one follows is cited above it.  Inputs are
 * taken as given (callers pass the FP32-rounded points when checking the
 * FP32 GPU path, so both sides see the same numbers).
 *
 * Kinds use the numbering of include/bbrcit_hip.h.
 */
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define BKO_EPANECHNIKOV 1
#define BKO_GAUSSIAN 2
#define BKO_EPANECHNIKOV_PRODUCT_2D 3
#define BKO_GAUSSIAN_PRODUCT_2D 4
#define BKO_EPANECHNIKOV_CONV_1D 5
#define BKO_EPANECHNIKOV_PRODUCT_CONV_2D 6
#define BKO_GAUSSIAN_CONV_1D 7
#define BKO_GAUSSIAN_PRODUCT_CONV_2D 8

static const double kPi = 3.14159265358979323846;

/* Normalisation C(h) of each kind.
 * EpanechnikovTraits<D,T>::normalization  KernelTraits.h:22-25
 * GaussianTraits<D,T>::normalization      KernelTraits.h:100-102
 * GaussianProduct2dTraits                 KernelTraits.h:175-177
 * EpanechnikovProduct2dTraits             KernelTraits.h:223-225
 * EpanechnikovConv1dTraits                KernelTraits.h:268-270
 * EpanechnikovProductConv2dTraits         KernelTraits.h:318-320
 * GaussianConv1dTraits                    KernelTraits.h:369-371
 * GaussianProductConv2dTraits             KernelTraits.h:413-415 */
double bko_normalization(int kind, int dim, double hx, double hy) {
  const double D = (double)dim;
  switch (kind) {
    case BKO_EPANECHNIKOV:
      return 0.5 * (D + 2.0) / (pow(kPi, D / 2.0) / tgamma(1.0 + D / 2.0)) / pow(hx, D);
    case BKO_GAUSSIAN:
      return pow(2.0 * kPi, -D / 2.0) / pow(hx, D);
    case BKO_GAUSSIAN_PRODUCT_2D:
      return 1.0 / (2.0 * kPi * hx * hy);
    case BKO_EPANECHNIKOV_PRODUCT_2D:
      return 0.5625 / (hx * hy);
    case BKO_EPANECHNIKOV_CONV_1D:
      return 0.6 / hx;
    case BKO_EPANECHNIKOV_PRODUCT_CONV_2D:
      return 0.36 / (hx * hy);
    case BKO_GAUSSIAN_CONV_1D:
      return 1.0 / (sqrt(4.0 * kPi) * hx);
    case BKO_GAUSSIAN_PRODUCT_CONV_2D:
      return 1.0 / (4.0 * kPi * hx * hy);
    default:
      return NAN;
  }
}

/* (2-|u|)^3 (u^2 + 6|u| + 4) / 32 for |u| < 2, else 0.
 * EpanechnikovConv1dTraits::kernel, KernelTraits.h:273-277 / 303-307. */
static double epan_conv_1d(double u) {
  const double a = fabs(u);
  const double c = a < 2.0 ? 2.0 - a : 0.0;
  return 0.03125 * c * c * c * (a * a + 6.0 * a + 4.0);
}

/* Unnormalised kernel U(x - y; h*a), with dx = x - y.
 * GaussianKernel::unnormalized_eval + point_arg_eval   GaussianKernel.h:118-142
 * EpanechnikovKernel::unnormalized_eval + point_arg_eval EpanechnikovKernel.h:122-146
 * GaussianProductKernel2d                               GaussianProductKernel2d.h:132-150
 * EpanechnikovProductKernel2d                           EpanechnikovProductKernel2d.h:309-317
 * EpanechnikovConvKernel1d                              EpanechnikovConvKernel1d.h:109-114
 * EpanechnikovProductConvKernel2d                       EpanechnikovProductConvKernel2d.h:211-217
 * GaussianConvKernel1d (exp(-0.25 arg))                 GaussianConvKernel1d.h:106-121, KernelTraits.h:373-376
 * GaussianProductConvKernel2d: the float/double specialisations the
 * library instantiates evaluate exp(-0.5 arg) (KernelTraits.h:431-433,
 * 444-446), not the exp(-0.25 arg) of the generic template
 * (KernelTraits.h:418-420); this restatement follows the specialisations. */
double bko_kernel_u(int kind, int dim, double hx, double hy, const double *dx, double a) {
  switch (kind) {
    case BKO_GAUSSIAN:
    case BKO_EPANECHNIKOV: {
      double r = 0.0;
      for (int d = 0; d < dim; ++d) r += dx[d] * dx[d];
      r /= (hx * hx * a * a);
      if (kind == BKO_GAUSSIAN) return exp(-0.5 * r);
      return 1.0 - r > 0.0 ? 1.0 - r : 0.0;
    }
    case BKO_GAUSSIAN_PRODUCT_2D:
    case BKO_GAUSSIAN_PRODUCT_CONV_2D: {
      const double r = dx[0] * dx[0] / (hx * hx * a * a) + dx[1] * dx[1] / (hy * hy * a * a);
      return exp(-0.5 * r);
    }
    case BKO_EPANECHNIKOV_PRODUCT_2D: {
      const double tx = 1.0 - dx[0] * dx[0] / (hx * hx * a * a);
      const double ty = 1.0 - dx[1] * dx[1] / (hy * hy * a * a);
      return (tx > 0.0 ? tx : 0.0) * (ty > 0.0 ? ty : 0.0);
    }
    case BKO_EPANECHNIKOV_CONV_1D:
      return epan_conv_1d(dx[0] / (hx * a));
    case BKO_EPANECHNIKOV_PRODUCT_CONV_2D:
      return epan_conv_1d(dx[0] / (hx * a)) * epan_conv_1d(dx[1] / (hy * a));
    case BKO_GAUSSIAN_CONV_1D:
      return exp(-0.25 * dx[0] * dx[0] / (hx * hx * a * a));
    default:
      return NAN;
  }
}

typedef struct {
  int kind, dim;
  double hx, hy;
  const double *ref, *mass, *abw;
  size_t nr;
  const double *qry;
  size_t q_begin, q_end;
  int exclude_self;
  const long long *self_idx; /* non-NULL: query q is ref[self_idx[q]] and pair (self_idx[q], q) is left out */
  double *out;
} direct_job;

/* Per query, the serial FP64 sum over every reference of mass * U, as
 * KernelDensity::direct_eval(p, kernel) does (KernelDensityImpl.h:1151-1173),
 * without the normalisation.  exclude_self leaves out the pair r == q
 * (the leave-one-out sums of KernelDensityImpl.h:55-59, 176-180, 244-248,
 * computed here by exclusion rather than by subtraction). */
static void *direct_worker(void *arg) {
  direct_job *j = (direct_job *)arg;
  double dx[64];
  for (size_t q = j->q_begin; q < j->q_end; ++q) {
    double s = 0.0;
    const size_t self = j->self_idx ? (size_t)j->self_idx[q] : q;
    const double *xq = j->self_idx ? j->ref + self * (size_t)j->dim : j->qry + q * (size_t)j->dim;
    for (size_t r = 0; r < j->nr; ++r) {
      if (j->exclude_self && r == self) continue;
      const double *xr = j->ref + r * (size_t)j->dim;
      for (int d = 0; d < j->dim; ++d) dx[d] = xq[d] - xr[d];
      const double a = j->abw ? j->abw[r] : 1.0;
      const double m = j->mass ? j->mass[r] : 1.0;
      s += m * bko_kernel_u(j->kind, j->dim, j->hx, j->hy, dx, a);
    }
    j->out[q] = s;
  }
  return NULL;
}

/* out[q] = sum_r mass_r U(x_q - x_r; h abw_r), q < nq (unnormalised).
 * Points are row-major (n x dim).  nthreads <= 0 uses one thread. */
static int direct_run(int kind, int dim, double hx, double hy, const double *ref, const double *mass,
                      const double *abw, size_t nr, const double *qry, size_t nq, int exclude_self,
                      const long long *self_idx, double *out, int nthreads) {
  if (dim < 1 || dim > 64) return -1;
  if (nthreads < 1) nthreads = 1;
  if ((size_t)nthreads > nq) nthreads = nq > 0 ? (int)nq : 1;
  direct_job *jobs = (direct_job *)calloc((size_t)nthreads, sizeof(direct_job));
  pthread_t *th = (pthread_t *)calloc((size_t)nthreads, sizeof(pthread_t));
  if (!jobs || !th) {
    free(jobs);
    free(th);
    return -2;
  }
  const size_t chunk = (nq + (size_t)nthreads - 1) / (size_t)nthreads;
  for (int t = 0; t < nthreads; ++t) {
    direct_job *j = &jobs[t];
    j->kind = kind;
    j->dim = dim;
    j->hx = hx;
    j->hy = hy;
    j->ref = ref;
    j->mass = mass;
    j->abw = abw;
    j->nr = nr;
    j->qry = qry;
    j->q_begin = (size_t)t * chunk < nq ? (size_t)t * chunk : nq;
    j->q_end = j->q_begin + chunk < nq ? j->q_begin + chunk : nq;
    j->exclude_self = exclude_self;
    j->self_idx = self_idx;
    j->out = out;
  }
  int *joined = (int *)calloc((size_t)nthreads, sizeof(int));
  for (int t = 1; t < nthreads; ++t) {
    if (joined && pthread_create(&th[t], NULL, direct_worker, &jobs[t]) == 0) {
      joined[t] = 1;
    } else {
      direct_worker(&jobs[t]); /* could not start a thread: run inline */
    }
  }
  direct_worker(&jobs[0]);
  for (int t = 1; t < nthreads; ++t)
    if (joined && joined[t]) pthread_join(th[t], NULL);
  free(joined);
  free(jobs);
  free(th);
  return 0;
}

int bko_direct(int kind, int dim, double hx, double hy, const double *ref, const double *mass,
               const double *abw, size_t nr, const double *qry, size_t nq, int exclude_self, double *out,
               int nthreads) {
  return direct_run(kind, dim, hx, hy, ref, mass, abw, nr, qry, nq, exclude_self, NULL, out, nthreads);
}

/* Leave-one-out sums at a subset of the points: out[k] = sum over r != idx[k]
 * of mass_r U(ref[idx[k]] - ref[r]; h abw_r), the entries idx[k] of
 * bko_direct(..., ref, nr, 1, ...) over the whole set (same loop, same order). */
int bko_direct_loo_at(int kind, int dim, double hx, double hy, const double *ref, const double *mass,
                      const double *abw, size_t nr, const long long *idx, size_t nq, double *out, int nthreads) {
  for (size_t k = 0; k < nq; ++k)
    if (idx[k] < 0 || (size_t)idx[k] >= nr) return -3;
  return direct_run(kind, dim, hx, hy, ref, mass, abw, nr, NULL, nq, 1, idx, out, nthreads);
}

/* KernelDensity::likelihood_cross_validate, KernelDensityImpl.h:215-252:
 *   sum_i w_i log(C * (S_i - m_i))  with S_i the all-pairs sum;
 * here S_i - m_i U(0) is the sum over j != i. */
double bko_likelihood_cv(int kind, int dim, double hx, double hy, const double *pts, const double *weight,
                         const double *mass, const double *abw, size_t n, int nthreads) {
  double *s = (double *)malloc(n * sizeof(double));
  if (!s) return NAN;
  bko_direct(kind, dim, hx, hy, pts, mass, abw, n, pts, n, 1, s, nthreads);
  const double c = bko_normalization(kind, dim, hx, hy);
  double cv = 0.0;
  for (size_t i = 0; i < n; ++i) cv += weight[i] * log(c * s[i]);
  free(s);
  return cv;
}

static int conv_kind(int kind, int dim) {
  /* ConvKernelAssociator, Kernels/ConvKernelAssociator.h:23-84 */
  switch (kind) {
    case BKO_EPANECHNIKOV:
      return dim == 1 ? BKO_EPANECHNIKOV_CONV_1D : -1;
    case BKO_GAUSSIAN:
      return dim == 1 ? BKO_GAUSSIAN_CONV_1D : -1;
    case BKO_EPANECHNIKOV_PRODUCT_2D:
      return BKO_EPANECHNIKOV_PRODUCT_CONV_2D;
    case BKO_GAUSSIAN_PRODUCT_2D:
      return BKO_GAUSSIAN_PRODUCT_CONV_2D;
    default:
      return -1;
  }
}

/* KernelDensity::lsq_convolution_cross_validate, KernelDensityImpl.h:142-209:
 *   sum_i w_i C* S*_i - 2 sum_i w_i C (S_i - m_i)
 * with S*_i the all-pairs sum of the convolution kernel (self included). */
double bko_lsq_conv_cv(int kind, int dim, double hx, double hy, const double *pts, const double *weight,
                       const double *mass, const double *abw, size_t n, int nthreads) {
  const int ck = conv_kind(kind, dim);
  if (ck < 0) return NAN;
  double *s = (double *)malloc(n * sizeof(double));
  double *sc = (double *)malloc(n * sizeof(double));
  if (!s || !sc) {
    free(s);
    free(sc);
    return NAN;
  }
  bko_direct(kind, dim, hx, hy, pts, mass, abw, n, pts, n, 1, s, nthreads);
  bko_direct(ck, dim, hx, hy, pts, mass, abw, n, pts, n, 0, sc, nthreads);
  const double c = bko_normalization(kind, dim, hx, hy);
  const double cc = bko_normalization(ck, dim, hx, hy);
  double sq = 0.0, loo = 0.0;
  for (size_t i = 0; i < n; ++i) {
    sq += weight[i] * cc * sc[i];
    loo += weight[i] * c * s[i];
  }
  free(s);
  free(sc);
  return sq - 2.0 * loo;
}

/* KernelDensity::adapt_density, KernelDensityImpl.h:291-359, with an exact
 * (direct) pilot instead of the dual-tree pilot:
 *   f_i = C sum_j w_j U(x_i - x_j)         (non-adaptive, self included)
 *   g   = exp(sum_i w_i log f_i)
 *   lambda_i = (f_i / g)^(-alpha)
 * Writes lambda (n values) and returns g. */
double bko_adapt(int kind, int dim, double hx, double hy, const double *pts, const double *weight, size_t n,
                 double alpha, double *lambda, int nthreads) {
  double *f = (double *)malloc(n * sizeof(double));
  if (!f) return NAN;
  bko_direct(kind, dim, hx, hy, pts, weight, NULL, n, pts, n, 0, f, nthreads);
  const double c = bko_normalization(kind, dim, hx, hy);
  double lg = 0.0;
  for (size_t i = 0; i < n; ++i) {
    f[i] *= c;
    lg += weight[i] * log(f[i]);
  }
  const double g = exp(lg);
  for (size_t i = 0; i < n; ++i) lambda[i] = pow(f[i] / g, -alpha);
  free(f);
  return g;
}
