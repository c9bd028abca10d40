// ref_driver.cc -- C entry points over the reference's own CPU implementation
// (dchao34/bbrcit_kde, header-only C++).  TEST INFRASTRUCTURE ONLY.
//
// This file contains no reference code: it includes the reference headers
// from where they lie (-I /root/reference/include, see oracle/Makefile) and
// instantiates KernelDensity<D, KernelT, double> the way the reference's own
// test programs do (e.g. test/test_kde10.cc, test/test_kde20.cc), so the
// restatement in kde_oracle.c and the golden fixtures can be checked against
// the reference itself.  The build goes to oracle/_ref/ only.
//
// Kinds use the numbering of include/bbrcit_hip.h; kernel_float selects the
// kernel's float type (0: double, 1: float) as in KernelDensity<2,
// Kernel<float>, double>, the type every reference GPU caller uses.
// Standard headers first: some reference headers use <random> and <vector>
// without including them, and the reference's test programs include these
// before the library headers (e.g. test/test_kde10.cc).
#include <random>
#include <vector>

#include <KernelDensity.h>
#include <Kernels/ConvKernelAssociator.h>
#include <Kernels/EpanechnikovConvKernel1d.h>
#include <Kernels/EpanechnikovKernel.h>
#include <Kernels/EpanechnikovProductConvKernel2d.h>
#include <Kernels/EpanechnikovProductKernel2d.h>
#include <Kernels/GaussianConvKernel1d.h>
#include <Kernels/GaussianKernel.h>
#include <Kernels/GaussianProductConvKernel2d.h>
#include <Kernels/GaussianProductKernel2d.h>

#include <chrono>
#include <cstddef>
#include <thread>
#include <exception>
#include <iostream>
#include <sstream>
#include <vector>

using namespace bbrcit;

namespace {

template <int D, class K>
struct Tag {
  static constexpr int dim = D;
  using Kernel = K;
};

template <int D, class T>
void set_bw(GaussianKernel<D, T> &k, double hx, double) { k.set_bandwidth(hx); }
template <int D, class T>
void set_bw(EpanechnikovKernel<D, T> &k, double hx, double) { k.set_bandwidth(hx); }
template <class T>
void set_bw(GaussianProductKernel2d<T> &k, double hx, double hy) { k.set_bandwidths(hx, hy); }
template <class T>
void set_bw(EpanechnikovProductKernel2d<T> &k, double hx, double hy) { k.set_bandwidths(hx, hy); }
template <class T>
void set_bw(EpanechnikovProductConvKernel2d<T> &k, double hx, double hy) { k.set_bandwidths(hx, hy); }
template <class T>
void set_bw(GaussianProductConvKernel2d<T> &k, double hx, double hy) { k.set_bandwidths(hx, hy); }
template <class T>
void set_bw(EpanechnikovConvKernel1d<T> &k, double hx, double) { k.set_hx(hx); }
template <class T>
void set_bw(GaussianConvKernel1d<T> &k, double hx, double) { k.set_hx(hx); }

// Every (kind, dim) the reference supports for D <= 2, and the generic
// Gaussian and Epanechnikov kernels at D = 3, 8 and 16.
template <class T, class Fn>
int dispatch_t(int kind, int dim, Fn fn) {
  switch (kind) {
    case 1:
      if (dim == 1) return fn(Tag<1, EpanechnikovKernel<1, T>>());
      if (dim == 2) return fn(Tag<2, EpanechnikovKernel<2, T>>());
      if (dim == 3) return fn(Tag<3, EpanechnikovKernel<3, T>>());
      if (dim == 8) return fn(Tag<8, EpanechnikovKernel<8, T>>());
      if (dim == 16) return fn(Tag<16, EpanechnikovKernel<16, T>>());
      break;
    case 2:
      if (dim == 1) return fn(Tag<1, GaussianKernel<1, T>>());
      if (dim == 2) return fn(Tag<2, GaussianKernel<2, T>>());
      if (dim == 3) return fn(Tag<3, GaussianKernel<3, T>>());
      if (dim == 8) return fn(Tag<8, GaussianKernel<8, T>>());
      if (dim == 16) return fn(Tag<16, GaussianKernel<16, T>>());
      break;
    case 3:
      if (dim == 2) return fn(Tag<2, EpanechnikovProductKernel2d<T>>());
      break;
    case 4:
      if (dim == 2) return fn(Tag<2, GaussianProductKernel2d<T>>());
      break;
    case 5:
      if (dim == 1) return fn(Tag<1, EpanechnikovConvKernel1d<T>>());
      break;
    case 6:
      if (dim == 2) return fn(Tag<2, EpanechnikovProductConvKernel2d<T>>());
      break;
    case 7:
      if (dim == 1) return fn(Tag<1, GaussianConvKernel1d<T>>());
      break;
    case 8:
      if (dim == 2) return fn(Tag<2, GaussianProductConvKernel2d<T>>());
      break;
  }
  return -1;
}

template <class Fn>
int dispatch(int kind, int dim, int kernel_float, Fn fn) {
  return kernel_float ? dispatch_t<float>(kind, dim, fn) : dispatch_t<double>(kind, dim, fn);
}

// Kinds that have a convolution kernel (ConvKernelAssociator specialisations).
template <class T, class Fn>
int dispatch_conv_t(int kind, int dim, Fn fn) {
  if (kind == 1 && dim == 1) return fn(Tag<1, EpanechnikovKernel<1, T>>());
  if (kind == 2 && dim == 1) return fn(Tag<1, GaussianKernel<1, T>>());
  if (kind == 3 && dim == 2) return fn(Tag<2, EpanechnikovProductKernel2d<T>>());
  if (kind == 4 && dim == 2) return fn(Tag<2, GaussianProductKernel2d<T>>());
  return -1;
}

template <class Fn>
int dispatch_conv(int kind, int dim, int kernel_float, Fn fn) {
  return kernel_float ? dispatch_conv_t<float>(kind, dim, fn) : dispatch_conv_t<double>(kind, dim, fn);
}

// The 2-D kinds only (lsq_numint_cross_validate takes KernelDensity<2,...>).
template <class T, class Fn>
int dispatch2_t(int kind, Fn fn) {
  switch (kind) {
    case 1: return fn(Tag<2, EpanechnikovKernel<2, T>>());
    case 2: return fn(Tag<2, GaussianKernel<2, T>>());
    case 3: return fn(Tag<2, EpanechnikovProductKernel2d<T>>());
    case 4: return fn(Tag<2, GaussianProductKernel2d<T>>());
  }
  return -1;
}

template <class Fn>
int dispatch2(int kind, int kernel_float, Fn fn) {
  return kernel_float ? dispatch2_t<float>(kind, fn) : dispatch2_t<double>(kind, fn);
}

template <class KDE>
std::vector<typename KDE::DataPointType> make_points(const double *x, const double *w, size_t n) {
  using P = typename KDE::DataPointType;
  constexpr int D = KDE::dim();
  std::vector<P> pts;
  pts.reserve(n);
  for (size_t i = 0; i < n; ++i) {
    typename P::PointType g;
    for (int d = 0; d < D; ++d) g[d] = x[i * D + d];
    pts.emplace_back(g, typename KDE::AttributeType(w ? w[i] : 1.0));
  }
  return pts;
}

template <class F>
int guarded(F f) {
  try {
    return f();
  } catch (const std::exception &e) {
    std::cerr << "ref_driver: " << e.what() << std::endl;
    return -2;
  }
}

// The kinds whose kernels can be sampled (simulate): EpanechnikovKernel<1>,
// GaussianKernel<1|2>, EpanechnikovProductKernel2d, GaussianProductKernel2d.
template <class T, class Fn>
int dispatch_sim_t(int kind, int dim, Fn fn) {
  if (kind == 1 && dim == 1) return fn(Tag<1, EpanechnikovKernel<1, T>>());
  if (kind == 2 && dim == 1) return fn(Tag<1, GaussianKernel<1, T>>());
  if (kind == 2 && dim == 2) return fn(Tag<2, GaussianKernel<2, T>>());
  if (kind == 3 && dim == 2) return fn(Tag<2, EpanechnikovProductKernel2d<T>>());
  if (kind == 4 && dim == 2) return fn(Tag<2, GaussianProductKernel2d<T>>());
  return -1;
}

template <class Fn>
int dispatch_sim(int kind, int dim, int kernel_float, Fn fn) {
  return kernel_float ? dispatch_sim_t<float>(kind, dim, fn) : dispatch_sim_t<double>(kind, dim, fn);
}

}  // namespace

extern "C" {

// Kernel value: normalization() and unnormalized_eval(p, q, a).
int ref_kernel(int kind, int dim, int kernel_float, double hx, double hy, const double *p, const double *q,
               double a, double *norm_out, double *u_out) {
  return dispatch(kind, dim, kernel_float, [&](auto tag) -> int {
    using K = typename decltype(tag)::Kernel;
    using T = typename K::FloatType;
    constexpr int D = decltype(tag)::dim;
    K k;
    set_bw(k, hx, hy);
    Point<D, T> pp, qq;
    for (int d = 0; d < D; ++d) {
      pp[d] = (T)p[d];
      qq[d] = (T)q[d];
    }
    *norm_out = (double)k.normalization();
    *u_out = (double)k.unnormalized_eval(pp, qq, (T)a);
    return 0;
  });
}

// KernelDensity::direct_eval(queries): values in query order.
int ref_direct_eval(int kind, int dim, int kernel_float, double hx, double hy, const double *pts,
                    const double *w, size_t n, int leaf_nmax, const double *qry, size_t nq, double *out) {
  return guarded([&] {
    return dispatch(kind, dim, kernel_float, [&](auto tag) -> int {
      using K = typename decltype(tag)::Kernel;
      constexpr int D = decltype(tag)::dim;
      using KDE = KernelDensity<D, K, double>;
      KDE kde(make_points<KDE>(pts, w, n), leaf_nmax);
      set_bw(kde.kernel(), hx, hy);
      auto q = make_points<KDE>(qry, nullptr, nq);
      kde.direct_eval(q);
      for (size_t i = 0; i < nq; ++i) out[i] = q[i].attributes().value();
      return 0;
    });
  });
}

// Timing of KernelDensity::direct_eval alone: the KernelDensity (its kd tree)
// and the query points are built first, outside the timed region; then the
// queries are split into `threads` contiguous chunks evaluated concurrently
// with the per-point direct_eval, as test/test_kde_cppthread.cc:88-106 splits
// its queries over std::thread workers.  *seconds_out = wall time of the
// evaluation only; out receives the values in query order.
int ref_direct_eval_timed(int kind, int dim, int kernel_float, double hx, double hy, const double *pts,
                          const double *w, size_t n, int leaf_nmax, const double *qry, size_t nq, int threads,
                          double *out, double *seconds_out) {
  return guarded([&] {
    return dispatch(kind, dim, kernel_float, [&](auto tag) -> int {
      using K = typename decltype(tag)::Kernel;
      constexpr int D = decltype(tag)::dim;
      using KDE = KernelDensity<D, K, double>;
      KDE kde(make_points<KDE>(pts, w, n), leaf_nmax);
      set_bw(kde.kernel(), hx, hy);
      auto q = make_points<KDE>(qry, nullptr, nq);
      const int t = threads < 1 ? 1 : threads;
      auto segment = [&](size_t lo, size_t hi) {
        for (size_t i = lo; i < hi; ++i) kde.direct_eval(q[i]);
      };
      const auto t0 = std::chrono::steady_clock::now();
      std::vector<std::thread> workers;
      for (int c = 1; c < t; ++c) workers.emplace_back(segment, nq * c / t, nq * (c + 1) / t);
      segment(0, nq / t);
      for (auto &th : workers) th.join();
      const auto t1 = std::chrono::steady_clock::now();
      *seconds_out = std::chrono::duration<double>(t1 - t0).count();
      for (size_t i = 0; i < nq; ++i) out[i] = q[i].attributes().value();
      return 0;
    });
  });
}

// n_samples calls of the reference's KernelDensity::simulate(e, p) with
// std::mt19937_64(seed): out row-major (n_samples x dim).
int ref_simulate(int kind, int dim, int kernel_float, double hx, double hy, const double *pts, const double *w,
                 size_t n, int leaf_nmax, unsigned long long seed, size_t n_samples, double *out) {
  return guarded([&] {
    return dispatch_sim(kind, dim, kernel_float, [&](auto tag) -> int {
      using K = typename decltype(tag)::Kernel;
      constexpr int D = decltype(tag)::dim;
      using KDE = KernelDensity<D, K, double>;
      KDE kde(make_points<KDE>(pts, w, n), leaf_nmax);
      set_bw(kde.kernel(), hx, hy);
      std::mt19937_64 e(seed);
      std::vector<double> p;
      for (size_t k = 0; k < n_samples; ++k) {
        kde.simulate(e, p);
        for (int d = 0; d < D; ++d) out[k * D + d] = p[d];
      }
      return 0;
    });
  });
}

// The data tree's point order: coordinates, weight, mass, abw per point.
// Returns the number of points after duplicate merging.
long ref_tree_points(int kind, int dim, int kernel_float, double hx, double hy, const double *pts,
                     const double *w, size_t n, int leaf_nmax, double *coords_out, double *w_out,
                     double *mass_out, double *abw_out) {
  return guarded([&] {
    return dispatch(kind, dim, kernel_float, [&](auto tag) -> int {
      using K = typename decltype(tag)::Kernel;
      constexpr int D = decltype(tag)::dim;
      using KDE = KernelDensity<D, K, double>;
      KDE kde(make_points<KDE>(pts, w, n), leaf_nmax);
      set_bw(kde.kernel(), hx, hy);
      const auto &tp = kde.points();
      for (size_t i = 0; i < tp.size(); ++i) {
        for (int d = 0; d < D; ++d) coords_out[i * D + d] = tp[i][d];
        w_out[i] = tp[i].attributes().weight();
        mass_out[i] = tp[i].attributes().mass();
        abw_out[i] = tp[i].attributes().abw();
      }
      return (int)tp.size();
    });
  });
}

// Dual-tree KernelDensity::eval(queries, rel, abs, leaf): the queries come
// back in the query tree's order (sorted, merged, partitioned).  Returns
// their number and writes coordinates, lower and upper bounds.
long ref_tree_eval(int kind, int dim, int kernel_float, double hx, double hy, const double *pts, const double *w,
                   size_t n, int leaf_nmax, const double *qry, size_t nq, double rel, double abs_err,
                   int q_leaf_nmax, double *q_out, double *lower, double *upper) {
  return guarded([&] {
    return dispatch(kind, dim, kernel_float, [&](auto tag) -> int {
      using K = typename decltype(tag)::Kernel;
      constexpr int D = decltype(tag)::dim;
      using KDE = KernelDensity<D, K, double>;
      KDE kde(make_points<KDE>(pts, w, n), leaf_nmax);
      set_bw(kde.kernel(), hx, hy);
      auto q = make_points<KDE>(qry, nullptr, nq);
      std::ostringstream sink;
      std::streambuf *old = std::cerr.rdbuf(sink.rdbuf());
      kde.eval(q, rel, abs_err, q_leaf_nmax);
      std::cerr.rdbuf(old);
      for (size_t i = 0; i < q.size(); ++i) {
        for (int d = 0; d < D; ++d) q_out[i * D + d] = q[i][d];
        lower[i] = q[i].attributes().lower();
        upper[i] = q[i].attributes().upper();
      }
      return (int)q.size();
    });
  });
}

// likelihood_cross_validate(rel, abs).
int ref_likelihood_cv(int kind, int dim, int kernel_float, double hx, double hy, const double *pts,
                      const double *w, size_t n, int leaf_nmax, double rel, double abs_err, double *out) {
  return guarded([&] {
    return dispatch(kind, dim, kernel_float, [&](auto tag) -> int {
      using K = typename decltype(tag)::Kernel;
      constexpr int D = decltype(tag)::dim;
      using KDE = KernelDensity<D, K, double>;
      KDE kde(make_points<KDE>(pts, w, n), leaf_nmax);
      set_bw(kde.kernel(), hx, hy);
      std::ostringstream sink;
      std::streambuf *old = std::cerr.rdbuf(sink.rdbuf());
      *out = kde.likelihood_cross_validate(rel, abs_err);
      std::cerr.rdbuf(old);
      return 0;
    });
  });
}

// lsq_convolution_cross_validate(rel, abs).
int ref_lsq_conv_cv(int kind, int dim, int kernel_float, double hx, double hy, const double *pts,
                    const double *w, size_t n, int leaf_nmax, double rel, double abs_err, double *out) {
  return guarded([&] {
    return dispatch_conv(kind, dim, kernel_float, [&](auto tag) -> int {
      using K = typename decltype(tag)::Kernel;
      constexpr int D = decltype(tag)::dim;
      using KDE = KernelDensity<D, K, double>;
      KDE kde(make_points<KDE>(pts, w, n), leaf_nmax);
      set_bw(kde.kernel(), hx, hy);
      std::ostringstream sink;
      std::streambuf *old = std::cerr.rdbuf(sink.rdbuf());
      *out = kde.lsq_convolution_cross_validate(rel, abs_err);
      std::cerr.rdbuf(old);
      return 0;
    });
  });
}

// lsq_numint_cross_validate(kde, start_x, end_x, steps_x, ...), 2-D only.
int ref_lsq_numint_cv(int kind, int kernel_float, double hx, double hy, const double *pts, const double *w,
                      size_t n, int leaf_nmax, double start_x, double end_x, int steps_x, double start_y,
                      double end_y, int steps_y, double rel, double abs_err, int q_leaf_nmax, double *out) {
  return guarded([&] {
    return dispatch2(kind, kernel_float, [&](auto tag) -> int {
      using K = typename decltype(tag)::Kernel;
      constexpr int D = decltype(tag)::dim;
      using KDE = KernelDensity<D, K, double>;
      KDE kde(make_points<KDE>(pts, w, n), leaf_nmax);
      set_bw(kde.kernel(), hx, hy);
      std::ostringstream sink;
      std::streambuf *old = std::cerr.rdbuf(sink.rdbuf());
      *out = lsq_numint_cross_validate(kde, start_x, end_x, steps_x, start_y, end_y, steps_y, rel, abs_err,
                                       q_leaf_nmax);
      std::cerr.rdbuf(old);
      return 0;
    });
  });
}

// adapt_density(alpha, rel, abs): the data points in tree order with their
// adapted abw and mass.  Returns the number of points.
long ref_adapt(int kind, int dim, int kernel_float, double hx, double hy, const double *pts, const double *w,
               size_t n, int leaf_nmax, double alpha, double rel, double abs_err, double *coords_out,
               double *w_out, double *abw_out, double *mass_out) {
  return guarded([&] {
    return dispatch(kind, dim, kernel_float, [&](auto tag) -> int {
      using K = typename decltype(tag)::Kernel;
      constexpr int D = decltype(tag)::dim;
      using KDE = KernelDensity<D, K, double>;
      KDE kde(make_points<KDE>(pts, w, n), leaf_nmax);
      set_bw(kde.kernel(), hx, hy);
      std::ostringstream sink;
      std::streambuf *old = std::cerr.rdbuf(sink.rdbuf());
      kde.adapt_density(alpha, rel, abs_err);
      std::cerr.rdbuf(old);
      const auto &tp = kde.points();
      for (size_t i = 0; i < tp.size(); ++i) {
        for (int d = 0; d < D; ++d) coords_out[i * D + d] = tp[i][d];
        w_out[i] = tp[i].attributes().weight();
        abw_out[i] = tp[i].attributes().abw();
        mass_out[i] = tp[i].attributes().mass();
      }
      return (int)tp.size();
    });
  });
}

}  // extern "C"
