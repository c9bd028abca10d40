// api_driver.cc -- extern "C" entry points over the C++ drop-in API
// (bbrcit_kde_amd/include: KernelDensity, Kdtree, HipDirectKde, kernels),
// loaded by the tests with ctypes.  Test infrastructure only.
//
// Kinds use the numbering of include/bbrcit_hip.h; kernel_float selects the
// kernel's float type (0: double, 1: float).
#include <random>
#include <vector>

#include <CvGrid.h>
#include <HipDirectKde.h>
#include <KernelDensity.h>
#include <Kernels/ConvKernelAssociator.h>
#include <MarginalDensityAssociator.h>

#include <exception>
#include <iostream>
#include <sstream>
#include <stdexcept>

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

template <class T, class Fn>
int dispatch_t(int kind, int dim, Fn fn) {
  switch (kind) {
    case 1:
      if (dim == 1) return fn(Tag<1, EpanechnikovKernel<1, T>>());
      if (dim == 2) return fn(Tag<2, EpanechnikovKernel<2, T>>());
      break;
    case 2:
      if (dim == 1) return fn(Tag<1, GaussianKernel<1, T>>());
      if (dim == 2) return fn(Tag<2, GaussianKernel<2, T>>());
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
int dispatch(int kind, int dim, int kf, Fn fn) {
  return kf ? dispatch_t<float>(kind, dim, fn) : dispatch_t<double>(kind, dim, fn);
}

template <class T, class Fn>
int dispatch_conv_t(int kind, int dim, Fn fn) {
  if (kind == 1 && dim == 1) return fn(Tag<1, EpanechnikovKernel<1, T>>());
  if (kind == 2 && dim == 1) return fn(Tag<1, GaussianKernel<1, T>>());
  if (kind == 3 && dim == 2) return fn(Tag<2, EpanechnikovProductKernel2d<T>>());
  if (kind == 4 && dim == 2) return fn(Tag<2, GaussianProductKernel2d<T>>());
  return -1;
}

template <class Fn>
int dispatch_conv(int kind, int dim, int kf, Fn fn) {
  return kf ? dispatch_conv_t<float>(kind, dim, fn) : dispatch_conv_t<double>(kind, dim, fn);
}

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

thread_local std::string g_err;

template <class F>
int guarded(F f) {
  try {
    g_err.clear();
    return f();
  } catch (const std::out_of_range &e) {
    g_err = e.what();
    return -102;
  } catch (const std::invalid_argument &e) {
    g_err = e.what();
    return -101;
  } catch (const std::exception &e) {
    g_err = e.what();
    return -100;
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

const char *api_last_error() { return g_err.c_str(); }

// Kdtree order of the data points (host only).
long api_tree_points(int kind, int dim, int kf, double hx, double hy, const double *pts, const double *w, size_t n,
                     int leaf_nmax, double *coords_out, double *w_out, double *mass_out) {
  return guarded([&] {
    return dispatch(kind, dim, kf, [&](auto tag) -> int {
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
      }
      return (int)tp.size();
    });
  });
}

// Kernel normalization() and unnormalized_eval(p, q, a) (host only).
int api_kernel(int kind, int dim, int kf, double hx, double hy, const double *p, const double *q, double a,
               double *norm_out, double *u_out) {
  return guarded([&] {
    return dispatch(kind, dim, kf, [&](auto tag) -> int {
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
  });
}

// KernelDensity::direct_eval(queries): values in query order.
int api_direct_eval(int kind, int dim, int kf, double hx, double hy, const double *pts, const double *w, size_t n,
                    int leaf_nmax, const double *qry, size_t nq, double *out) {
  return guarded([&] {
    return dispatch(kind, dim, kf, [&](auto tag) -> int {
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

// MarginalDensityAssociator: the dim_idx marginal of a 2-D product-kernel
// density (kind 3 or 4), adapted first with alpha when alpha > 0, evaluated
// at the 1-D queries q1; the 2-D density itself at the 2-D queries q2.
int api_marginal_eval(int kind, int kf, double hx, double hy, const double *pts, const double *w, size_t n,
                      int leaf_nmax, double alpha, int dim_idx, const double *q1, size_t n1, double *out1,
                      const double *q2, size_t n2, double *out2) {
  return guarded([&] {
    auto run = [&](auto tag) -> int {
      using K = typename decltype(tag)::Kernel;
      using KDE = KernelDensity<2, K, double>;
      KDE kde(make_points<KDE>(pts, w, n), leaf_nmax);
      set_bw(kde.kernel(), hx, hy);
      if (alpha > 0) kde.adapt_density(alpha);
      using MDA = MarginalDensityAssociator<KDE>;
      auto marginal = dim_idx == 0 ? MDA::marginal_density_x(kde) : MDA::marginal_density_y(kde);
      using M = decltype(marginal);
      auto q = make_points<M>(q1, nullptr, n1);
      marginal.direct_eval(q);
      for (size_t i = 0; i < n1; ++i) out1[i] = q[i].attributes().value();
      auto r = make_points<KDE>(q2, nullptr, n2);
      kde.direct_eval(r);
      for (size_t i = 0; i < n2; ++i) out2[i] = r[i].attributes().value();
      return 0;
    };
    if (kind == 3) return kf ? run(Tag<2, EpanechnikovProductKernel2d<float>>()) : run(Tag<2, EpanechnikovProductKernel2d<double>>());
    if (kind == 4) return kf ? run(Tag<2, GaussianProductKernel2d<float>>()) : run(Tag<2, GaussianProductKernel2d<double>>());
    return -1;
  });
}

// Dual-tree KernelDensity::eval(queries, rel, abs, leaf): queries in the
// query tree's order with their bounds.
long api_tree_eval(int kind, int dim, int kf, double hx, double hy, const double *pts, const double *w, size_t n,
                   int leaf_nmax, const double *qry, size_t nq, double rel, double abs_err, int q_leaf_nmax,
                   double *q_out, double *lower, double *upper) {
  return guarded([&] {
    return dispatch(kind, dim, kf, [&](auto tag) -> int {
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

// Samples of KernelDensity::simulate with std::mt19937_64(seed), out row-major
// (n_samples x dim): batched != 0 takes the batched call (picks and shifts on
// the GPU), else n_samples single-point calls on the host.
int api_simulate(int kind, int dim, int kf, double hx, double hy, const double *pts, const double *w, size_t n,
                 int leaf_nmax, unsigned long long seed, size_t n_samples, int batched, double *out) {
  return guarded([&] {
    return dispatch_sim(kind, dim, kf, [&](auto tag) -> int {
      using K = typename decltype(tag)::Kernel;
      constexpr int D = decltype(tag)::dim;
      using KDE = KernelDensity<D, K, double>;
      KDE kde(make_points<KDE>(pts, w, n), leaf_nmax);
      set_bw(kde.kernel(), hx, hy);
      std::mt19937_64 e(seed);
      if (batched) {
        std::vector<double> all;
        kde.simulate(e, n_samples, all);
        std::copy(all.begin(), all.end(), out);
      } else {
        std::vector<double> p;
        for (size_t k = 0; k < n_samples; ++k) {
          kde.simulate(e, p);
          for (int d = 0; d < D; ++d) out[k * D + d] = p[d];
        }
      }
      return 0;
    });
  });
}

// The tree of KernelDensity<2, EpanechnikovKernel<2, float>, double> over n
// points (after adapt_density(alpha) when alpha > 0), built with the device
// steps on or off (kdtree_device_build).  points_out: 6 values per tree point
// (x, y, weight, mass, lower abw, upper abw); nodes_out: 13 values per node in
// pre-order (start, end, split, cell x lo/hi, y lo/hi, weight, mass, lower
// abw, upper abw, lower, upper).  Returns the node count (nodes_out holds
// max_nodes rows) or a negative status.
long api_tree_dump(const double *pts, const double *w, size_t n, int leaf_nmax, int use_device, double alpha,
                   double *points_out, size_t max_nodes, double *nodes_out) {
  return guarded([&]() -> int {
    using KDE = KernelDensity<2, EpanechnikovKernel<2, float>, double>;
    const bool saved = kdtree_device_build();
    kdtree_device_build() = use_device != 0;
    try {
      KDE kde(make_points<KDE>(pts, w, n), leaf_nmax);
      kde.kernel().set_bandwidth(0.2f);
      if (alpha > 0) kde.adapt_density(alpha);
      kdtree_device_build() = saved;
      const auto &tp = kde.points();
      for (size_t i = 0; i < tp.size(); ++i) {
        const auto &a = tp[i].attributes();
        double *o = points_out + 6 * i;
        o[0] = tp[i][0];
        o[1] = tp[i][1];
        o[2] = a.weight();
        o[3] = a.mass();
        o[4] = a.lower_abw();
        o[5] = a.upper_abw();
      }
      size_t k = 0;
      kde.data_tree().visit_nodes([&](size_t i, size_t j, double split, const auto &cell, const auto &a, bool) {
        if (k < max_nodes) {
          double *o = nodes_out + 13 * k;
          o[0] = (double)i;
          o[1] = (double)j;
          o[2] = split;
          o[3] = cell[0].lower();
          o[4] = cell[0].upper();
          o[5] = cell[1].lower();
          o[6] = cell[1].upper();
          o[7] = a.weight();
          o[8] = a.mass();
          o[9] = a.lower_abw();
          o[10] = a.upper_abw();
          o[11] = a.lower();
          o[12] = a.upper();
        }
        ++k;
      });
      return (int)k;
    } catch (...) {
      kdtree_device_build() = saved;
      throw;
    }
  });
}

int api_likelihood_cv(int kind, int dim, int kf, double hx, double hy, const double *pts, const double *w, size_t n,
                      int leaf_nmax, double *out) {
  return guarded([&] {
    return dispatch(kind, dim, kf, [&](auto tag) -> int {
      using K = typename decltype(tag)::Kernel;
      constexpr int D = decltype(tag)::dim;
      using KDE = KernelDensity<D, K, double>;
      KDE kde(make_points<KDE>(pts, w, n), leaf_nmax);
      set_bw(kde.kernel(), hx, hy);
      *out = kde.likelihood_cross_validate();
      return 0;
    });
  });
}

int api_lsq_conv_cv(int kind, int dim, int kf, double hx, double hy, const double *pts, const double *w, size_t n,
                    int leaf_nmax, double *out) {
  return guarded([&] {
    return dispatch_conv(kind, dim, kf, [&](auto tag) -> int {
      using K = typename decltype(tag)::Kernel;
      constexpr int D = decltype(tag)::dim;
      using KDE = KernelDensity<D, K, double>;
      KDE kde(make_points<KDE>(pts, w, n), leaf_nmax);
      set_bw(kde.kernel(), hx, hy);
      *out = kde.lsq_convolution_cross_validate();
      return 0;
    });
  });
}

int api_lsq_numint_cv(int kind, int kf, double hx, double hy, const double *pts, const double *w, size_t n,
                      int leaf_nmax, double sx, double ex, int nx, double sy, double ey, int ny, double *out) {
  return guarded([&] {
    auto fn = [&](auto tag) -> int {
      using K = typename decltype(tag)::Kernel;
      using KDE = KernelDensity<2, K, double>;
      KDE kde(make_points<KDE>(pts, w, n), leaf_nmax);
      set_bw(kde.kernel(), hx, hy);
      *out = lsq_numint_cross_validate(kde, sx, ex, nx, sy, ey, ny);
      return 0;
    };
    return kf ? dispatch2_t<float>(kind, fn) : dispatch2_t<double>(kind, fn);
  });
}

// adapt_density(alpha): data points in tree order with their abw and mass;
// then the adaptive density at the queries (direct_eval) and its LCV.
long api_adapt(int kind, int dim, int kf, double hx, double hy, const double *pts, const double *w, size_t n,
               int leaf_nmax, double alpha, double *coords_out, double *abw_out, double *mass_out,
               const double *qry, size_t nq, double *q_values, double *lcv_out) {
  return guarded([&] {
    return dispatch(kind, dim, kf, [&](auto tag) -> int {
      using K = typename decltype(tag)::Kernel;
      constexpr int D = decltype(tag)::dim;
      using KDE = KernelDensity<D, K, double>;
      KDE kde(make_points<KDE>(pts, w, n), leaf_nmax);
      set_bw(kde.kernel(), hx, hy);
      kde.adapt_density(alpha);
      const auto &tp = kde.points();
      for (size_t i = 0; i < tp.size(); ++i) {
        for (int d = 0; d < D; ++d) coords_out[i * D + d] = tp[i][d];
        abw_out[i] = tp[i].attributes().abw();
        mass_out[i] = tp[i].attributes().mass();
      }
      if (nq > 0) {
        auto q = make_points<KDE>(qry, nullptr, nq);
        kde.direct_eval(q);
        for (size_t i = 0; i < nq; ++i) q_values[i] = q[i].attributes().value();
      }
      if (lcv_out) *lcv_out = kde.likelihood_cross_validate();
      return (int)tp.size();
    });
  });
}

// HipDirectKde over (refs, queries) with unit masses scaled by 1/n:
// unnormalized_eval on [r_i, r_j] x [q_i, q_j]; returns the exception class
// on misuse (-102 out_of_range, -101 invalid_argument).
int api_hip_direct_range(int kind, int dim, int kf, double hx, double hy, const double *pts, size_t n,
                         const double *qry, size_t nq, size_t r_i, size_t r_j, size_t q_i, size_t q_j,
                         int normalize, double *out) {
  return guarded([&] {
    return dispatch(kind, dim, kf, [&](auto tag) -> int {
      using K = typename decltype(tag)::Kernel;
      using T = typename K::FloatType;
      constexpr int D = decltype(tag)::dim;
      using KDE = KernelDensity<D, K, double>;
      auto r = make_points<KDE>(pts, nullptr, n);
      for (auto &p : r) p.attributes().set_mass(1.0 / n);
      auto q = make_points<KDE>(qry, nullptr, nq);
      HipDirectKde<D, T, K> dev(r, q);
      set_bw(dev.kernel(), hx, hy);
      if (dev.reference_size() != n || dev.query_size() != nq) throw std::logic_error("sizes");
      std::vector<T> res;
      if (normalize)
        dev.eval(r_i, r_j, q_i, q_j, res);
      else
        dev.unnormalized_eval(r_i, r_j, q_i, q_j, res);
      HipDirectKde<D, T, K> copy(dev);  // copy construction re-uploads
      std::vector<T> res2;
      copy.unnormalized_eval(r_i, r_j, q_i, q_j, res2);
      const double c = normalize ? (double)dev.kernel().normalization() : 1.0;
      for (size_t k = 0; k <= q_j - q_i; ++k) {
        out[k] = (double)res[k];
        if (std::abs((double)res2[k] * c - (double)res[k]) > 1e-6 * std::abs((double)res[k]) + 1e-300)
          throw std::logic_error("copy differs");
      }
      return 0;
    });
  });
}

// Grid search over (hx, hy) pairs with the product kernels (kind 3 or 4),
// kernel_float 0: scores from lsq_convolution_cross_validate_grid /
// likelihood_cross_validate_grid (one batched device pipeline) and, for
// comparison, from the per-bandwidth KernelDensity calls grid_search makes.
int api_cv_grid(int kind, const double *pts, const double *w, size_t n, int leaf_nmax, const double *hx,
                const double *hy, size_t npairs, int adapt, double alpha, double *lcv_grid, double *lscv_grid,
                double *lcv_loop, double *lscv_loop) {
  return guarded([&] {
    auto fn = [&](auto kernel_tag) -> int {
      using K = decltype(kernel_tag);
      using KDE = KernelDensity<2, K, double>;
      KDE kde(make_points<KDE>(pts, w, n), leaf_nmax);
      if (adapt) {
        kde.kernel().set_bandwidths(hx[0], hy[0]);
        kde.adapt_density(alpha);
      }
      std::vector<std::pair<double, double>> grid;
      for (size_t k = 0; k < npairs; ++k) grid.emplace_back(hx[k], hy[k]);
      const std::vector<double> l = likelihood_cross_validate_grid(kde, grid);
      const std::vector<double> q = lsq_convolution_cross_validate_grid(kde, grid);
      for (size_t k = 0; k < npairs; ++k) {
        lcv_grid[k] = l[k];
        lscv_grid[k] = q[k];
        kde.kernel().set_bandwidths(hx[k], hy[k]);
        lcv_loop[k] = kde.likelihood_cross_validate();
        lscv_loop[k] = kde.lsq_convolution_cross_validate();
      }
      return 0;
    };
    if (kind == 3) return fn(EpanechnikovProductKernel2d<double>());
    if (kind == 4) return fn(GaussianProductKernel2d<double>());
    throw std::invalid_argument("api_cv_grid: kind must be 3 or 4");
  });
}

// The reference's two bandwidth-grid layouts (grid_search/cv_utils_impl.h).
int api_grid_layouts(const double *v1, size_t n1, const double *v2, size_t n2, double sx, double ex, int nx,
                     double sy, double ey, int ny, double *crossed, double *generated) {
  return guarded([&] {
    const auto c = make_crossed_vector_pairs(std::vector<double>(v1, v1 + n1), std::vector<double>(v2, v2 + n2));
    for (size_t k = 0; k < c.size(); ++k) {
      crossed[2 * k] = c[k].first;
      crossed[2 * k + 1] = c[k].second;
    }
    const auto g = generate_2d_bandwidth_grid(sx, ex, nx, sy, ey, ny);
    for (size_t k = 0; k < g.size(); ++k) {
      generated[2 * k] = g[k].first;
      generated[2 * k + 1] = g[k].second;
    }
    return 0;
  });
}

}  // extern "C"
