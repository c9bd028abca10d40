// KernelDensity.h -- kernel density estimator over a kd-tree of weighted points,
// evaluated on an MI355X.
//
// Public interface of the reference's include/KernelDensity.h:44-311 (GPU-build
// signatures, which add a trailing block_size to the CPU-build ones, so both
// kinds of call sites compile).  Semantics kept from
// include/KernelDensityImpl.h:
//   * construction normalises the weights to sum 1 and sets
//     mass = weight * abw^-D (388-455);
//   * density values are C(h) * sum_r mass_r U(q - r; h abw_r), stored as the
//     lower/upper bounds of each query's attributes, value() = midpoint;
//   * direct_eval keeps the query order; eval(vector) returns the queries in
//     the query tree's order (sorted, duplicates merged, partitioned);
//   * the dual-tree evaluation meets the reference's bounds and tolerance
//     contract (681-1110).
// Device work goes through the C ABI (bbrcit_hip.h):
//   * every direct sum is one GPU evaluation (queries sorted along a Z curve
//     on the device, results in the caller's order);
//   * the dual-tree evaluation is one bounded GPU evaluation
//     (bk_eval_bounded): the pruning is decided per (query block, reference
//     tile) on the device against the same bounds contract, instead of a
//     host recursion with one synchronous launch per leaf pair;
//   * cross-validation scores and the adaptive pass are exact direct sums on
//     the GPU with FP64 reductions (the leave-one-out pair excluded by index);
//     rel_err / abs_err are accepted and met trivially.
#ifndef BBRCIT_AMD_KERNELDENSITY_H_
#define BBRCIT_AMD_KERNELDENSITY_H_

#include <bbrcit_hip.h>

#include <algorithm>
#include <cmath>
#include <iomanip>
#include <iostream>
#include <limits>
#include <memory>
#include <random>
#include <stdexcept>
#include <thread>
#include <utility>
#include <vector>

#include "Attributes/AdaKdeAttributes.h"
#include "HipDirectKde.h"
#include "KdeTraits.h"
#include "Kdtree.h"
#include "Kernels/ConvKernelAssociator.h"
#include "Kernels/EpanechnikovKernel.h"

namespace bbrcit {

template <int D, typename KernelT = EpanechnikovKernel<D, double>, typename FloatT = double,
          typename AttrT = AdaKdeAttributes<FloatT>>
class KernelDensity;

template <typename KT, typename FT, typename AT>
FT lsq_numint_cross_validate(const KernelDensity<2, KT, FT, AT> &kde, FT start_x, FT end_x, int steps_x,
                             FT start_y, FT end_y, int steps_y, FT rel_err = 1e-6, FT abs_err = 1e-8,
                             int qtree_leaf_nmax = 1024, size_t block_size = 128);

template <int D, typename KernelT, typename FloatT, typename AttrT>
class KernelDensity {
 public:
  using FloatType = FloatT;
  using KdtreeType = Kdtree<D, AttrT, FloatT>;
  using DataPointType = typename KdtreeType::DataPointType;
  using KernelType = KernelT;
  using KernelFloatType = typename KernelType::FloatType;
  using AttributeType = AttrT;
  static constexpr int dim() { return D; }

 private:
  using TreeNodeType = typename KdtreeType::Node;
  using GeomPointType = typename DataPointType::PointType;
  template <typename K>
  using Device = HipDirectKde<D, KernelFloatType, K>;

 public:
  KernelDensity() = default;
  KernelDensity(const std::vector<DataPointType> &data, int leaf_nmax = 2) {
    std::vector<DataPointType> pts = data;
    init(std::move(pts), leaf_nmax);
  }
  KernelDensity(std::vector<DataPointType> &&data, int leaf_nmax = 2) { init(std::move(data), leaf_nmax); }

  KernelDensity(const KernelDensity &o) : kernel_(o.kernel_), data_tree_(o.data_tree_), cum_weights_(o.cum_weights_) {}
  KernelDensity(KernelDensity &&o) noexcept = default;
  KernelDensity &operator=(KernelDensity o) {
    swap(*this, o);
    return *this;
  }
  ~KernelDensity() = default;

  friend void swap(KernelDensity &a, KernelDensity &b) noexcept {
    using std::swap;
    swap(a.kernel_, b.kernel_);
    swap(a.data_tree_, b.data_tree_);
    swap(a.cum_weights_, b.cum_weights_);
    swap(a.self_dev_, b.self_dev_);
  }

  const KernelType &kernel() const { return kernel_; }
  KernelType &kernel() { return kernel_; }
  void set_kernel(const KernelType &k) { kernel_ = k; }

  size_t size() const { return data_tree_.size(); }
  const std::vector<DataPointType> &points() const { return data_tree_.points(); }
  const KdtreeType &data_tree() const { return data_tree_; }

  // ---- direct evaluation ----------------------------------------------------
  FloatT direct_eval(DataPointType &p) const {
    std::vector<DataPointType> one{p};
    direct_eval(one);
    p.attributes().set_lower(one[0].attributes().lower());
    p.attributes().set_upper(one[0].attributes().upper());
    return one[0].attributes().value();
  }

  void direct_eval(std::vector<DataPointType> &queries, size_t block_size = 128) const {
    (void)block_size;
    if (queries.empty()) return;
    require_data();
    Device<KernelType> dev(points(), queries);
    dev.kernel() = kernel_;
    std::vector<double> s;
    eval_all(dev, queries.size(), s);
    const double c = (double)kernel_.normalization();
    for (size_t i = 0; i < queries.size(); ++i) {
      queries[i].attributes().set_lower((FloatT)(s[i] * c));
      queries[i].attributes().set_upper((FloatT)(s[i] * c));
    }
  }

  // ---- tree evaluation --------------------------------------------------------
  // one point: evaluated exactly (a direct sum meets any rel/abs tolerance)
  FloatT eval(DataPointType &p, FloatType rel_err, FloatType abs_err) const {
    (void)rel_err;
    (void)abs_err;
    return direct_eval(p);
  }

  void eval(std::vector<DataPointType> &queries, FloatType rel_err, FloatType abs_err, int leaf_nmax = 2,
            size_t block_size = 128) const {
    KdtreeType qtree(std::move(queries), leaf_nmax);
    eval(qtree, rel_err, abs_err, block_size);
    queries = std::move(qtree.points_);
  }

  void eval(KdtreeType &query_tree, FloatType rel_err, FloatType abs_err, size_t block_size = 128) const {
    (void)block_size;
    tree_eval(query_tree, kernel_, rel_err, abs_err);
  }

  // ---- adaptive densities ------------------------------------------------------
  void adapt_density(FloatType alpha, FloatType rel_err = 1e-6, FloatType abs_err = 1e-6,
                     size_t block_size = 128) {
    (void)rel_err;
    (void)abs_err;
    (void)block_size;
    if (alpha < 0 || alpha > 1) throw std::invalid_argument("adapt_density: alpha must be between [0, 1]. ");
    unadapt_density();
    if (alpha == 0) return;
    require_data();
    auto &pts = data_tree_.points_;
    std::vector<double> w(pts.size()), lambda(pts.size());
    for (size_t i = 0; i < pts.size(); ++i) w[i] = (double)pts[i].attributes().weight();
    Device<KernelType> &dev = self_device();
    const bk_kernel_desc d = kernel_.desc();
    double g = 0.0;
    hip_detail::check(bk_adapt(dev.context(), &d, (double)alpha, w.data(), &g, lambda.data()));
    for (size_t i = 0; i < pts.size(); ++i) {
      auto &a = pts[i].attributes();
      a.set_lower_abw((FloatT)lambda[i]);
      a.set_upper_abw((FloatT)lambda[i]);
      a.set_mass((FloatT)(w[i] * std::pow(lambda[i], -(double)D)));
    }
    data_tree_.refresh_node_attributes(data_tree_.root_);
    self_dev_.reset();  // masses and abw changed
  }

  void unadapt_density() {
    auto &pts = data_tree_.points_;
    for (auto &p : pts) {
      p.attributes().set_lower_abw(ConstantTraits<FloatT>::one());
      p.attributes().set_upper_abw(ConstantTraits<FloatT>::one());
      p.attributes().set_mass(p.attributes().weight());
    }
    data_tree_.refresh_node_attributes(data_tree_.root_);
    self_dev_.reset();
  }

  // ---- cross-validation ----------------------------------------------------------
  // sum_i w_i log(f_{-i}(x_i)), f_{-i} the density without point i
  FloatType likelihood_cross_validate(FloatType rel_err = 1e-6, FloatType abs_err = 1e-6,
                                      size_t block_size = 128) const {
    (void)rel_err;
    (void)abs_err;
    (void)block_size;
    require_data();
    const std::vector<double> w = weights();
    const bk_kernel_desc d = kernel_.desc();
    double score = 0.0;
    hip_detail::check(bk_likelihood_cv(self_device().context(), &d, w.data(), &score));
    return (FloatType)score;
  }

  // sum_i w_i f*(x_i) - 2 sum_i w_i f_{-i}(x_i) with the convolution kernel K*
  FloatType lsq_convolution_cross_validate(FloatType rel_err = 1e-6, FloatType abs_err = 1e-6,
                                           size_t block_size = 128) const {
    (void)rel_err;
    (void)abs_err;
    (void)block_size;
    using Conv = typename ConvKernelAssociator<KernelType>::ConvKernelType;
    (void)sizeof(Conv);
    require_data();
    const std::vector<double> w = weights();
    const bk_kernel_desc d = kernel_.desc();
    double score = 0.0;
    hip_detail::check(bk_lsq_conv_cv(self_device().context(), &d, w.data(), &score));
    return (FloatType)score;
  }

  // ---- sampling (host) ---------------------------------------------------------------
  // a point of the density: pick a data point with probability weight, then
  // add a draw of the kernel with that point's bandwidth factor
  // (KernelDensityImpl.h:99-138)
  template <typename RNG>
  void simulate(RNG &e, std::vector<FloatType> &p) const {
    require_data();
    std::uniform_real_distribution<FloatType> u(0, 1);
    auto it = std::upper_bound(cum_weights_.begin(), cum_weights_.end(), u(e));
    if (it == cum_weights_.end()) --it;
    const DataPointType &ref = points()[it - cum_weights_.begin()];
    std::vector<KernelFloatType> k(D, KernelFloatType(0));
    kernel_.simulate(e, k, (KernelFloatType)ref.attributes().abw());
    p.resize(D);
    for (int i = 0; i < D; ++i) p[i] = ref[i] + (FloatType)k[i];
  }

  // n samples at once into out (row-major, n x D): the same values as n
  // calls of simulate(e, p).  The draws are taken from e on the host in the
  // order those calls take them (the pick's uniform, then the kernel's unit
  // draws, sample after sample); the picks (a binary search over the
  // cumulative weights per sample) and the shifts run on the GPU
  // (bk_sampler_draw).  Not in the reference, which samples one point per call.
  template <typename RNG>
  void simulate(RNG &e, size_t n, std::vector<FloatType> &out) const {
    require_data();
    static_assert(D <= 2, "batched sampling supports D = 1 and 2");
    out.resize(n * D);
    if (n == 0) return;
    std::uniform_real_distribution<FloatType> pick(0, 1);
    std::vector<double> u(n), z(n * D);
    KernelFloatType unit[D];
    for (size_t k = 0; k < n; ++k) {
      u[k] = (double)pick(e);
      kernel_.draw_unit(e, unit);
      for (int d = 0; d < D; ++d) z[k * D + d] = (double)unit[d];
    }
    const auto &pts = points();
    std::vector<double> xy(pts.size() * D), cum(cum_weights_.begin(), cum_weights_.end()), abw(pts.size());
    for (size_t i = 0; i < pts.size(); ++i) {
      for (int d = 0; d < D; ++d) xy[i * D + d] = (double)pts[i][d];
      abw[i] = (double)pts[i].attributes().abw();
    }
    const bk_kernel_desc kd = kernel_.desc();
    const double h[2] = {kd.hx, kd.hy};
    bk_sampler *s = nullptr;
    hip_detail::check(bk_sampler_create(0, D, pts.size(), xy.data(), cum.data(), abw.data(), &s));
    std::vector<double> res(n * D);
    const int st = bk_sampler_draw(s, std::is_same<KernelFloatType, float>::value ? 1 : 0, h, n, u.data(), z.data(),
                                   res.data(), nullptr);
    bk_sampler_destroy(s);
    hip_detail::check(st);
    for (size_t k = 0; k < n * D; ++k) out[k] = (FloatType)res[k];
  }

  template <typename RNG>
  DataPointType simulate(RNG &e) const {
    std::vector<FloatType> p;
    simulate(e, p);
    DataPointType q;
    for (int i = 0; i < D; ++i) q[i] = p[i];
    return q;
  }

 private:
  template <typename KT, typename FT, typename AT>
  friend FT lsq_numint_cross_validate(const KernelDensity<2, KT, FT, AT> &, FT, FT, int, FT, FT, int, FT, FT, int,
                                      size_t);

  // device context over the data points with the query set aliased to them;
  // built on first use, dropped when masses or bandwidth factors change
  struct SelfCache {
    std::unique_ptr<Device<KernelType>> dev;
    SelfCache() = default;
    SelfCache(const SelfCache &) {}  // copies start without a device context
    SelfCache(SelfCache &&) noexcept = default;
    SelfCache &operator=(const SelfCache &) {
      dev.reset();
      return *this;
    }
    SelfCache &operator=(SelfCache &&) noexcept = default;
    void reset() { dev.reset(); }
    friend void swap(SelfCache &a, SelfCache &b) noexcept { a.dev.swap(b.dev); }
  };

  void init(std::vector<DataPointType> &&pts, int leaf_nmax) {
    FloatT total = ConstantTraits<FloatT>::zero();
    for (const auto &p : pts) total += p.attributes().weight();
    // per point: weight / total, mass = weight * abw^-D; independent per
    // point, so large sets split over host threads give the same values
    // (pow(1, -D) is exactly 1, so a unit factor skips the call)
    auto normalise = [&pts, total](size_t lo, size_t hi) {
      for (size_t i = lo; i < hi; ++i) {
        auto &a = pts[i].attributes();
        a.set_weight(a.weight() / total);
        const FloatT f = a.abw() == ConstantTraits<FloatT>::one() ? ConstantTraits<FloatT>::one()
                                                                   : (FloatT)std::pow(a.abw(), -D);
        a.set_mass(a.weight() * f);
      }
    };
    const size_t n = pts.size();
    const unsigned hw = std::thread::hardware_concurrency();
    const size_t t = n < (size_t(1) << 18) || hw < 2 ? 1 : std::min<size_t>(hw, 16);
    if (t == 1) {
      normalise(0, n);
    } else {
      std::vector<std::thread> workers;
      for (size_t c = 0; c < t; ++c) workers.emplace_back(normalise, n * c / t, n * (c + 1) / t);
      for (auto &w : workers) w.join();
    }
    data_tree_ = KdtreeType(std::move(pts), leaf_nmax);
    cum_weights_.clear();
    cum_weights_.reserve(points().size());
    FloatT c = 0;
    for (const auto &p : points()) {
      c += p.attributes().weight();
      cum_weights_.push_back(c);
    }
    if (!cum_weights_.empty()) cum_weights_.back() = 1;
  }

  void require_data() const {
    if (data_tree_.empty()) throw std::logic_error("KernelDensity<>: no data points");
  }

  std::vector<double> weights() const {
    std::vector<double> w(points().size());
    for (size_t i = 0; i < w.size(); ++i) w[i] = (double)points()[i].attributes().weight();
    return w;
  }

  Device<KernelType> &self_device() const {
    if (!self_dev_.dev) {
      self_dev_.dev.reset(new Device<KernelType>(points(), SelfQueries{}));
    }
    self_dev_.dev->kernel() = kernel_;
    return *self_dev_.dev;
  }

  // every reference against every query; the device takes the queries in
  // Z-curve order (its workgroups then skip the reference tiles out of reach)
  template <typename K>
  void eval_all(Device<K> &dev, size_t nq, std::vector<double> &s) const {
    dev.unnormalized_eval_range(0, size() - 1, 0, nq - 1, s, true);
  }

  // ---- dual tree (KernelDensityImpl.h:681-1110) ------------------------------------
  // The whole traversal runs on the GPU (bk_eval_bounded): queries in the
  // query tree's order are taken in blocks, references in tiles of the data
  // tree's order, and each (block, tile) pair is pruned or summed there.
  // The contract is the reference's: per query, lower <= f <= upper with
  // upper - lower <= max(abs_err, rel_err * lower), value() the midpoint.
  template <typename K>
  void tree_eval(KdtreeType &qtree, const K &kernel, FloatType rel_err, FloatType abs_err) const {
    if (qtree.empty()) return;
    require_data();
    const FloatType norm = (FloatType)kernel.normalization();
    Device<K> dev(points(), qtree.points());
    dev.kernel() = kernel;
    hip_detail::HostBuffer lo, up;
    dev.unnormalized_eval_bounded(0, qtree.size() - 1, (double)rel_err, (double)(abs_err / norm), lo, up);
    // bounds into the query points (large sets on host threads), then the
    // reference's precision report for the queries that miss the tolerance,
    // in query order
    auto &qp = qtree.points_;
    auto store = [&](size_t a0, size_t a1) {
      for (size_t i = a0; i < a1; ++i) {
        auto &a = qp[i].attributes();
        a.set_lower((FloatType)lo[i] * norm);
        a.set_upper((FloatType)up[i] * norm);
      }
    };
    const size_t nq = qp.size();
    const unsigned hw = std::thread::hardware_concurrency();
    const size_t t = nq < (size_t(1) << 18) || hw < 2 ? 1 : std::min<size_t>(hw, 16);
    if (t == 1) {
      store(0, nq);
    } else {
      std::vector<std::thread> workers;
      for (size_t c = 0; c < t; ++c) workers.emplace_back(store, nq * c / t, nq * (c + 1) / t);
      for (auto &w : workers) w.join();
    }
    for (size_t i = 0; i < nq; ++i) {
      const auto &a = qp[i].attributes();
      if (!within_tolerance(a.upper(), a.lower(), rel_err, abs_err))
        report_error(std::cerr, qp[i].point(), a.upper(), a.lower(), rel_err, abs_err);
    }
  }

  // the tolerance test of report_error
  static bool within_tolerance(FloatType upper, FloatType lower, FloatType rel_err, FloatType abs_err) {
    const FloatType gap = std::abs(upper - lower);
    if (gap <= abs_err) return true;
    return lower && std::abs((upper - lower) / lower) <= rel_err;
  }

  void report_error(std::ostream &os, const GeomPointType &p, FloatType upper, FloatType lower, FloatType rel_err,
                    FloatType abs_err) const {
    if (within_tolerance(upper, lower, rel_err, abs_err)) return;
    const FloatType gap = std::abs(upper - lower);
    os << std::setprecision(6) << (lower ? "Relative" : "Absolute") << " precision loss when querying " << p
       << ": " << std::endl
       << std::setprecision(15) << "\tlower:   " << lower << std::endl
       << "\tupper:   " << upper << std::endl
       << "\tabs_err: " << gap << " (c.f. " << abs_err << ")" << std::endl;
    if (lower) os << "\trel_err: " << gap / lower << " (c.f. " << rel_err << ")" << std::endl;
    os << std::endl;
  }

  KernelType kernel_{};
  KdtreeType data_tree_;
  std::vector<FloatType> cum_weights_;
  mutable SelfCache self_dev_;
};

// Least-squares cross-validation by numerical integration (2-D):
//   sum over the grid of f(g)^2 dx dy - 2 sum_i w_i f_{-i}(x_i),
// grid points start + i * delta with the end excluded (KernelDensityImpl.h:14-97).
// The grid values are direct GPU sums; the leave-one-out term excludes the
// self pair by index.
template <typename KT, typename FT, typename AT>
FT lsq_numint_cross_validate(const KernelDensity<2, KT, FT, AT> &kde, FT start_x, FT end_x, int steps_x,
                             FT start_y, FT end_y, int steps_y, FT rel_err, FT abs_err, int qtree_leaf_nmax,
                             size_t block_size) {
  (void)rel_err;
  (void)abs_err;
  (void)qtree_leaf_nmax;
  (void)block_size;
  using KDE = KernelDensity<2, KT, FT, AT>;
  using P = typename KDE::DataPointType;
  kde.require_data();
  const double dx = (end_x - start_x) / steps_x;
  const double dy = (end_y - start_y) / steps_y;
  std::vector<P> grid;
  grid.reserve((size_t)steps_x * steps_y);
  for (int j = 0; j < steps_y; ++j) {
    for (int i = 0; i < steps_x; ++i) {
      P p;
      p[0] = start_x + i * dx;
      p[1] = start_y + j * dy;
      grid.push_back(p);
    }
  }
  kde.direct_eval(grid);
  double sq = 0.0;
  for (const auto &g : grid) {
    const double v = g.attributes().value();
    sq += v * v * dx * dy;
  }
  // leave-one-out term: self sums without the pair (i, i)
  const std::vector<double> w = kde.weights();
  auto &dev = kde.self_device();
  const bk_kernel_desc d = kde.kernel().desc();
  std::vector<double> loo(w.size());
  hip_detail::check(bk_self_eval(dev.context(), &d, 1, loo.data(), 0));
  const double c = (double)kde.kernel().normalization();
  double llo = 0.0;
  for (size_t i = 0; i < w.size(); ++i) llo += w[i] * loo[i] * c;
  return (FT)(sq - 2.0 * llo);
}

}  // namespace bbrcit

#endif
