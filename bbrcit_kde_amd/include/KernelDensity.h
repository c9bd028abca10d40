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
//   * the dual-tree evaluation prunes with the reference's bounds and
//     tolerance rules (681-1110).
// Device work goes through the C ABI (bbrcit_hip.h):
//   * every direct sum is one GPU evaluation;
//   * the dual tree's leaf-leaf base cases are collected during the host
//     traversal and evaluated together by bk_eval_leaf_pairs, instead of one
//     synchronous launch per leaf pair; pruning decisions therefore see the
//     node bounds without the pending base cases, which are looser but valid,
//     so the tree prunes no more than the reference does;
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
    for (auto &p : pts) {
      auto &a = p.attributes();
      a.set_weight(a.weight() / total);
      a.set_mass(a.weight() * std::pow(a.abw(), -D));
    }
    data_tree_ = KdtreeType(std::move(pts), leaf_nmax);
    cum_weights_.clear();
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

  template <typename K>
  void eval_all(Device<K> &dev, size_t nq, std::vector<double> &s) const {
    std::vector<bk_range_pair> one(1);
    one[0].r_i = 0;
    one[0].r_j = size() - 1;
    one[0].q_i = 0;
    one[0].q_j = nq - 1;
    dev.unnormalized_eval_pairs(one, s);
  }

  // ---- dual tree (KernelDensityImpl.h:681-1110) ------------------------------------
  // Base cases collected during the traversal, evaluated on the GPU in one
  // launch per batch and applied to the per-query bounds afterwards.
  template <typename K>
  struct LeafBatch {
    Device<K> *dev;
    KdtreeType *qtree;
    std::vector<bk_range_pair> pairs;
    std::vector<FloatT> dmass;  // mass of the reference leaf of each pair
    size_t pending = 0;
    static constexpr size_t kMaxPending = size_t(1) << 24;  // queries per flush

    void add(const TreeNodeType *Dn, const TreeNodeType *Qn) {
      bk_range_pair p;
      p.r_i = Dn->start_idx_;
      p.r_j = Dn->end_idx_;
      p.q_i = Qn->start_idx_;
      p.q_j = Qn->end_idx_;
      pairs.push_back(p);
      dmass.push_back(Dn->attr_.mass());
      pending += Qn->size();
      if (pending >= kMaxPending) flush();
    }

    // per query: upper += S - M_D, lower += S, lower <= upper
    // (the GPU base case of KernelDensityImpl.h:921-983)
    void flush() {
      if (pairs.empty()) return;
      std::vector<double> s;
      dev->unnormalized_eval_pairs(pairs, s);
      size_t off = 0;
      for (size_t k = 0; k < pairs.size(); ++k) {
        for (size_t i = pairs[k].q_i; i <= pairs[k].q_j; ++i, ++off) {
          auto &a = qtree->points_[i].attributes();
          FloatT up = a.upper() + (FloatT)s[off] - dmass[k];
          const FloatT lo = a.lower() + (FloatT)s[off];
          if (lo > up) up = lo;
          a.set_lower(lo);
          a.set_upper(up);
        }
      }
      pairs.clear();
      dmass.clear();
      pending = 0;
    }
  };

  template <typename K>
  void tree_eval(KdtreeType &qtree, const K &kernel, FloatType rel_err, FloatType abs_err) const {
    if (qtree.empty()) return;
    require_data();
    for (auto &q : qtree.points_) {
      q.attributes().set_lower(0);
      q.attributes().set_upper(data_tree_.root_->attr_.mass());
    }
    qtree.refresh_node_attributes(qtree.root_);
    const FloatType norm = (FloatType)kernel.normalization();
    Device<K> dev(points(), qtree.points());
    dev.kernel() = kernel;
    LeafBatch<K> batch{&dev, &qtree, {}, {}, 0};
    dual_tree(data_tree_.root_, qtree.root_, kernel, FloatType(1), FloatType(0), rel_err, abs_err / norm, qtree,
              batch);
    batch.flush();
    for (auto &q : qtree.points_) {
      q.attributes().set_lower(q.attributes().lower() * norm);
      q.attributes().set_upper(q.attributes().upper() * norm);
      report_error(std::cerr, q.point(), q.attributes().upper(), q.attributes().lower(), rel_err, abs_err);
    }
  }

  template <typename K>
  void dual_tree(const TreeNodeType *Dn, TreeNodeType *Qn, const K &kernel, FloatType du, FloatType dl,
                 FloatType rel_err, FloatType abs_err, KdtreeType &qtree, LeafBatch<K> &batch) const {
    FloatType du_new, dl_new;
    estimate_contributions(Dn, Qn->bbox_, kernel, du_new, dl_new);

    if (can_approximate(Dn, du_new, dl_new, du, dl, Qn->attr_.upper(), Qn->attr_.lower(), rel_err, abs_err)) {
      tighten_node(Dn, Qn, du_new, dl_new, du, dl);
      for (auto i = Qn->start_idx_; i <= Qn->end_idx_; ++i) {
        auto &a = qtree.points_[i].attributes();
        FloatType up = a.upper(), lo = a.lower();
        tighten_bounds(Dn, du_new, dl_new, FloatType(1), FloatType(0), up, lo);
        a.set_upper(up);
        a.set_lower(lo);
      }
      return;
    }

    if (Qn->is_leaf() && Dn->is_leaf()) {
      batch.add(Dn, Qn);
      return;
    }

    if (Qn->is_leaf()) {
      tighten_node(Dn, Qn, du_new, dl_new, du, dl);
      const TreeNodeType *near = Dn->left_, *far = Dn->right_;
      order_by_distance(&near, &far, Qn->bbox_);
      dual_tree(near, Qn, kernel, du_new, dl_new, rel_err, abs_err, qtree, batch);
      dual_tree(far, Qn, kernel, du_new, dl_new, rel_err, abs_err, qtree, batch);
      return;
    }

    tighten_node(Dn, Qn->left_, du_new, dl_new, du, dl);
    tighten_node(Dn, Qn->right_, du_new, dl_new, du, dl);
    if (Dn->is_leaf()) {
      dual_tree(Dn, Qn->left_, kernel, du_new, dl_new, rel_err, abs_err, qtree, batch);
      dual_tree(Dn, Qn->right_, kernel, du_new, dl_new, rel_err, abs_err, qtree, batch);
    } else {
      for (TreeNodeType *Qc : {Qn->left_, Qn->right_}) {
        const TreeNodeType *near = Dn->left_, *far = Dn->right_;
        order_by_distance(&near, &far, Qc->bbox_);
        dual_tree(near, Qc, kernel, du_new, dl_new, rel_err, abs_err, qtree, batch);
        dual_tree(far, Qc, kernel, du_new, dl_new, rel_err, abs_err, qtree, batch);
      }
    }
    Qn->attr_.set_lower(std::min(Qn->left_->attr_.lower(), Qn->right_->attr_.lower()));
    Qn->attr_.set_upper(std::max(Qn->left_->attr_.upper(), Qn->right_->attr_.upper()));
  }

  // bounds of the kernel over the node: its closest and farthest distance
  // to obj on every axis, at the node's largest / smallest abw
  template <typename ObjT, typename K>
  void estimate_contributions(const TreeNodeType *Dn, const ObjT &obj, const K &kernel, FloatType &du,
                              FloatType &dl) const {
    GeomPointType near, far;
    const GeomPointType origin;
    for (int i = 0; i < D; ++i) {
      near[i] = Dn->bbox_.min_dist(i, obj);
      far[i] = Dn->bbox_.max_dist(i, obj);
    }
    du = (FloatType)kernel.unnormalized_eval(near, origin, (KernelFloatType)Dn->attr_.upper_abw());
    dl = (FloatType)kernel.unnormalized_eval(far, origin, (KernelFloatType)Dn->attr_.lower_abw());
  }

  // add the node's new contribution bounds, removing the previous ones
  void tighten_bounds(const TreeNodeType *Dn, FloatType du_new, FloatType dl_new, FloatType du, FloatType dl,
                      FloatType &upper, FloatType &lower) const {
    lower += Dn->attr_.mass() * (dl_new - dl);
    upper += Dn->attr_.mass() * (du_new - du);
    if (lower > upper) upper = lower;
  }

  void tighten_node(const TreeNodeType *Dn, TreeNodeType *Qn, FloatType du_new, FloatType dl_new, FloatType du,
                    FloatType dl) const {
    FloatType up = Qn->attr_.upper(), lo = Qn->attr_.lower();
    tighten_bounds(Dn, du_new, dl_new, du, dl, up, lo);
    Qn->attr_.set_upper(up);
    Qn->attr_.set_lower(lo);
  }

  // prune when the absolute error stays within abs_err (exclusion or
  // approximation) or the relative error within rel_err (Deng & Moore,
  // Gray & Moore; KernelDensityImpl.h:1043-1075)
  bool can_approximate(const TreeNodeType *Dn, FloatType du_new, FloatType dl_new, FloatType du, FloatType dl,
                       FloatType upper, FloatType lower, FloatType rel_err, FloatType abs_err) const {
    const FloatType abs_tol = 2 * abs_err / data_tree_.size();
    if (std::abs(du_new) <= abs_tol) return true;
    if (std::abs(du_new - dl_new) <= abs_tol) return true;
    tighten_bounds(Dn, du_new, dl_new, du, dl, upper, lower);
    return std::abs(upper - lower) <= abs_err || std::abs(upper - lower) <= std::abs(lower) * rel_err;
  }

  template <typename ObjT>
  static void order_by_distance(const TreeNodeType **near, const TreeNodeType **far, const ObjT &obj) {
    if ((*near)->bbox_.min_dist(obj) > (*far)->bbox_.min_dist(obj)) std::swap(*near, *far);
  }

  void report_error(std::ostream &os, const GeomPointType &p, FloatType upper, FloatType lower, FloatType rel_err,
                    FloatType abs_err) const {
    const FloatType gap = std::abs(upper - lower);
    if (gap <= abs_err) return;
    if (lower && std::abs((upper - lower) / lower) <= rel_err) return;
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
