// Kdtree.h -- D-dimensional kd-tree over decorated points.
//
// Interface of the reference's include/Kdtree.h:31-156.  Construction keeps
// the reference's steps, because the order of points_ is what the device
// sums and every index range refer to (Kdtree.h:337-471):
//   1. sort the points lexicographically (ExactLexicoLess) and merge exact
//      duplicates into one point whose attributes are merged;
//   2. split [i, j] at the lower median m = i + (j - i)/2 of the coordinate
//      d (std::nth_element), d cycling through the axes, until a range holds
//      at most leaf_nmax points;
//   3. each node stores the partition cell of its range (the root's box is
//      the points' bounding box, grown from the reference's initial values:
//      lowest edge numeric_limits::max(), highest edge numeric_limits::min())
//      and the merge of its points' attributes.
// Large sets (kParallelMin points and more) are built without changing the
// result: the sort runs on the HIP device when there is one (a stable radix
// sort, bk_lexico_order) or else as sorted chunks merged pairwise on several
// host threads (the same order, since ExactLexicoLess is a strict order on
// distinct points; only the order among exact duplicates, which are merged
// next, may differ), and sibling subtrees, whose ranges are disjoint, are
// partitioned concurrently with the same std::nth_element calls.
#ifndef BBRCIT_AMD_KDTREE_H_
#define BBRCIT_AMD_KDTREE_H_

#include <algorithm>
#include <future>
#include <iostream>
#include <iterator>
#include <limits>
#include <memory>
#include <stack>
#include <thread>
#include <utility>
#include <vector>

#include <bbrcit_hip.h>
#include "Attributes/PointWeights.h"
#include "DecoratedPoint.h"
#include "FloatUtils.h"
#include "Rectangle.h"

namespace bbrcit {

template <int D, typename KT, typename FT, typename AT>
class KernelDensity;

template <int D, typename AttrT = PointWeights<int>, typename FloatT = double>
class Kdtree {
 public:
  using DataPointType = DecoratedPoint<D, AttrT, FloatT>;
  using RectangleType = Rectangle<D, FloatT>;
  using AttributesType = AttrT;
  using FloatType = FloatT;
  static constexpr int dim() { return D; }

 protected:
  using IndexType = typename std::vector<DataPointType>::size_type;
  template <int DIM, typename KT, typename FT, typename AT>
  friend class KernelDensity;

 public:
  Kdtree() : root_(nullptr), leaf_nmax_(2) {}
  Kdtree(const std::vector<DataPointType> &data, int leaf_nmax = 2)
      : points_(data), root_(nullptr), leaf_nmax_(leaf_nmax) { build(); }
  Kdtree(std::vector<DataPointType> &&data, int leaf_nmax = 2)
      : points_(std::move(data)), root_(nullptr), leaf_nmax_(leaf_nmax) { build(); }

  Kdtree(const Kdtree &o) : points_(o.points_), root_(copy_nodes(o.root_)), leaf_nmax_(o.leaf_nmax_) {}
  Kdtree(Kdtree &&o) noexcept : points_(std::move(o.points_)), root_(o.root_), leaf_nmax_(o.leaf_nmax_) {
    o.root_ = nullptr;
  }
  Kdtree &operator=(Kdtree o) {
    swap(*this, o);
    return *this;
  }
  virtual ~Kdtree() { free_nodes(root_); }

  friend void swap(Kdtree &a, Kdtree &b) {
    using std::swap;
    swap(a.points_, b.points_);
    swap(a.root_, b.root_);
    swap(a.leaf_nmax_, b.leaf_nmax_);
  }

  bool empty() const { return root_ == nullptr; }
  IndexType size() const { return root_ ? root_->size() : 0; }
  int leaf_nmax() const { return leaf_nmax_; }
  const std::vector<DataPointType> &points() const { return points_; }

  void print_points(std::ostream &os) const {
    for (const auto &p : points_) os << p << std::endl;
  }

  // cells at depth d, and leaf cells shallower than d
  void report_partitions(int d, std::vector<RectangleType> &out) const { collect_cells(root_, d, out); }
  void print_partitions(int d, std::ostream &os) const {
    std::vector<RectangleType> cells;
    report_partitions(d, cells);
    for (const auto &c : cells) os << c << std::endl;
  }

  void range_search(const RectangleType &q, std::vector<DataPointType> &out) const {
    std::vector<IndexType> idx;
    collect_range(root_, q, idx);
    for (IndexType i : idx) out.emplace_back(points_[i]);
  }
  void print_range_search(const RectangleType &q, std::ostream &os) const {
    std::vector<IndexType> idx;
    collect_range(root_, q, idx);
    for (IndexType i : idx) os << points_[i] << std::endl;
  }

  // index ranges [start, end] of the leaves, left to right
  void report_leaves(std::vector<std::pair<IndexType, IndexType>> &out) const {
    if (!root_) return;
    std::stack<const Node *> s;
    s.push(root_);
    while (!s.empty()) {
      const Node *n = s.top();
      s.pop();
      if (n->is_leaf()) {
        out.emplace_back(n->start_idx_, n->end_idx_);
      } else {
        s.push(n->right_);
        s.push(n->left_);
      }
    }
  }

 protected:
  struct Node {
    FloatType split_ = FloatType();
    Node *left_ = nullptr, *right_ = nullptr;
    IndexType start_idx_ = 0, end_idx_ = 0;  // closed range of points_
    AttributesType attr_ = AttributesType();
    RectangleType bbox_;
    bool is_leaf() const { return left_ == nullptr && right_ == nullptr; }
    IndexType size() const { return end_idx_ - start_idx_ + 1; }
  };

  std::vector<DataPointType> points_;
  Node *root_;
  int leaf_nmax_;

  // recompute node attributes below n from the points' attributes
  void refresh_node_attributes(Node *n) {
    if (!n) return;
    if (n->is_leaf()) {
      n->attr_ = merged_attributes(n->start_idx_, n->end_idx_);
    } else {
      refresh_node_attributes(n->left_);
      refresh_node_attributes(n->right_);
      n->attr_ = merge(n->left_->attr_, n->right_->attr_);
    }
  }

 private:
  static constexpr size_t kParallelMin = size_t(1) << 18;

  static int build_threads() {
    const unsigned t = std::thread::hardware_concurrency();
    return t == 0 ? 1 : (int)std::min(t, 16u);
  }

  void build() {
    merge_duplicates();
    if (points_.empty()) return;
    int spawn_depth = 0;
    if (points_.size() >= kParallelMin)
      while ((1 << spawn_depth) < build_threads()) ++spawn_depth;
    root_ = make_node(0, points_.size() - 1, 0, bounding_box(), spawn_depth);
  }

  // true when HIP device 0 runs bk_lexico_order (tried once per process)
  static bool device_sort_available() {
    static const bool ok = [] {
      double x = 0.0;
      long long o = 0;
      return bk_lexico_order(0, 1, 1, &x, &o) == BK_OK;
    }();
    return ok;
  }

  // Large sets on a machine with a HIP device: the order from a device radix
  // sort (bk_lexico_order: ExactLexicoLess, duplicates in input order), then
  // the points moved into it on several threads.  false: nothing done.
  bool device_sort() {
    const size_t n = points_.size();
    if (n < kParallelMin || n > 0x7fffffffULL || !device_sort_available()) return false;
    const int t = build_threads();
    auto on_threads = [t, n](auto &&body) {
      std::vector<std::future<void>> jobs;
      for (int c = 0; c < t; ++c) {
        const size_t a = n * (size_t)c / (size_t)t, b = n * (size_t)(c + 1) / (size_t)t;
        jobs.push_back(std::async(std::launch::async, [&body, a, b] { body(a, b); }));
      }
      for (auto &j : jobs) j.get();
    };
    std::vector<double> coords(n * D);
    on_threads([&](size_t a, size_t b) {
      for (size_t i = a; i < b; ++i)
        for (int d = 0; d < D; ++d) coords[i * D + d] = (double)points_[i][d];
    });
    std::vector<long long> order(n);
    if (bk_lexico_order(0, D, n, coords.data(), order.data()) != BK_OK) return false;
    std::vector<DataPointType> sorted(n);
    on_threads([&](size_t a, size_t b) {
      for (size_t k = a; k < b; ++k) sorted[k] = std::move(points_[(size_t)order[k]]);
    });
    points_.swap(sorted);
    return true;
  }

  // chunks sorted on their own threads, then merged pairwise
  void sort_points() {
    if (device_sort()) return;
    const int t = build_threads();
    if (points_.size() < kParallelMin || t < 2) {
      std::sort(points_.begin(), points_.end(), ExactLexicoLess<DataPointType>);
      return;
    }
    const size_t n = points_.size();
    std::vector<size_t> bounds;
    for (int c = 0; c <= t; ++c) bounds.push_back(n * (size_t)c / (size_t)t);
    {
      std::vector<std::future<void>> jobs;
      for (int c = 0; c < t; ++c)
        jobs.push_back(std::async(std::launch::async, [this, &bounds, c] {
          std::sort(points_.begin() + bounds[c], points_.begin() + bounds[c + 1], ExactLexicoLess<DataPointType>);
        }));
      for (auto &j : jobs) j.get();
    }
    std::vector<DataPointType> buf(n);
    std::vector<DataPointType> *src = &points_, *dst = &buf;
    while (bounds.size() > 2) {
      std::vector<size_t> next;
      std::vector<std::future<void>> jobs;
      for (size_t c = 0; c + 1 < bounds.size(); c += 2) {
        const size_t a = bounds[c], b = bounds[c + 1], e = (c + 2 < bounds.size()) ? bounds[c + 2] : b;
        next.push_back(a);
        jobs.push_back(std::async(std::launch::async, [src, dst, a, b, e] {
          std::merge(std::make_move_iterator(src->begin() + a), std::make_move_iterator(src->begin() + b),
                     std::make_move_iterator(src->begin() + b), std::make_move_iterator(src->begin() + e),
                     dst->begin() + a, ExactLexicoLess<DataPointType>);
        }));
      }
      next.push_back(bounds.back());
      for (auto &j : jobs) j.get();
      bounds.swap(next);
      std::swap(src, dst);
    }
    if (src != &points_) points_.swap(buf);
  }

  void merge_duplicates() {
    if (points_.empty()) return;
    sort_points();
    size_t last = 0;
    for (size_t k = 1; k < points_.size(); ++k) {
      if (ExactEqual(points_[last], points_[k])) {
        points_[last].set_attributes(merge(points_[last].attributes(), points_[k].attributes()));
      } else {
        ++last;
        swap(points_[last], points_[k]);
      }
    }
    points_.resize(last + 1);
  }

  RectangleType bounding_box() const {
    std::vector<FloatType> lo(D, std::numeric_limits<FloatType>::max());
    std::vector<FloatType> hi(D, std::numeric_limits<FloatType>::min());
    for (const auto &p : points_) {
      for (int d = 0; d < D; ++d) {
        lo[d] = std::min(p[d], lo[d]);
        hi[d] = std::max(p[d], hi[d]);
      }
    }
    RectangleType r;
    for (int d = 0; d < D; ++d) r.resize(d, typename RectangleType::EdgeType(lo[d], hi[d]));
    return r;
  }

  AttributesType merged_attributes(IndexType i, IndexType j) const {
    AttributesType a = points_[i].attributes();
    for (IndexType k = i + 1; k <= j; ++k) a.merge(points_[k].attributes());
    return a;
  }

  // spawn_depth > 0: the left subtree is built on a thread of its own
  Node *make_node(IndexType i, IndexType j, int d, const RectangleType &cell, int spawn_depth = 0) {
    Node *n = new Node();
    n->start_idx_ = i;
    n->end_idx_ = j;
    n->bbox_ = cell;
    if (j - i + 1 <= (IndexType)leaf_nmax_) {
      n->attr_ = merged_attributes(i, j);
      return n;
    }
    const IndexType m = i + (j - i) / 2;
    std::nth_element(points_.begin() + i, points_.begin() + m, points_.begin() + j + 1,
                     [d](const DataPointType &a, const DataPointType &b) { return a[d] < b[d]; });
    n->split_ = points_[m][d];
    const RectangleType lo_cell = cell.lower_halfspace(d, n->split_), hi_cell = cell.upper_halfspace(d, n->split_);
    if (spawn_depth > 0) {
      auto left = std::async(std::launch::async, [this, i, m, d, &lo_cell, spawn_depth] {
        return make_node(i, m, (d + 1) % D, lo_cell, spawn_depth - 1);
      });
      n->right_ = make_node(m + 1, j, (d + 1) % D, hi_cell, spawn_depth - 1);
      n->left_ = left.get();
    } else {
      n->left_ = make_node(i, m, (d + 1) % D, lo_cell);
      n->right_ = make_node(m + 1, j, (d + 1) % D, hi_cell);
    }
    n->attr_ = merge(n->left_->attr_, n->right_->attr_);
    return n;
  }

  static Node *copy_nodes(const Node *n) {
    if (!n) return nullptr;
    Node *c = new Node(*n);
    c->left_ = copy_nodes(n->left_);
    c->right_ = copy_nodes(n->right_);
    return c;
  }

  static void free_nodes(Node *n) {
    if (!n) return;
    free_nodes(n->left_);
    free_nodes(n->right_);
    delete n;
  }

  void collect_cells(const Node *n, int depth, std::vector<RectangleType> &out) const {
    if (!n) return;
    if (depth == 0 || n->is_leaf()) {
      out.push_back(n->bbox_);
      return;
    }
    collect_cells(n->left_, depth - 1, out);
    collect_cells(n->right_, depth - 1, out);
  }

  void collect_range(const Node *n, const RectangleType &q, std::vector<IndexType> &out) const {
    if (!n) return;
    if (n->is_leaf()) {
      for (IndexType i = n->start_idx_; i <= n->end_idx_; ++i)
        if (q.contains(points_[i])) out.push_back(i);
      return;
    }
    for (const Node *c : {n->left_, n->right_}) {
      if (q.contains(c->bbox_)) {
        for (IndexType i = c->start_idx_; i <= c->end_idx_; ++i) out.push_back(i);
      } else {
        collect_range(c, q, out);
      }
    }
  }
};

}  // namespace bbrcit

#endif
