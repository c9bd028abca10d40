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
#include <type_traits>
#include <utility>
#include <vector>

#include <bbrcit_hip.h>
#include "Attributes/AdaKdeAttributes.h"
#include "Attributes/PointWeights.h"
#include "DecoratedPoint.h"
#include "FloatUtils.h"
#include "HostVector.h"
#include "Rectangle.h"

namespace bbrcit {

template <int D, typename KT, typename FT, typename AT>
class KernelDensity;

namespace kdtree_detail {

// Attribute types whose merge the device reproduces exactly: their values
// merge by summation, minimum or maximum and are held as double.  pack()
// writes a point's values into the three groups, unpack() rebuilds the
// attributes from merged values.
template <typename AttrT>
struct DeviceMerge {
  static constexpr bool ok = false;
};

template <>
struct DeviceMerge<PointWeights<double>> {
  static constexpr bool ok = true;
  static constexpr int n_sum = 1, n_min = 0, n_max = 0;
  static void pack(const PointWeights<double> &a, double *s, double *, double *) { s[0] = a.weight(); }
  static PointWeights<double> unpack(const double *s, const double *, const double *) {
    return PointWeights<double>(s[0]);
  }
};

template <>
struct DeviceMerge<AdaKdeAttributes<double>> {
  static constexpr bool ok = true;
  static constexpr int n_sum = 2, n_min = 2, n_max = 2;
  static void pack(const AdaKdeAttributes<double> &a, double *s, double *mn, double *mx) {
    s[0] = a.weight();
    s[1] = a.mass();
    mn[0] = a.lower_abw();
    mn[1] = a.lower();
    mx[0] = a.upper_abw();
    mx[1] = a.upper();
  }
  static AdaKdeAttributes<double> unpack(const double *s, const double *mn, const double *mx) {
    AdaKdeAttributes<double> a(s[0]);
    a.set_mass(s[1]);
    a.set_lower_abw(mn[0]);
    a.set_lower(mn[1]);
    a.set_upper_abw(mx[0]);
    a.set_upper(mx[1]);
    return a;
  }
};

}  // namespace kdtree_detail

// Whether large trees (2^18 points and more) build their duplicate merge,
// root box and node attributes on the HIP device when there is one (default
// true); the host path gives the same tree.  For tests and comparisons.
inline bool &kdtree_device_build() {
  static bool on = true;
  return on;
}

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

  // f(start, end, split, cell, attributes, is_leaf) for every node, in
  // pre-order (an addition to the reference's interface, for inspection)
  template <typename F>
  void visit_nodes(F &&f) const {
    if (!root_) return;
    std::stack<const Node *> st;
    st.push(root_);
    while (!st.empty()) {
      const Node *n = st.top();
      st.pop();
      f(n->start_idx_, n->end_idx_, n->split_, n->bbox_, n->attr_, n->is_leaf());
      if (!n->is_leaf()) {
        st.push(n->right_);
        st.push(n->left_);
      }
    }
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

  // recompute node attributes below n from the points' attributes (the
  // whole tree of a large set with sibling subtrees on host threads)
  void refresh_node_attributes(Node *n) {
    if (!n) return;
    if (n == root_ && points_.size() >= kParallelMin) {
      int spawn_depth = 0;
      while ((1 << spawn_depth) < build_threads()) ++spawn_depth;
      refresh_parallel(n, spawn_depth);
      return;
    }
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
    if (use_device_build() && device_build()) return;
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

  using Merge = kdtree_detail::DeviceMerge<AttrT>;

  bool use_device_build() const {
    return std::is_same<FloatT, double>::value && kdtree_device_build() && points_.size() >= kParallelMin &&
           points_.size() <= 0x7fffffffULL && device_sort_available();
  }

  // body(c, a, b) for the slices c of [0, n), one per build_threads() thread
  // (an empty slice when n is smaller)
  template <typename F>
  static void on_threads_indexed(size_t n, F &&body) {
    const int t = build_threads();
    std::vector<std::future<void>> jobs;
    for (int c = 0; c < t; ++c) {
      const size_t a = n * (size_t)c / (size_t)t, b = n * (size_t)(c + 1) / (size_t)t;
      jobs.push_back(std::async(std::launch::async, [&body, c, a, b] { body(c, a, b); }));
    }
    for (auto &j : jobs) j.get();
  }

  // body(a, b) over [0, n) on up to build_threads() host threads
  template <typename F>
  static void on_threads(size_t n, F &&body) {
    const int t = build_threads();
    std::vector<std::future<void>> jobs;
    for (int c = 0; c < t; ++c) {
      const size_t a = n * (size_t)c / (size_t)t, b = n * (size_t)(c + 1) / (size_t)t;
      jobs.push_back(std::async(std::launch::async, [&body, a, b] { body(a, b); }));
    }
    for (auto &j : jobs) j.get();
  }

  // A point's coordinates and its index among the distinct points: the
  // partition runs on these records, whose comparisons (coordinate d) and
  // moves are those of the same std::nth_element calls on the points, so the
  // order is the same; the points move into it once at the end.
  struct Record {
    FloatT c[D];
    IndexType i;
  };

  // Merged attributes of the runs of duplicates (run u: positions
  // start[u] .. end - 1 of order), on the device for the attribute types it
  // merges (bk_run_merge), else with the host merge; dup lists the runs.
  template <typename IdxVec, typename DupVec>
  void merge_runs(const IdxVec &order, const IdxVec &start, size_t n, const DupVec &dup,
                  std::vector<AttrT> &merged) const {
    merged.resize(dup.size());
    if (dup.empty()) return;
    auto run_end = [&](size_t u) { return u + 1 < start.size() ? (size_t)start[u + 1] : n; };
    if (merge_runs_device(order, start, run_end, dup, merged, std::integral_constant<bool, Merge::ok>())) return;
    for (size_t k = 0; k < dup.size(); ++k) {
      const size_t u = dup[k];
      AttrT a = points_[(size_t)order[(size_t)start[u]]].attributes();
      for (size_t p = (size_t)start[u] + 1; p < run_end(u); ++p) a = merge(a, points_[(size_t)order[p]].attributes());
      merged[k] = a;
    }
  }

  template <typename IdxVec, typename DupVec, typename RunEnd>
  bool merge_runs_device(const IdxVec &, const IdxVec &, RunEnd &&, const DupVec &, std::vector<AttrT> &,
                         std::false_type) const {
    return false;
  }
  template <typename IdxVec, typename DupVec, typename RunEnd>
  bool merge_runs_device(const IdxVec &order, const IdxVec &start, RunEnd &&run_end, const DupVec &dup,
                         std::vector<AttrT> &merged, std::true_type) const {
    constexpr int ns = Merge::n_sum, nmn = Merge::n_min, nmx = Merge::n_max;
    std::vector<long long> off(dup.size() + 1, 0);
    for (size_t k = 0; k < dup.size(); ++k) off[k + 1] = off[k] + (long long)(run_end(dup[k]) - (size_t)start[dup[k]]);
    const size_t rows = (size_t)off.back();
    std::vector<double> s(rows * ns), mn(rows * nmn), mx(rows * nmx);
    double ds[ns > 0 ? ns : 1], dmn[nmn > 0 ? nmn : 1], dmx[nmx > 0 ? nmx : 1];
    for (size_t k = 0, r = 0; k < dup.size(); ++k)
      for (size_t p = (size_t)start[dup[k]]; p < run_end(dup[k]); ++p, ++r) {
        Merge::pack(points_[(size_t)order[p]].attributes(), ds, dmn, dmx);
        for (int j = 0; j < ns; ++j) s[r * ns + j] = ds[j];
        for (int j = 0; j < nmn; ++j) mn[r * nmn + j] = dmn[j];
        for (int j = 0; j < nmx; ++j) mx[r * nmx + j] = dmx[j];
      }
    std::vector<double> os(dup.size() * ns), omn(dup.size() * nmn), omx(dup.size() * nmx);
    if (bk_run_merge(0, dup.size(), off.data(), ns, ns ? s.data() : nullptr, nmn, nmn ? mn.data() : nullptr, nmx,
                     nmx ? mx.data() : nullptr, ns ? os.data() : nullptr, nmn ? omn.data() : nullptr,
                     nmx ? omx.data() : nullptr) != BK_OK)
      return false;
    for (size_t k = 0; k < dup.size(); ++k)
      merged[k] = Merge::unpack(ns ? &os[k * ns] : nullptr, nmn ? &omn[k * nmn] : nullptr,
                                nmx ? &omx[k * nmx] : nullptr);
    return true;
  }

  // Large sets with a HIP device: the sort, the runs of duplicates and the
  // root box in one device pass (bk_lexico_runs), the runs' attribute merge
  // on the device (bk_run_merge), the partition on host threads over
  // compact records, the points moved into tree order once, node attributes
  // on host threads.  false: nothing done.
  bool device_build() {
    // host scratch arrays are not zero-filled (HostVector): the threads that
    // fill them touch their pages
    const size_t n = points_.size();
    HostVector<double> xy(n * D);
    on_threads(n, [&](size_t a, size_t b) {
      for (size_t i = a; i < b; ++i)
        for (int d = 0; d < D; ++d) xy[i * D + d] = (double)points_[i][d];
    });
    size_t nu = 0;
    HostVector<long long> order(n), start(n);
    std::vector<double> box(2 * D);
    if (bk_lexico_runs(0, D, n, xy.data(), order.data(), &nu, start.data(), box.data()) != BK_OK) return false;
    start.resize(nu);
    // runs of more than one point (duplicates), in run order: counted per
    // slice on threads, then listed at each slice's offset
    auto run_len = [&](size_t u) { return (u + 1 < nu ? (size_t)start[u + 1] : n) - (size_t)start[u]; };
    const int t = build_threads();
    std::vector<size_t> slice_dups(t + 1, 0);
    on_threads_indexed(nu, [&](int c, size_t a, size_t b) {
      size_t k = 0;
      for (size_t u = a; u < b; ++u) k += run_len(u) > 1;
      slice_dups[c + 1] = k;
    });
    for (int c = 0; c < t; ++c) slice_dups[c + 1] += slice_dups[c];
    HostVector<size_t> dup(slice_dups[t]), dup_of(nu);
    on_threads_indexed(nu, [&](int c, size_t a, size_t b) {
      size_t k = slice_dups[c];
      for (size_t u = a; u < b; ++u) {
        if (run_len(u) > 1) {
          dup_of[u] = k;
          dup[k++] = u;
        } else {
          dup_of[u] = (size_t)-1;
        }
      }
    });
    std::vector<AttrT> merged;
    merge_runs(order, start, n, dup, merged);
    HostVector<Record> rec(nu);
    on_threads(nu, [&](size_t a, size_t b) {
      for (size_t u = a; u < b; ++u) {
        const size_t i = (size_t)order[(size_t)start[u]];
        for (int d = 0; d < D; ++d) rec[u].c[d] = (FloatT)xy[i * D + d];  // exact: FloatT is double here
        rec[u].i = u;
      }
    });
    RectangleType cell;
    for (int d = 0; d < D; ++d) cell.resize(d, typename RectangleType::EdgeType(box[2 * d], box[2 * d + 1]));
    int spawn_depth = 0;
    while ((1 << spawn_depth) < build_threads()) ++spawn_depth;
    Node *root = make_node_rec(rec, 0, nu - 1, 0, cell, spawn_depth);
    // the points into tree order: moved into raw storage on threads (k-th
    // slot from its source), then moved back over the first nu points
    using Slot = typename std::aligned_storage<sizeof(DataPointType), alignof(DataPointType)>::type;
    std::unique_ptr<Slot[]> raw(new Slot[nu]);
    DataPointType *tmp = reinterpret_cast<DataPointType *>(raw.get());
    on_threads(nu, [&](size_t a, size_t b) {
      for (size_t k = a; k < b; ++k) {
        const size_t u = rec[k].i;
        DataPointType *p = ::new (static_cast<void *>(tmp + k)) DataPointType(
            std::move(points_[(size_t)order[(size_t)start[u]]]));
        if (dup_of[u] != (size_t)-1) p->set_attributes(merged[dup_of[u]]);
      }
    });
    on_threads(nu, [&](size_t a, size_t b) {
      for (size_t k = a; k < b; ++k) {
        points_[k] = std::move(tmp[k]);
        tmp[k].~DataPointType();
      }
    });
    points_.resize(nu);
    root_ = root;
    refresh_parallel(root_, spawn_depth);
    return true;
  }

  // the partition of make_node over records; node attributes are set after
  // the points move into tree order
  Node *make_node_rec(HostVector<Record> &rec, IndexType i, IndexType j, int d, const RectangleType &cell,
                      int spawn_depth) {
    Node *n = new Node();
    n->start_idx_ = i;
    n->end_idx_ = j;
    n->bbox_ = cell;
    if (j - i + 1 <= (IndexType)leaf_nmax_) return n;
    const IndexType m = i + (j - i) / 2;
    std::nth_element(rec.begin() + i, rec.begin() + m, rec.begin() + j + 1,
                     [d](const Record &a, const Record &b) { return a.c[d] < b.c[d]; });
    n->split_ = rec[m].c[d];
    const RectangleType lo_cell = cell.lower_halfspace(d, n->split_), hi_cell = cell.upper_halfspace(d, n->split_);
    if (spawn_depth > 0) {
      auto left = std::async(std::launch::async, [this, &rec, i, m, d, &lo_cell, spawn_depth] {
        return make_node_rec(rec, i, m, (d + 1) % D, lo_cell, spawn_depth - 1);
      });
      n->right_ = make_node_rec(rec, m + 1, j, (d + 1) % D, hi_cell, spawn_depth - 1);
      n->left_ = left.get();
    } else {
      n->left_ = make_node_rec(rec, i, m, (d + 1) % D, lo_cell, 0);
      n->right_ = make_node_rec(rec, m + 1, j, (d + 1) % D, hi_cell, 0);
    }
    return n;
  }

  // node attributes below n (leaves: their points merged in index order;
  // internal nodes: left merged with right), sibling subtrees on their own
  // threads down to spawn_depth: the same values as the serial refresh
  void refresh_parallel(Node *n, int spawn_depth) {
    if (!n) return;
    if (n->is_leaf()) {
      n->attr_ = merged_attributes(n->start_idx_, n->end_idx_);
      return;
    }
    if (spawn_depth > 0) {
      auto left = std::async(std::launch::async, [this, n, spawn_depth] { refresh_parallel(n->left_, spawn_depth - 1); });
      refresh_parallel(n->right_, spawn_depth - 1);
      left.get();
    } else {
      refresh_parallel(n->left_, 0);
      refresh_parallel(n->right_, 0);
    }
    n->attr_ = merge(n->left_->attr_, n->right_->attr_);
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
