// HipDirectKde.h -- direct kernel-density sums on an MI355X, drop-in for the
// reference's CudaDirectKde<D, FloatT, KernelT> (include/CudaDirectKde.h:28-123).
//
// Same public members: construction over reference and query points (uploaded
// once), reference_size(), query_size(), kernel() (assigned by value before
// evaluations), unnormalized_eval / eval over closed index ranges with
// results[k - q_i] holding query k, and the same exceptions
// (std::out_of_range if r_j or q_j is past the end, std::invalid_argument if
// i > j).  HIP failures, which the reference ignores, raise
// std::runtime_error.  FloatT float selects the FP32 device path (the
// reference GPU callers' type), double the FP64 path.  block_size is accepted
// for source compatibility; launch shapes are chosen by the library.
//
// Extensions: unnormalized_eval_pairs evaluates many (reference range, query
// range) pairs in one launch; unnormalized_eval_bounded runs the whole
// dual-tree evaluation (bounds within a tolerance) on the device.
#ifndef BBRCIT_AMD_HIPDIRECTKDE_H_
#define BBRCIT_AMD_HIPDIRECTKDE_H_

#include <bbrcit_hip.h>

#include <algorithm>
#include <stdexcept>
#include <string>
#include <thread>
#include <type_traits>
#include <utility>
#include <vector>

#include "HostVector.h"
#include "Kernels/EpanechnikovKernel.h"

namespace bbrcit {

namespace hip_detail {

inline void check(int status) {
  if (status == BK_OK) return;
  const std::string msg = bk_last_error();
  if (status == BK_ERR_OUT_OF_RANGE) throw std::out_of_range(msg);
  if (status == BK_ERR_INVALID_ARG) throw std::invalid_argument(msg);
  throw std::runtime_error("bbrcit_hip: " + msg);
}

// A host buffer of doubles that resize() does not zero-fill.
using HostBuffer = HostVector<double>;

// Host points -> row-major coordinates, masses, abw (as double); large sets
// on several host threads (disjoint slices, so the result is the same).
template <int D, typename HostPointT, typename Vec>
void flatten(const std::vector<HostPointT> &pts, Vec &xy, Vec &m, Vec &a) {
  const size_t n = pts.size();
  xy.resize(n * D);
  m.resize(n);
  a.resize(n);
  auto slice = [&](size_t lo, size_t hi) {
    for (size_t i = lo; i < hi; ++i) {
      for (int d = 0; d < D; ++d) xy[i * D + d] = (double)pts[i][d];
      m[i] = (double)pts[i].attributes().mass();
      a[i] = (double)pts[i].attributes().abw();
    }
  };
  const unsigned hw = std::thread::hardware_concurrency();
  const size_t t = n < (size_t(1) << 18) || hw < 2 ? 1 : std::min<size_t>(hw, 16);
  if (t == 1) {
    slice(0, n);
    return;
  }
  std::vector<std::thread> workers;
  for (size_t c = 0; c < t; ++c) workers.emplace_back(slice, n * c / t, n * (c + 1) / t);
  for (auto &w : workers) w.join();
}

}  // namespace hip_detail

// Tag: the query set is the reference set itself (leave-one-out sums).
struct SelfQueries {};

template <int D, typename FloatT = double, typename KernelT = EpanechnikovKernel<D, FloatT>>
class HipDirectKde {
 public:
  using KernelType = KernelT;
  using FloatType = FloatT;

  HipDirectKde() = default;

  template <typename HostPointT>
  HipDirectKde(const std::vector<HostPointT> &ref_pts, const std::vector<HostPointT> &query_pts, int device = 0) {
    open(device);
    upload(BK_SET_REF, ref_pts);
    upload(BK_SET_QUERY, query_pts);
  }

  template <typename HostPointT>
  HipDirectKde(const std::vector<HostPointT> &ref_pts, SelfQueries, int device = 0) {
    open(device);
    upload(BK_SET_REF, ref_pts);
    hip_detail::check(bk_alias_queries_to_refs(ctx_));
    self_ = true;
  }

  HipDirectKde(const HipDirectKde &o) : kernel_(o.kernel_) {
    if (!o.ctx_) return;
    open(o.device_);
    copy_set(o, BK_SET_REF);
    if (o.self_) {
      hip_detail::check(bk_alias_queries_to_refs(ctx_));
      self_ = true;
    } else {
      copy_set(o, BK_SET_QUERY);
    }
  }
  HipDirectKde(HipDirectKde &&o) noexcept { swap(*this, o); }
  HipDirectKde &operator=(HipDirectKde o) {
    swap(*this, o);
    return *this;
  }
  ~HipDirectKde() {
    if (ctx_) bk_destroy(ctx_);
  }

  friend void swap(HipDirectKde &a, HipDirectKde &b) noexcept {
    using std::swap;
    swap(a.ctx_, b.ctx_);
    swap(a.device_, b.device_);
    swap(a.self_, b.self_);
    swap(a.kernel_, b.kernel_);
    swap(a.host_, b.host_);
  }

  size_t reference_size() const { return size_of(BK_SET_REF); }
  size_t query_size() const { return size_of(BK_SET_QUERY); }

  const KernelType &kernel() const { return kernel_; }
  KernelType &kernel() { return kernel_; }

  void unnormalized_eval(size_t r_i, size_t r_j, size_t q_i, size_t q_j, std::vector<FloatT> &results,
                         size_t block_size = 128) {
    (void)block_size;
    run(r_i, r_j, q_i, q_j, results, false);
  }

  // unnormalized_eval times kernel().normalization()
  void eval(size_t r_i, size_t r_j, size_t q_i, size_t q_j, std::vector<FloatT> &results,
            size_t block_size = 128) {
    (void)block_size;
    run(r_i, r_j, q_i, q_j, results, true);
  }

  // All pairs in one launch; out receives the results pair after pair.
  void unnormalized_eval_pairs(const std::vector<bk_range_pair> &pairs, std::vector<double> &out) {
    require_ctx();
    size_t total = 0;
    for (const auto &p : pairs) total += p.q_j - p.q_i + 1;
    out.resize(total);
    if (pairs.empty()) return;
    const bk_kernel_desc d = kernel_.desc();
    hip_detail::check(bk_eval_leaf_pairs(ctx_, &d, pairs.data(), pairs.size(), out.data(), 0));
  }

  // Unnormalised sums in double precision; with spatial, the library works
  // through the queries in Z-curve order on the device (BK_OPT_QUERY_ORDER),
  // the results still in query order.
  void unnormalized_eval_range(size_t r_i, size_t r_j, size_t q_i, size_t q_j, std::vector<double> &out,
                               bool spatial) {
    require_ctx();
    if (r_j >= reference_size() || q_j >= query_size())
      throw std::out_of_range("HipDirectKde<>: unnormalized_eval_range(): j_r and j_q should be less than the number "
                              "of reference and query points. ");
    if (r_i > r_j || q_i > q_j)
      throw std::invalid_argument("HipDirectKde<>: unnormalized_eval_range(): must have i_r <= j_r and i_q <= j_q. ");
    hip_detail::check(bk_set_option(ctx_, BK_OPT_QUERY_ORDER, spatial ? BK_QUERY_ORDER_SPATIAL : BK_QUERY_ORDER_GIVEN));
    out.resize(q_j - q_i + 1);
    const bk_kernel_desc d = kernel_.desc();
    hip_detail::check(bk_eval_range(ctx_, &d, r_i, r_j, q_i, q_j, out.data(), 0));
  }

  // Bounds of the unnormalised sums over every reference for queries
  // q_i..q_j: upper - lower <= max(abs_err, rel_err * lower)
  // (bk_eval_bounded).
  template <typename Vec>
  void unnormalized_eval_bounded(size_t q_i, size_t q_j, double rel_err, double abs_err, Vec &lower, Vec &upper) {
    require_ctx();
    if (q_j >= query_size() || reference_size() == 0)
      throw std::out_of_range("HipDirectKde<>: unnormalized_eval_bounded(): j_q should be less than the number of "
                              "query points. ");
    if (q_i > q_j) throw std::invalid_argument("HipDirectKde<>: unnormalized_eval_bounded(): must have i_q <= j_q. ");
    lower.resize(q_j - q_i + 1);
    upper.resize(q_j - q_i + 1);
    const bk_kernel_desc d = kernel_.desc();
    hip_detail::check(
        bk_eval_bounded(ctx_, &d, q_i, q_j, rel_err, abs_err, lower.data(), upper.data(), 0, nullptr));
  }

  bk_ctx *context() const { return ctx_; }

 private:
  static constexpr int kPrecision = std::is_same<FloatT, float>::value ? BK_PREC_F32 : BK_PREC_F64;

  // host copies kept only to implement copy construction
  struct HostSet {
    hip_detail::HostBuffer xy, m, a;
  };
  struct HostSets {
    HostSet ref, query;
  };

  void open(int device) {
    device_ = device;
    hip_detail::check(bk_create(&ctx_, device, kPrecision));
  }

  template <typename HostPointT>
  void upload(int set, const std::vector<HostPointT> &pts) {
    HostSet &h = set == BK_SET_REF ? host_.ref : host_.query;
    hip_detail::flatten<D>(pts, h.xy, h.m, h.a);
    hip_detail::check(bk_upload(ctx_, set, D, pts.size(), h.xy.data(), h.m.data(), h.a.data()));
  }

  void copy_set(const HipDirectKde &o, int set) {
    const HostSet &src = set == BK_SET_REF ? o.host_.ref : o.host_.query;
    HostSet &dst = set == BK_SET_REF ? host_.ref : host_.query;
    dst = src;
    hip_detail::check(bk_upload(ctx_, set, D, dst.m.size(), dst.xy.data(), dst.m.data(), dst.a.data()));
  }

  size_t size_of(int set) const {
    if (!ctx_) return 0;
    size_t n = 0;
    hip_detail::check(bk_size(ctx_, set, &n));
    return n;
  }

  void require_ctx() const {
    if (!ctx_) throw std::logic_error("HipDirectKde<>: no points (default-constructed object)");
  }

  void run(size_t r_i, size_t r_j, size_t q_i, size_t q_j, std::vector<FloatT> &results, bool normalize) {
    require_ctx();
    if (r_j >= reference_size() || q_j >= query_size())
      throw std::out_of_range(
          "HipDirectKde<>: unnormalized_eval(): j_r and j_q should be at most the number of reference and "
          "query points. ");
    if (r_i > r_j || q_i > q_j)
      throw std::invalid_argument("HipDirectKde<>: unnormalized_eval(): must have i_r <= j_r and i_q <= j_q. ");
    const size_t nq = q_j - q_i + 1;
    scratch_.resize(nq);
    const bk_kernel_desc d = kernel_.desc();
    hip_detail::check(bk_eval_range(ctx_, &d, r_i, r_j, q_i, q_j, scratch_.data(), 0));
    if (results.size() < nq) results.resize(nq);
    const double c = normalize ? (double)kernel_.normalization() : 1.0;
    for (size_t k = 0; k < nq; ++k) results[k] = (FloatT)(scratch_[k] * c);
  }

  bk_ctx *ctx_ = nullptr;
  int device_ = 0;
  bool self_ = false;
  KernelType kernel_{};
  HostSets host_;
  std::vector<double> scratch_;
};

}  // namespace bbrcit

#endif
