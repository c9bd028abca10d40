// CvGrid.h -- batched bandwidth grid search over one or several 2-D
// estimators on an MI355X.
//
// The reference's grid_search tool (grid_search/grid_search.cc:366-402) loops
// over a list of (hx, hy) pairs, calling kernel().set_bandwidths(hx, hy) and
// one cross-validation per pair.  This header keeps that caller's vocabulary
// and moves the whole loop into one device pipeline (bk_cv_grid_*):
//   * make_crossed_vector_pairs and generate_2d_bandwidth_grid build the
//     reference's two grid layouts (grid_search/cv_utils_impl.h:15-25: v1
//     outer, v2 inner; cv_utils_impl.h:43-59: (steps+1)^2 points, end
//     included, y outer, x inner);
//   * lsq_convolution_cross_validate_grid / likelihood_cross_validate_grid
//     return the score of every pair of such a list, in the list's order, for
//     one estimator;
//   * CvGrid evaluates several estimators (datasets) against one
//     cross-product grid at once.
// Scores follow KernelDensity::lsq_convolution_cross_validate and
// ::likelihood_cross_validate (KernelDensityImpl.h:142-252) with the data's
// current masses and abw (adapted densities included).  Only the 2-D product
// kernels are supported (GaussianProductKernel2d, EpanechnikovProductKernel2d,
// the kernels with set_bandwidths).
#ifndef BBRCIT_AMD_CVGRID_H_
#define BBRCIT_AMD_CVGRID_H_

#include <bbrcit_hip.h>

#include <algorithm>
#include <cstddef>
#include <map>
#include <stdexcept>
#include <utility>
#include <vector>

#include "HipDirectKde.h"

namespace bbrcit {

inline std::vector<std::pair<double, double>> make_crossed_vector_pairs(const std::vector<double> &v1,
                                                                       const std::vector<double> &v2) {
  std::vector<std::pair<double, double>> out;
  out.reserve(v1.size() * v2.size());
  for (double a : v1)
    for (double b : v2) out.emplace_back(a, b);
  return out;
}

inline std::vector<std::pair<double, double>> generate_2d_bandwidth_grid(double start_x, double end_x, int steps_x,
                                                                        double start_y, double end_y,
                                                                        int steps_y) {
  std::vector<std::pair<double, double>> out;
  const double dx = (end_x - start_x) / steps_x, dy = (end_y - start_y) / steps_y;
  for (int j = 0; j <= steps_y; ++j)
    for (int i = 0; i <= steps_x; ++i) out.emplace_back(start_x + i * dx, start_y + j * dy);
  return out;
}

// Scores of one run: index (set * n_hx + ix) * n_hy + iy.
struct CvGridScores {
  size_t n_sets = 0, n_hx = 0, n_hy = 0;
  std::vector<double> lcv, lscv;
  double lcv_at(size_t set, size_t ix, size_t iy) const { return lcv[(set * n_hx + ix) * n_hy + iy]; }
  double lscv_at(size_t set, size_t ix, size_t iy) const { return lscv[(set * n_hx + ix) * n_hy + iy]; }
};

// One plan over several estimators of the same type (their points, weights
// and abw are copied to the device once).
template <typename KDE>
class CvGrid {
 public:
  static constexpr int kMaxAxis = 64;

  explicit CvGrid(const std::vector<const KDE *> &kdes, int device = 0) {
    if (kdes.empty()) throw std::invalid_argument("CvGrid: no estimators");
    const bk_kernel_desc d = kdes.front()->kernel().desc();
    if (d.kind != BK_GAUSSIAN_PRODUCT_2D && d.kind != BK_EPANECHNIKOV_PRODUCT_2D)
      throw std::invalid_argument("CvGrid: needs GaussianProductKernel2d or EpanechnikovProductKernel2d");
    kind_ = d.kind;
    std::vector<size_t> sizes;
    std::vector<double> xy, w, a;
    for (const KDE *k : kdes) {
      if (!k || k->size() == 0) throw std::invalid_argument("CvGrid: empty estimator");
      for (size_t i = 0; i < k->size(); ++i) {
        const auto &p = k->points()[i];
        xy.push_back((double)p[0]);
        xy.push_back((double)p[1]);
        w.push_back((double)p.attributes().weight());
        a.push_back((double)p.attributes().abw());
      }
      sizes.push_back(k->size());
    }
    hip_detail::check(bk_create(&ctx_, device, BK_PREC_F32));
    const int st = bk_cv_grid_create(ctx_, kind_, sizes.size(), sizes.data(), xy.data(), w.data(), a.data(), &plan_);
    if (st != BK_OK) {
      bk_destroy(ctx_);
      ctx_ = nullptr;
      hip_detail::check(st);
    }
    n_sets_ = sizes.size();
  }
  ~CvGrid() {
    if (plan_) bk_cv_grid_destroy(plan_);
    if (ctx_) bk_destroy(ctx_);
  }
  CvGrid(const CvGrid &) = delete;
  CvGrid &operator=(const CvGrid &) = delete;

  // Scores for the cross product hx x hy (any lengths; evaluated in blocks of
  // at most 64 x 64 bandwidths).
  CvGridScores run(const std::vector<double> &hx, const std::vector<double> &hy) {
    if (hx.empty() || hy.empty()) throw std::invalid_argument("CvGrid::run: empty bandwidth list");
    CvGridScores s;
    s.n_sets = n_sets_;
    s.n_hx = hx.size();
    s.n_hy = hy.size();
    s.lcv.assign(n_sets_ * hx.size() * hy.size(), 0.0);
    s.lscv.assign(s.lcv.size(), 0.0);
    for (size_t x0 = 0; x0 < hx.size(); x0 += kMaxAxis)
      for (size_t y0 = 0; y0 < hy.size(); y0 += kMaxAxis) {
        const size_t nx = std::min<size_t>(kMaxAxis, hx.size() - x0), ny = std::min<size_t>(kMaxAxis, hy.size() - y0);
        std::vector<double> l(n_sets_ * nx * ny), q(l.size());
        hip_detail::check(bk_cv_grid_run(plan_, nx, hx.data() + x0, ny, hy.data() + y0, l.data(), q.data()));
        for (size_t d = 0; d < n_sets_; ++d)
          for (size_t i = 0; i < nx; ++i)
            for (size_t j = 0; j < ny; ++j) {
              const size_t src = (d * nx + i) * ny + j, dst = (d * hx.size() + x0 + i) * hy.size() + y0 + j;
              s.lcv[dst] = l[src];
              s.lscv[dst] = q[src];
            }
      }
    return s;
  }

  void set_option(int option, int value) { hip_detail::check(bk_cv_grid_set_option(plan_, option, value)); }

  // Time every schedule of the pair kernel on whole runs of the blocks of
  // at most 64 x 64 bandwidths that run() evaluates for hx x hy, and record
  // the fastest per block shape in the autotuner cache (bk_cv_grid_tune).
  void tune(const std::vector<double> &hx, const std::vector<double> &hy) {
    if (hx.empty() || hy.empty()) throw std::invalid_argument("CvGrid::tune: empty bandwidth list");
    for (size_t x0 = 0; x0 < hx.size(); x0 += kMaxAxis)
      for (size_t y0 = 0; y0 < hy.size(); y0 += kMaxAxis) {
        const size_t nx = std::min<size_t>(kMaxAxis, hx.size() - x0), ny = std::min<size_t>(kMaxAxis, hy.size() - y0);
        hip_detail::check(bk_cv_grid_tune(plan_, nx, hx.data() + x0, ny, hy.data() + y0, nullptr, nullptr));
      }
  }

 private:
  bk_ctx *ctx_ = nullptr;
  bk_cv_grid *plan_ = nullptr;
  int kind_ = 0;
  size_t n_sets_ = 0;
};

namespace cv_grid_detail {

// Scores of a list of (hx, hy) pairs, evaluated as the cross product of the
// distinct values.
template <typename KDE>
CvGridScores run_pairs(const KDE &kde, const std::vector<std::pair<double, double>> &pairs, std::vector<size_t> &ix,
                       std::vector<size_t> &iy, std::vector<double> &hx, std::vector<double> &hy) {
  std::map<double, size_t> mx, my;
  for (const auto &p : pairs) {
    mx.emplace(p.first, 0);
    my.emplace(p.second, 0);
  }
  for (auto &e : mx) {
    e.second = hx.size();
    hx.push_back(e.first);
  }
  for (auto &e : my) {
    e.second = hy.size();
    hy.push_back(e.first);
  }
  for (const auto &p : pairs) {
    ix.push_back(mx[p.first]);
    iy.push_back(my[p.second]);
  }
  CvGrid<KDE> grid(std::vector<const KDE *>{&kde});
  return grid.run(hx, hy);
}

}  // namespace cv_grid_detail

template <typename KDE>
std::vector<double> lsq_convolution_cross_validate_grid(const KDE &kde,
                                                        const std::vector<std::pair<double, double>> &pairs) {
  std::vector<size_t> ix, iy;
  std::vector<double> hx, hy;
  const CvGridScores s = cv_grid_detail::run_pairs(kde, pairs, ix, iy, hx, hy);
  std::vector<double> out(pairs.size());
  for (size_t k = 0; k < pairs.size(); ++k) out[k] = s.lscv_at(0, ix[k], iy[k]);
  return out;
}

template <typename KDE>
std::vector<double> likelihood_cross_validate_grid(const KDE &kde,
                                                   const std::vector<std::pair<double, double>> &pairs) {
  std::vector<size_t> ix, iy;
  std::vector<double> hx, hy;
  const CvGridScores s = cv_grid_detail::run_pairs(kde, pairs, ix, iy, hx, hy);
  std::vector<double> out(pairs.size());
  for (size_t k = 0; k < pairs.size(); ++k) out[k] = s.lcv_at(0, ix[k], iy[k]);
  return out;
}

}  // namespace bbrcit

#endif  // BBRCIT_AMD_CVGRID_H_
