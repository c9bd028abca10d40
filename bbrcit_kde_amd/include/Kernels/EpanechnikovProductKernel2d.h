// EpanechnikovProductKernel2d.h -- 2-D Epanechnikov product kernel with bandwidths (hx, hy).
// Interface of the reference's include/Kernels/EpanechnikovProductKernel2d.h: normalization(),
// unnormalized_eval(p, q[, a]) = U(p - q; hx a, hy a), bandwidth get/set, simulate.
#ifndef BBRCIT_AMD_EPANECHNIKOVPRODUCTKERNEL2D_H_
#define BBRCIT_AMD_EPANECHNIKOVPRODUCTKERNEL2D_H_

#include <random>
#include <vector>

#include "../KdeTraits.h"
#include "KernelDesc.h"
#include "KernelTraits.h"
#include "Utils.h"

namespace bbrcit {

template <typename T = double>
class EpanechnikovProductKernel2d {
 public:
  using FloatType = T;
  static constexpr int dim() { return 2; }

  EpanechnikovProductKernel2d() : hx_(T(1)), hy_(T(1)) {}
  EpanechnikovProductKernel2d(T hx, T hy) : hx_(hx), hy_(hy) {}

  T normalization() const { return T(0.5625) / (hx_ * hy_); }

  template <typename PointT>
  T unnormalized_eval(const PointT &p, const PointT &q) const {
    return unnormalized_eval(p, q, ConstantTraits<T>::one());
  }

  template <typename PointT>
  T unnormalized_eval(const PointT &p, const PointT &q, T a) const {
    const T dx = T(p[0]) - T(q[0]), dy = T(p[1]) - T(q[1]);
    return kernel_detail::epan_profile<T>(dx * dx / (hx_ * hx_ * a * a)) *
           kernel_detail::epan_profile<T>(dy * dy / (hy_ * hy_ * a * a));
  }

  template <typename RNG>
  void simulate(RNG &e, std::vector<T> &p, T a = ConstantTraits<T>::one()) const;

  // The unit draws of simulate: per axis (x, then y) the Epanechnikov choice
  // of three uniform(-1, 1) draws, taken in the order v3, v2, v1 -- the
  // order in which g++ evaluates the reference's
  // epanechnikov_choice(d(e), d(e), d(e)) (arguments right to left).
  template <typename RNG>
  void draw_unit(RNG &e, T *z) const {
    std::uniform_real_distribution<T> u(T(-1), T(1));
    for (int i = 0; i < 2; ++i) {
      const T v3 = u(e), v2 = u(e), v1 = u(e);
      z[i] = epanechnikov_choice(v1, v2, v3);
    }
  }

  T hx() const { return hx_; }
  T hy() const { return hy_; }
  void set_hx(T h) { hx_ = h; }
  void set_hy(T h) { hy_ = h; }
  void set_bandwidths(T hx, T hy) {
    hx_ = hx;
    hy_ = hy;
  }

  bk_kernel_desc desc() const { return make_kernel_desc(BK_EPANECHNIKOV_PRODUCT_2D, 2, (double)hx_, (double)hy_); }

 private:
  T hx_, hy_;
};

// three uniform(-1, 1) draws per axis (epanechnikov_choice)
template <typename T>
template <typename RNG>
void EpanechnikovProductKernel2d<T>::simulate(RNG &e, std::vector<T> &p, T a) const {
  T z[2];
  draw_unit(e, z);
  p.resize(2);
  p[0] = a * hx_ * z[0];
  p[1] = a * hy_ * z[1];
}

}  // namespace bbrcit

#endif
