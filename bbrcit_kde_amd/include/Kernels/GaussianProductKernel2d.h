// GaussianProductKernel2d.h -- 2-D Gaussian product kernel with bandwidths (hx, hy).
// Interface of the reference's include/Kernels/GaussianProductKernel2d.h: normalization(),
// unnormalized_eval(p, q[, a]) = U(p - q; hx a, hy a), bandwidth get/set, simulate.
#ifndef BBRCIT_AMD_GAUSSIANPRODUCTKERNEL2D_H_
#define BBRCIT_AMD_GAUSSIANPRODUCTKERNEL2D_H_

#include <random>
#include <vector>

#include "../KdeTraits.h"
#include "KernelDesc.h"
#include "KernelTraits.h"
#include "Utils.h"

namespace bbrcit {

template <typename T = double>
class GaussianProductKernel2d {
 public:
  using FloatType = T;
  static constexpr int dim() { return 2; }

  GaussianProductKernel2d() : hx_(T(1)), hy_(T(1)) {}
  GaussianProductKernel2d(T hx, T hy) : hx_(hx), hy_(hy) {}

  T normalization() const { return T(0.15915494309189533577) / (hx_ * hy_); }

  template <typename PointT>
  T unnormalized_eval(const PointT &p, const PointT &q) const {
    return unnormalized_eval(p, q, ConstantTraits<T>::one());
  }

  template <typename PointT>
  T unnormalized_eval(const PointT &p, const PointT &q, T a) const {
    const T dx = T(p[0]) - T(q[0]), dy = T(p[1]) - T(q[1]);
    const T arg = dx * dx / (hx_ * hx_ * a * a) + dy * dy / (hy_ * hy_ * a * a);
    return kernel_detail::gauss_profile<T>(arg, T(0.5));
  }

  template <typename RNG>
  void simulate(RNG &e, std::vector<T> &p, T a = ConstantTraits<T>::one()) const;

  // The unit draws of simulate (two standard normals, x then y, from one
  // distribution object shared by every call as the reference's static).
  template <typename RNG>
  void draw_unit(RNG &e, T *z) const {
    static std::normal_distribution<T> n(T(0), T(1));
    z[0] = n(e);
    z[1] = n(e);
  }

  T hx() const { return hx_; }
  T hy() const { return hy_; }
  void set_hx(T h) { hx_ = h; }
  void set_hy(T h) { hy_ = h; }
  void set_bandwidths(T hx, T hy) {
    hx_ = hx;
    hy_ = hy;
  }

  bk_kernel_desc desc() const { return make_kernel_desc(BK_GAUSSIAN_PRODUCT_2D, 2, (double)hx_, (double)hy_); }

 private:
  T hx_, hy_;
};

// one N(0, (h a)^2) draw per axis
template <typename T>
template <typename RNG>
void GaussianProductKernel2d<T>::simulate(RNG &e, std::vector<T> &p, T a) const {
  T z[2];
  draw_unit(e, z);
  p.resize(2);
  p[0] = a * hx_ * z[0];
  p[1] = a * hy_ * z[1];
}

}  // namespace bbrcit

#endif
