// GaussianProductConvKernel2d.h -- convolution kernel associated with the 2-D Gaussian product kernel (reference-compatible: exp(-arg/2), C = 1/(4 pi hx hy)).
// Interface of the reference's include/Kernels/GaussianProductConvKernel2d.h: normalization(),
// unnormalized_eval(p, q[, a]) = U(p - q; hx a, hy a), bandwidth get/set.
#ifndef BBRCIT_AMD_GAUSSIANPRODUCTCONVKERNEL2D_H_
#define BBRCIT_AMD_GAUSSIANPRODUCTCONVKERNEL2D_H_

#include <random>
#include <vector>

#include "../KdeTraits.h"
#include "KernelDesc.h"
#include "KernelTraits.h"
#include "Utils.h"

namespace bbrcit {

template <typename T = double>
class GaussianProductConvKernel2d {
 public:
  using FloatType = T;
  static constexpr int dim() { return 2; }

  GaussianProductConvKernel2d() : hx_(T(1)), hy_(T(1)) {}
  GaussianProductConvKernel2d(T hx, T hy) : hx_(hx), hy_(hy) {}

  T normalization() const { return T(0.07957747154594766788) / (hx_ * hy_); }

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

  T hx() const { return hx_; }
  T hy() const { return hy_; }
  void set_hx(T h) { hx_ = h; }
  void set_hy(T h) { hy_ = h; }
  void set_bandwidths(T hx, T hy) {
    hx_ = hx;
    hy_ = hy;
  }

  bk_kernel_desc desc() const { return make_kernel_desc(BK_GAUSSIAN_PRODUCT_CONV_2D, 2, (double)hx_, (double)hy_); }

 private:
  T hx_, hy_;
};

}  // namespace bbrcit

#endif
