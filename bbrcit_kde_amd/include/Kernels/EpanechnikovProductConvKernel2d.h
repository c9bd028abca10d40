// EpanechnikovProductConvKernel2d.h -- self-convolution of the 2-D Epanechnikov product kernel.
// Interface of the reference's include/Kernels/EpanechnikovProductConvKernel2d.h: normalization(),
// unnormalized_eval(p, q[, a]) = U(p - q; hx a, hy a), bandwidth get/set.
#ifndef BBRCIT_AMD_EPANECHNIKOVPRODUCTCONVKERNEL2D_H_
#define BBRCIT_AMD_EPANECHNIKOVPRODUCTCONVKERNEL2D_H_

#include <random>
#include <vector>

#include "../KdeTraits.h"
#include "KernelDesc.h"
#include "KernelTraits.h"
#include "Utils.h"

namespace bbrcit {

template <typename T = double>
class EpanechnikovProductConvKernel2d {
 public:
  using FloatType = T;
  static constexpr int dim() { return 2; }

  EpanechnikovProductConvKernel2d() : hx_(T(1)), hy_(T(1)) {}
  EpanechnikovProductConvKernel2d(T hx, T hy) : hx_(hx), hy_(hy) {}

  T normalization() const { return T(0.36) / (hx_ * hy_); }

  template <typename PointT>
  T unnormalized_eval(const PointT &p, const PointT &q) const {
    return unnormalized_eval(p, q, ConstantTraits<T>::one());
  }

  template <typename PointT>
  T unnormalized_eval(const PointT &p, const PointT &q, T a) const {
    return kernel_detail::epan_conv_profile<T>((T(p[0]) - T(q[0])) / (hx_ * a)) *
           kernel_detail::epan_conv_profile<T>((T(p[1]) - T(q[1])) / (hy_ * a));
  }

  T hx() const { return hx_; }
  T hy() const { return hy_; }
  void set_hx(T h) { hx_ = h; }
  void set_hy(T h) { hy_ = h; }
  void set_bandwidths(T hx, T hy) {
    hx_ = hx;
    hy_ = hy;
  }

  bk_kernel_desc desc() const { return make_kernel_desc(BK_EPANECHNIKOV_PRODUCT_CONV_2D, 2, (double)hx_, (double)hy_); }

 private:
  T hx_, hy_;
};

}  // namespace bbrcit

#endif
