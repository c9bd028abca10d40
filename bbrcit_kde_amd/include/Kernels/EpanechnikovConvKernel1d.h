// EpanechnikovConvKernel1d.h -- self-convolution of the 1-D Epanechnikov kernel.
// Interface of the reference's include/Kernels/EpanechnikovConvKernel1d.h: normalization(),
// unnormalized_eval(p, q[, a]) = U(p - q; h a), bandwidth get/set.
#ifndef BBRCIT_AMD_EPANECHNIKOVCONVKERNEL1D_H_
#define BBRCIT_AMD_EPANECHNIKOVCONVKERNEL1D_H_

#include <random>
#include <vector>

#include "../KdeTraits.h"
#include "KernelDesc.h"
#include "KernelTraits.h"
#include "Utils.h"

namespace bbrcit {

template <typename T = double>
class EpanechnikovConvKernel1d {
 public:
  using FloatType = T;
  static constexpr int dim() { return 1; }

  EpanechnikovConvKernel1d() : hx_(T(1)), hy_(T(1)) {}
  EpanechnikovConvKernel1d(T hx) : hx_(hx), hy_(hx) {}

  T normalization() const { return T(0.6) / hx_; }

  template <typename PointT>
  T unnormalized_eval(const PointT &p, const PointT &q) const {
    return unnormalized_eval(p, q, ConstantTraits<T>::one());
  }

  template <typename PointT>
  T unnormalized_eval(const PointT &p, const PointT &q, T a) const {
    return kernel_detail::epan_conv_profile<T>((T(p[0]) - T(q[0])) / (hx_ * a));
  }

  T hx() const { return hx_; }
  void set_hx(T h) {
    hx_ = h;
    hy_ = h;
  }

  bk_kernel_desc desc() const { return make_kernel_desc(BK_EPANECHNIKOV_CONV_1D, 1, (double)hx_, (double)hy_); }

 private:
  T hx_, hy_;
};

}  // namespace bbrcit

#endif
