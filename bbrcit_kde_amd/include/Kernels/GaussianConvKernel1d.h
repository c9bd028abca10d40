// GaussianConvKernel1d.h -- self-convolution of the 1-D Gaussian kernel.
// Interface of the reference's include/Kernels/GaussianConvKernel1d.h: normalization(),
// unnormalized_eval(p, q[, a]) = U(p - q; h a), bandwidth get/set.
#ifndef BBRCIT_AMD_GAUSSIANCONVKERNEL1D_H_
#define BBRCIT_AMD_GAUSSIANCONVKERNEL1D_H_

#include <random>
#include <vector>

#include "../KdeTraits.h"
#include "KernelDesc.h"
#include "KernelTraits.h"
#include "Utils.h"

namespace bbrcit {

template <typename T = double>
class GaussianConvKernel1d {
 public:
  using FloatType = T;
  static constexpr int dim() { return 1; }

  GaussianConvKernel1d() : hx_(T(1)), hy_(T(1)) {}
  GaussianConvKernel1d(T hx) : hx_(hx), hy_(hx) {}

  T normalization() const { return T(0.28209479177387814347) / hx_; }

  template <typename PointT>
  T unnormalized_eval(const PointT &p, const PointT &q) const {
    return unnormalized_eval(p, q, ConstantTraits<T>::one());
  }

  template <typename PointT>
  T unnormalized_eval(const PointT &p, const PointT &q, T a) const {
    const T d = T(p[0]) - T(q[0]);
    return kernel_detail::gauss_profile<T>(d * d / (hx_ * hx_ * a * a), T(0.25));
  }

  T hx() const { return hx_; }
  void set_hx(T h) {
    hx_ = h;
    hy_ = h;
  }

  bk_kernel_desc desc() const { return make_kernel_desc(BK_GAUSSIAN_CONV_1D, 1, (double)hx_, (double)hy_); }

 private:
  T hx_, hy_;
};

}  // namespace bbrcit

#endif
