// GaussianKernel.h -- isotropic GaussianKernel in D dimensions with bandwidth h.
// Interface of the reference's include/Kernels/GaussianKernel.h: normalization(),
// unnormalized_eval(p, q[, a]) = U(p - q; h a), bandwidth get/set, simulate.
#ifndef BBRCIT_AMD_GAUSSIANKERNEL_H_
#define BBRCIT_AMD_GAUSSIANKERNEL_H_

#include <random>
#include <vector>

#include "../KdeTraits.h"
#include "KernelDesc.h"
#include "KernelTraits.h"
#include "Utils.h"

namespace bbrcit {

template <int D, typename T = double>
class GaussianKernel {
 public:
  using FloatType = T;
  static constexpr int dim() { return D; }

  GaussianKernel() : h_(T(1)) {}
  GaussianKernel(T bandwidth) : h_(bandwidth) {}

  T normalization() const { return kernel_detail::gaussian_norm<T>(D, h_); }

  template <typename PointT>
  T unnormalized_eval(const PointT &p, const PointT &q) const {
    return unnormalized_eval(p, q, ConstantTraits<T>::one());
  }

  // U(p - q; h a) from the squared distance over (h a)^2
  template <typename PointT>
  T unnormalized_eval(const PointT &p, const PointT &q, T a) const {
    T s = T(0);
    for (int i = 0; i < D; ++i) {
      const T d = T(p[i]) - T(q[i]);
      s += d * d;
    }
    return profile(s / (h_ * h_ * a * a));
  }

  template <typename RNG>
  void simulate(RNG &e, std::vector<T> &p, T a = ConstantTraits<T>::one()) const;

  // The unit draws of simulate (D standard normals; simulate scales them by
  // a * h).  One distribution object serves every call, as the reference's
  // function-local static does (GaussianKernel.h:146-151): for odd D a call
  // uses the second normal of the pair the previous call drew.
  template <typename RNG>
  void draw_unit(RNG &e, T *z) const {
    static std::normal_distribution<T> n(T(0), T(1));
    for (int i = 0; i < D; ++i) z[i] = n(e);
  }

  T bandwidth() const { return h_; }
  void set_bandwidth(T h) { h_ = h; }

  bk_kernel_desc desc() const { return make_kernel_desc(BK_GAUSSIAN, D, (double)h_, (double)h_); }

 private:
  static T profile(T arg) { return kernel_detail::gauss_profile<T>(arg, T(0.5)); }
  T h_;
};

// one N(0, (h a)^2) draw per coordinate
template <int D, typename T>
template <typename RNG>
void GaussianKernel<D, T>::simulate(RNG &e, std::vector<T> &p, T a) const {
  T z[D];
  draw_unit(e, z);
  p.resize(D);
  for (int i = 0; i < D; ++i) p[i] = a * h_ * z[i];
}

}  // namespace bbrcit

#endif
