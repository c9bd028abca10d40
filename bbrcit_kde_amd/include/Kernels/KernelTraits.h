// KernelTraits.h -- normalisations C(h) and profile functions of the kernel
// families, in the kernel's float type T.
// Constants and formulas of the reference's include/Kernels/KernelTraits.h:
//   Epanechnikov    C = (D+2)/2 * Gamma(1+D/2) / pi^(D/2) / h^D,  U = max(1-arg, 0)   (19-88)
//   Gaussian        C = (2 pi)^(-D/2) / h^D,                      U = exp(-arg/2)      (97-164)
//   Gaussian x2     C = 1/(2 pi hx hy)                                                (170-213)
//   Epan. x2        C = (9/16)/(hx hy)                                                (218-257)
//   Epan. conv 1-D  C = 0.6/h,  U = (2-|u|)^3 (u^2+6|u|+4)/32, |u| < 2                (263-308)
//   Epan. conv x2   C = 0.36/(hx hy)                                                  (313-358)
//   Gauss. conv 1-D C = 1/(sqrt(4 pi) h), U = exp(-arg/4)                             (364-403)
//   Gauss. conv x2  C = 1/(4 pi hx hy), U = exp(-arg/2) for T = float, double         (408-447)
// The last line keeps the reference's float/double specialisations
// (exp(-0.5 arg), KernelTraits.h:432, 445); the mathematical convolution of
// two Gaussian product kernels would be exp(-0.25 arg).  The 1-D double
// Gaussian normalisation takes its bandwidth as float, as the reference's
// GaussianTraits<1,double>::normalization(float) does (KernelTraits.h:143).
#ifndef BBRCIT_AMD_KERNELTRAITS_H_
#define BBRCIT_AMD_KERNELTRAITS_H_

#include <algorithm>
#include <cmath>
#include <type_traits>

namespace bbrcit {
namespace kernel_detail {

constexpr double kPi = 3.14159265358979323846;

template <typename T>
T epanechnikov_norm(int D, T h) {
  if (D == 1) return T(0.75) / h;
  if (D == 2) return T(0.63661977236758134308) / (h * h);
  const double c = 0.5 * (D + 2) / (std::pow(kPi, D / 2.0) / std::tgamma(1.0 + D / 2.0));
  return T(c / std::pow((double)h, D));
}

template <typename T>
T gaussian_norm(int D, T h) {
  if (D == 1) {
    const float hf = (float)h;  // reference: normalization(float h) for T = double
    return T(0.39894228040143267794) / (std::is_same<T, double>::value ? T(hf) : h);
  }
  if (D == 2) return T(0.15915494309189533577) / (h * h);
  return T(std::pow(2.0 * kPi, -D / 2.0) / std::pow((double)h, D));
}

template <typename T>
T epan_profile(T arg) {
  return std::max(T(1) - arg, T(0));
}

template <typename T>
T gauss_profile(T arg, T c) {
  return std::exp(-c * arg);
}

template <typename T>
T epan_conv_profile(T u) {
  const T a = std::abs(u);
  const T c = std::max(T(2) - a, T(0));
  return T(0.03125) * c * c * c * (a * a + T(6) * a + T(4));
}

}  // namespace kernel_detail
}  // namespace bbrcit

#endif
