// Utils.h -- sampling helper for Epanechnikov kernels: of three uniform(-1,1)
// draws, returns v2 if |v3| is the largest, else v3 (the reference's
// include/Kernels/Utils.h:4-7; a standard Epanechnikov sampler).
#ifndef BBRCIT_AMD_KERNEL_UTILS_H_
#define BBRCIT_AMD_KERNEL_UTILS_H_

#include <cmath>

template <typename T>
inline T epanechnikov_choice(T v1, T v2, T v3) {
  return (std::abs(v3) >= std::abs(v2) && std::abs(v3) >= std::abs(v1)) ? v2 : v3;
}

#endif
