// ConvKernelAssociator.h -- maps a kernel class to the class of its
// convolution kernel K*(x) = integral K(t) K(x - t) dt, as the reference's
// include/Kernels/ConvKernelAssociator.h:17-84 does (the same four pairs).
#ifndef BBRCIT_AMD_CONVKERNELASSOCIATOR_H_
#define BBRCIT_AMD_CONVKERNELASSOCIATOR_H_

#include "EpanechnikovConvKernel1d.h"
#include "EpanechnikovKernel.h"
#include "EpanechnikovProductConvKernel2d.h"
#include "EpanechnikovProductKernel2d.h"
#include "GaussianConvKernel1d.h"
#include "GaussianKernel.h"
#include "GaussianProductConvKernel2d.h"
#include "GaussianProductKernel2d.h"

namespace bbrcit {

template <typename KernelT>
class ConvKernelAssociator {};

template <typename T>
class ConvKernelAssociator<EpanechnikovKernel<1, T>> {
 public:
  using KernelType = EpanechnikovKernel<1, T>;
  using ConvKernelType = EpanechnikovConvKernel1d<T>;
  static ConvKernelType make_convolution_kernel(const KernelType &k) { return ConvKernelType(k.bandwidth()); }
};

template <typename T>
class ConvKernelAssociator<EpanechnikovProductKernel2d<T>> {
 public:
  using KernelType = EpanechnikovProductKernel2d<T>;
  using ConvKernelType = EpanechnikovProductConvKernel2d<T>;
  static ConvKernelType make_convolution_kernel(const KernelType &k) { return ConvKernelType(k.hx(), k.hy()); }
};

template <typename T>
class ConvKernelAssociator<GaussianKernel<1, T>> {
 public:
  using KernelType = GaussianKernel<1, T>;
  using ConvKernelType = GaussianConvKernel1d<T>;
  static ConvKernelType make_convolution_kernel(const KernelType &k) { return ConvKernelType(k.bandwidth()); }
};

template <typename T>
class ConvKernelAssociator<GaussianProductKernel2d<T>> {
 public:
  using KernelType = GaussianProductKernel2d<T>;
  using ConvKernelType = GaussianProductConvKernel2d<T>;
  static ConvKernelType make_convolution_kernel(const KernelType &k) { return ConvKernelType(k.hx(), k.hy()); }
};

}  // namespace bbrcit

#endif
