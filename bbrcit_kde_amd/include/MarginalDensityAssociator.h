// MarginalDensityAssociator.h -- 1-D marginals of a 2-D product-kernel density.
// Interface of the reference's include/MarginalDensityAssociator.h:13-128.
//
// The marginal of a KernelDensity<2, ProductKernel> along one axis is the 1-D
// density of that coordinate with the same weights and abw, the 1-D kernel of
// that axis (GaussianProductKernel2d -> GaussianKernel<1>,
// EpanechnikovProductKernel2d -> EpanechnikovKernel<1>) and mass w / abw
// (one power of abw: the other axis integrates out).  Its evaluations run on
// the GPU like any other KernelDensity<1, ...>.
#ifndef BBRCIT_AMD_MARGINALDENSITYASSOCIATOR_H_
#define BBRCIT_AMD_MARGINALDENSITYASSOCIATOR_H_

#include <cstddef>
#include <vector>

#include "Attributes/AdaKdeAttributes.h"
#include "DecoratedPoint.h"
#include "KernelDensity.h"
#include "Kernels/EpanechnikovKernel.h"
#include "Kernels/EpanechnikovProductKernel2d.h"
#include "Kernels/GaussianKernel.h"
#include "Kernels/GaussianProductKernel2d.h"

namespace bbrcit {

// coordinate dim_idx (0: x, 1: y) of a 2-D point with its weight and abw
// bounds; mass = weight / abw
template <typename FloatT, typename AttrT>
DecoratedPoint<1, AttrT, FloatT> marginalize_decorated_point2d(const DecoratedPoint<2, AttrT, FloatT> &pt2d,
                                                               size_t dim_idx) {
  DecoratedPoint<1, AttrT, FloatT> pt1d;
  pt1d[0] = pt2d[dim_idx];
  auto &a = pt1d.attributes();
  a.set_weight(pt2d.attributes().weight());
  a.set_lower_abw(pt2d.attributes().lower_abw());
  a.set_upper_abw(pt2d.attributes().upper_abw());
  a.set_mass(a.weight() / a.abw());
  return pt1d;
}

template <typename T>
class MarginalKernelAssociator {};

template <typename T>
class MarginalKernelAssociator<EpanechnikovProductKernel2d<T>> {
 public:
  using KernelType = EpanechnikovProductKernel2d<T>;
  using MarginalKernelType = EpanechnikovKernel<1, T>;
  static MarginalKernelType marginal_kernel(const KernelType &kern, size_t dim_idx) {
    return MarginalKernelType(dim_idx == 0 ? kern.hx() : kern.hy());
  }
};

template <typename T>
class MarginalKernelAssociator<GaussianProductKernel2d<T>> {
 public:
  using KernelType = GaussianProductKernel2d<T>;
  using MarginalKernelType = GaussianKernel<1, T>;
  static MarginalKernelType marginal_kernel(const KernelType &kern, size_t dim_idx) {
    return MarginalKernelType(dim_idx == 0 ? kern.hx() : kern.hy());
  }
};

template <typename KdeT, typename KernelT = typename KdeT::KernelType, typename FloatT = typename KdeT::FloatType,
          typename AttrT = typename KdeT::AttributeType>
class MarginalDensityAssociator {};

template <typename KernelT, typename FloatT, typename AttrT>
class MarginalDensityAssociator<KernelDensity<2, KernelT, FloatT, AttrT>, KernelT, FloatT, AttrT> {
 public:
  using KernelDensityType = KernelDensity<2, KernelT, FloatT, AttrT>;
  using MarginalKernelType = typename MarginalKernelAssociator<KernelT>::MarginalKernelType;
  using MarginalDensityType = KernelDensity<1, MarginalKernelType, FloatT, AttrT>;
  using Point1d = typename MarginalDensityType::DataPointType;

  static MarginalDensityType marginal_density(const KernelDensityType &kde, size_t dim_idx) {
    std::vector<Point1d> points;
    points.reserve(kde.points().size());
    for (const auto &p : kde.points()) points.push_back(marginalize_decorated_point2d(p, dim_idx));
    MarginalDensityType marginal(std::move(points), kde.data_tree().leaf_nmax());
    marginal.kernel() = MarginalKernelAssociator<KernelT>::marginal_kernel(kde.kernel(), dim_idx);
    return marginal;
  }
  static MarginalDensityType marginal_density_x(const KernelDensityType &kde) { return marginal_density(kde, 0); }
  static MarginalDensityType marginal_density_y(const KernelDensityType &kde) { return marginal_density(kde, 1); }
};

}  // namespace bbrcit

#endif  // BBRCIT_AMD_MARGINALDENSITYASSOCIATOR_H_
