// DecoratedPoint.h -- a Point<> with attributes attached.
// Interface of the reference's include/DecoratedPoint.h:244-315.
#ifndef BBRCIT_AMD_DECORATEDPOINT_H_
#define BBRCIT_AMD_DECORATEDPOINT_H_

#include <iostream>
#include <utility>

#include "Attributes/PointWeights.h"
#include "Point.h"

namespace bbrcit {

template <int Dim, typename AttrT = PointWeights<int>, typename FloatT = double>
class DecoratedPoint {
 public:
  using PointType = Point<Dim, FloatT>;
  using AttributesType = AttrT;
  using FloatType = FloatT;
  static constexpr int dim() { return Dim; }

  DecoratedPoint() = default;
  DecoratedPoint(const PointType &p) : p_(p), a_() {}
  DecoratedPoint(PointType &&p) : p_(std::move(p)), a_() {}
  DecoratedPoint(const PointType &p, const AttrT &a) : p_(p), a_(a) {}

  const AttrT &attributes() const { return a_; }
  AttrT &attributes() { return a_; }
  const PointType &point() const { return p_; }
  PointType &point() { return p_; }
  void set_attributes(const AttrT &a) { a_ = a; }
  void set_point(const PointType &p) { p_ = p; }

  const FloatType &operator[](size_t i) const { return p_[(int)i]; }
  FloatType &operator[](size_t i) { return p_[(int)i]; }

  friend void swap(DecoratedPoint &a, DecoratedPoint &b) {
    using std::swap;
    swap(a.p_, b.p_);
    swap(a.a_, b.a_);
  }

  friend std::istream &operator>>(std::istream &is, DecoratedPoint &d) {
    is >> d.p_;
    is >> d.a_;
    if (!is) d = DecoratedPoint();
    return is;
  }

 private:
  PointType p_;
  AttrT a_;
};

template <int Dim, typename AttrT, typename FloatT>
std::ostream &operator<<(std::ostream &os, const DecoratedPoint<Dim, AttrT, FloatT> &d) {
  return os << "{ " << d.point() << ", " << d.attributes() << " }";
}

}  // namespace bbrcit

#endif
