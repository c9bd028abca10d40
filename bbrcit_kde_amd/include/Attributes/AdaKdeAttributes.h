// AdaKdeAttributes.h -- per point / per tree node attributes of the kernel
// density: weight, mass, bounds of the adaptive bandwidth factor and bounds
// of the density value.
// Interface and merge rule of the reference's
// include/Attributes/AdaKdeAttributes.h:21-106: weight and mass add, the
// lower bounds take the minimum, the upper bounds the maximum; abw() and
// value() are the midpoints of their bounds.
#ifndef BBRCIT_AMD_ADAKDEATTRIBUTES_H_
#define BBRCIT_AMD_ADAKDEATTRIBUTES_H_

#include <algorithm>
#include <iostream>

#include "../KdeTraits.h"
#include "AttributesInterface.h"

namespace bbrcit {

template <typename T>
class AdaKdeAttributes {
 public:
  AdaKdeAttributes() : AdaKdeAttributes(ConstantTraits<T>::one()) {}
  AdaKdeAttributes(T w)
      : weight_(w), mass_(w), abw_lo_(ConstantTraits<T>::one()), abw_hi_(ConstantTraits<T>::one()),
        lo_(ConstantTraits<T>::zero()), hi_(ConstantTraits<T>::zero()) {}

  const T &weight() const { return weight_; }
  void set_weight(const T &w) { weight_ = w; }

  const T &mass() const { return mass_; }
  void set_mass(const T &m) { mass_ = m; }

  T abw() const { return abw_lo_ + (abw_hi_ - abw_lo_) / 2; }
  const T &lower_abw() const { return abw_lo_; }
  const T &upper_abw() const { return abw_hi_; }
  void set_lower_abw(const T &v) { abw_lo_ = v; }
  void set_upper_abw(const T &v) { abw_hi_ = v; }

  T value() const { return lo_ + (hi_ - lo_) / 2; }
  const T &lower() const { return lo_; }
  const T &upper() const { return hi_; }
  void set_lower(const T &v) { lo_ = v; }
  void set_upper(const T &v) { hi_ = v; }

  AdaKdeAttributes &merge(const AdaKdeAttributes &o) {
    weight_ += o.weight_;
    mass_ += o.mass_;
    abw_lo_ = std::min(abw_lo_, o.abw_lo_);
    abw_hi_ = std::max(abw_hi_, o.abw_hi_);
    lo_ = std::min(lo_, o.lo_);
    hi_ = std::max(hi_, o.hi_);
    return *this;
  }

 private:
  T weight_, mass_;
  T abw_lo_, abw_hi_;
  T lo_, hi_;
};

template <typename T>
std::ostream &operator<<(std::ostream &os, const AdaKdeAttributes<T> &a) {
  return os << "(" << a.weight() << ")";
}

}  // namespace bbrcit

#endif
