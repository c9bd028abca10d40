// Interval.h -- a closed interval [lower, upper] of the real line.
// Interface of the reference's include/Interval.h:14-150 (including its
// distance conventions: max_dist to a point measures to the far endpoint as
// seen from the midpoint).
#ifndef BBRCIT_AMD_INTERVAL_H_
#define BBRCIT_AMD_INTERVAL_H_

#include <algorithm>
#include <iostream>
#include <stdexcept>

namespace bbrcit {

template <typename T = double>
class Interval {
 public:
  using FloatType = T;

  Interval() : lo_(T()), hi_(T()) {}
  Interval(const T &lo, const T &hi) { resize(lo, hi); }

  T length() const { return hi_ - lo_; }
  T lower() const { return lo_; }
  T upper() const { return hi_; }
  T middle() const { return lo_ + (hi_ - lo_) / 2; }

  void resize(const T &lo, const T &hi) {
    if (hi < lo)
      throw std::logic_error("Interval<>: resize(const T &low, const T &high): "
                             "Arguments must satisfy low <= high. ");
    lo_ = lo;
    hi_ = hi;
  }

  bool contains(const T &p) const { return lo_ <= p && p <= hi_; }
  bool contains(const Interval &o) const { return lo_ <= o.lo_ && o.hi_ <= hi_; }

  T min_dist(const T &p) const {
    if (contains(p)) return T(0);
    return p < lo_ ? lo_ - p : p - hi_;
  }
  T max_dist(const T &p) const { return p < middle() ? hi_ - p : p - lo_; }
  T min_dist(const Interval &o) const {
    if (lo_ <= o.hi_ && o.lo_ <= hi_) return T(0);
    return hi_ < o.lo_ ? o.lo_ - hi_ : lo_ - o.hi_;
  }
  T max_dist(const Interval &o) const { return std::max(o.hi_ - lo_, hi_ - o.lo_); }

 private:
  T lo_, hi_;
};

template <typename T>
bool intersect(const Interval<T> &a, const Interval<T> &b) {
  return a.lower() <= b.upper() && b.lower() <= a.upper();
}

template <typename T>
std::ostream &operator<<(std::ostream &os, const Interval<T> &i) {
  return os << "(" << i.lower() << ", " << i.upper() << ")";
}

}  // namespace bbrcit

#endif
