// Rectangle.h -- an axis-aligned box of D-dimensional space.
// Interface of the reference's include/Rectangle.h:18-88: edges, halfspaces,
// containment, and the per-axis / total min and max distances the dual-tree
// bounds use.
#ifndef BBRCIT_AMD_RECTANGLE_H_
#define BBRCIT_AMD_RECTANGLE_H_

#include <array>
#include <cmath>
#include <sstream>
#include <stdexcept>
#include <utility>

#include "Interval.h"
#include "Point.h"

namespace bbrcit {

template <int D, typename T = double>
class Rectangle {
 public:
  using FloatType = T;
  using EdgeType = Interval<T>;
  static constexpr int dim() { return D; }

  Rectangle() { e_.fill(EdgeType(T(), T())); }
  Rectangle(const Point<D, T> &a, const Point<D, T> &b) { resize(a, b); }

  const EdgeType &operator[](int i) const { return e_.at(i); }
  EdgeType &operator[](int i) { return e_.at(i); }

  void resize(size_t d, const EdgeType &e) { e_.at(d) = e; }
  void resize(const Point<D, T> &a, const Point<D, T> &b) {
    for (int i = 0; i < D; ++i) {
      const auto mm = std::minmax(a[i], b[i]);
      e_[i] = EdgeType(mm.first, mm.second);
    }
  }

  Rectangle lower_halfspace(size_t d, const T &v) const { return split(d, v, true); }
  Rectangle upper_halfspace(size_t d, const T &v) const { return split(d, v, false); }

  template <typename GT>
  bool contains(const GT &g) const {
    for (int i = 0; i < D; ++i)
      if (!e_[i].contains(g[i])) return false;
    return true;
  }

  template <typename GT>
  T min_dist(size_t i, const GT &g) const { return e_[i].min_dist(g[i]); }
  template <typename GT>
  T max_dist(size_t i, const GT &g) const { return e_[i].max_dist(g[i]); }

  template <typename GT>
  T min_dist(const GT &g) const {
    T s = T(0);
    for (int i = 0; i < D; ++i) {
      const T v = e_[i].min_dist(g[i]);
      s += v * v;
    }
    return std::sqrt(s);
  }
  template <typename GT>
  T max_dist(const GT &g) const {
    T s = T(0);
    for (int i = 0; i < D; ++i) {
      const T v = e_[i].max_dist(g[i]);
      s += v * v;
    }
    return std::sqrt(s);
  }

  friend void swap(Rectangle &a, Rectangle &b) { a.e_.swap(b.e_); }

 private:
  Rectangle split(size_t d, const T &v, bool lower) const {
    if (!e_.at(d).contains(v)) {
      std::ostringstream os;
      os << "Rectangle<>: " << (lower ? "lower" : "upper")
         << "_halfspace(size_t d, const T &v): partition point " << v << " must be contained in edge "
         << e_[d];
      throw std::length_error(os.str());
    }
    Rectangle r(*this);
    r.e_[d] = lower ? EdgeType(e_[d].lower(), v) : EdgeType(v, e_[d].upper());
    return r;
  }

  std::array<EdgeType, D> e_;
};

template <int D, typename T>
bool intersect(const Rectangle<D, T> &a, const Rectangle<D, T> &b) {
  for (int i = 0; i < D; ++i)
    if (!intersect(a[i], b[i])) return false;
  return true;
}

template <int D, typename T>
std::ostream &operator<<(std::ostream &os, const Rectangle<D, T> &r) {
  os << "{ " << r[0];
  for (int i = 1; i < D; ++i) os << ", " << r[i];
  return os << " }";
}

}  // namespace bbrcit

#endif
