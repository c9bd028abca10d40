// Point.h -- a point of D-dimensional Euclidean space.
// Interface of the reference's include/Point.h:15-89 (construction from an
// initializer list, indexing, vector arithmetic, stream I/O); storage is a
// fixed-size array here instead of a std::vector.
#ifndef BBRCIT_AMD_POINT_H_
#define BBRCIT_AMD_POINT_H_

#include <array>
#include <initializer_list>
#include <iostream>
#include <stdexcept>

namespace bbrcit {

template <int D, typename T = double>
class Point {
 public:
  using FloatType = T;
  static constexpr int dim() { return D; }

  Point() { c_.fill(T()); }
  Point(const std::initializer_list<T> &li) { reset(li); }

  // coordinate i takes element i of li; missing coordinates become T()
  void reset(const std::initializer_list<T> &li) {
    c_.fill(T());
    int i = 0;
    for (const T &v : li) {
      if (i >= D) break;
      c_[i++] = v;
    }
  }

  const T &operator[](int i) const {
    if (i < 0 || i >= D) throw std::out_of_range("Point<>: index out of range");
    return c_[i];
  }
  T &operator[](int i) {
    if (i < 0 || i >= D) throw std::out_of_range("Point<>: index out of range");
    return c_[i];
  }

  Point &operator*=(double s) {
    for (auto &v : c_) v *= s;
    return *this;
  }
  Point &operator/=(double s) {
    if (!s) throw std::domain_error("Point<>: operator/=(): division by zero. ");
    for (auto &v : c_) v /= s;
    return *this;
  }
  Point &operator+=(const Point &o) {
    for (int i = 0; i < D; ++i) c_[i] += o.c_[i];
    return *this;
  }
  Point &operator-=(const Point &o) {
    for (int i = 0; i < D; ++i) c_[i] -= o.c_[i];
    return *this;
  }

  friend void swap(Point &a, Point &b) { a.c_.swap(b.c_); }

 private:
  std::array<T, D> c_;
};

template <int D, typename T>
Point<D, T> operator+(Point<D, T> a, const Point<D, T> &b) { return a += b; }
template <int D, typename T>
Point<D, T> operator-(Point<D, T> a, const Point<D, T> &b) { return a -= b; }
template <int D, typename T>
Point<D, T> operator*(Point<D, T> a, double s) { return a *= s; }
template <int D, typename T>
Point<D, T> operator*(double s, Point<D, T> a) { return a *= s; }
template <int D, typename T>
Point<D, T> operator/(Point<D, T> a, double s) { return a /= s; }

// reads D whitespace-separated coordinates; on failure the point is reset
template <int D, typename T>
std::istream &operator>>(std::istream &is, Point<D, T> &p) {
  for (int i = 0; i < D && is; ++i) is >> p[i];
  if (!is) p = Point<D, T>();
  return is;
}

// prints (x0, x1, ..., xD-1)
template <int D, typename T>
std::ostream &operator<<(std::ostream &os, const Point<D, T> &p) {
  os << "(" << p[0];
  for (int i = 1; i < D; ++i) os << ", " << p[i];
  return os << ")";
}

}  // namespace bbrcit

#endif
