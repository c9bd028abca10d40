// Point1d.h -- compact 1-D point {x, m, abw}, the reference's device point
// type (include/Point1d.h:19-85), here a host value type.
#ifndef BBRCIT_AMD_POINT1D_H_
#define BBRCIT_AMD_POINT1D_H_

namespace bbrcit {

template <typename T>
class Point1d {
 public:
  using FloatType = T;

  Point1d() = default;
  Point1d(T x, T m) : x_(x), m_(m), abw_(T(1)) {}
  Point1d(T x, T m, T abw) : x_(x), m_(m), abw_(abw) {}

  template <typename HostDecPointT>
  Point1d &operator=(const HostDecPointT &p) {
    x_ = (T)p[0];
    m_ = (T)p.attributes().mass();
    abw_ = (T)p.attributes().abw();
    return *this;
  }

  T x() const { return x_; }
  T m() const { return m_; }
  T abw() const { return abw_; }
  T operator[](int) const { return x_; }

 private:
  T x_ = T(0), m_ = T(1), abw_ = T(1);
};

}  // namespace bbrcit

#endif
