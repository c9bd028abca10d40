// Point2d.h -- compact 2-D point {x, y, m, abw}, the reference's device point
// type (include/Point2d.h:19-91).  Here it is a host value type: device data
// is uploaded through the C ABI as SoA arrays.  Assigning from a decorated
// point copies coordinates, mass and abw, rounding to T.
#ifndef BBRCIT_AMD_POINT2D_H_
#define BBRCIT_AMD_POINT2D_H_

namespace bbrcit {

template <typename T>
class Point2d {
 public:
  using FloatType = T;

  Point2d() = default;
  Point2d(T x, T y, T m) : x_(x), y_(y), m_(m), abw_(T(1)) {}
  Point2d(T x, T y, T m, T abw) : x_(x), y_(y), m_(m), abw_(abw) {}

  template <typename HostDecPointT>
  Point2d &operator=(const HostDecPointT &p) {
    x_ = (T)p[0];
    y_ = (T)p[1];
    m_ = (T)p.attributes().mass();
    abw_ = (T)p.attributes().abw();
    return *this;
  }

  T x() const { return x_; }
  T y() const { return y_; }
  T m() const { return m_; }
  T abw() const { return abw_; }

  // coordinate access, so kernels' generic point forms accept Point2d
  T operator[](int i) const { return i == 0 ? x_ : y_; }

 private:
  T x_ = T(0), y_ = T(0), m_ = T(1), abw_ = T(1);
};

}  // namespace bbrcit

#endif
