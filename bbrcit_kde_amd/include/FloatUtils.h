// FloatUtils.h -- floating-point comparisons and exact point orderings.
// Same functions as the reference's include/FloatUtils.h:15-75.
#ifndef BBRCIT_AMD_FLOATUTILS_H_
#define BBRCIT_AMD_FLOATUTILS_H_

#include <algorithm>
#include <cmath>
#include <limits>
#include <type_traits>

namespace bbrcit {

// |lhs-rhs| below ulp machine epsilons of |lhs+rhs|, or below the smallest normal.
template <typename T>
typename std::enable_if<!std::numeric_limits<T>::is_integer, bool>::type
cppref_almost_equal(T lhs, T rhs, int ulp) {
  const T d = std::abs(lhs - rhs);
  return d < std::numeric_limits<T>::epsilon() * ulp * std::abs(lhs + rhs) ||
         d < std::numeric_limits<T>::min();
}

// Absolute difference at most the smallest normal, or relative difference
// at most machine epsilon.
template <typename T>
typename std::enable_if<!std::numeric_limits<T>::is_integer, bool>::type
almost_equal(T lhs, T rhs) {
  const T d = std::abs(lhs - rhs);
  if (d <= std::numeric_limits<T>::min()) return true;
  return d <= std::numeric_limits<T>::epsilon() * std::max(std::abs(lhs), std::abs(rhs));
}

// Absolute difference at most abs_err, or relative difference at most rel_err.
template <typename T>
typename std::enable_if<!std::numeric_limits<T>::is_integer, bool>::type
approximately_equal(T lhs, T rhs, T rel_err, T abs_err = std::numeric_limits<T>::min()) {
  const T d = std::abs(lhs - rhs);
  if (d <= abs_err) return true;
  return d <= rel_err * std::max(std::abs(lhs), std::abs(rhs));
}

// Coordinate-wise equality with ==, intended for exact duplicate detection.
template <typename PointT>
bool ExactEqual(const PointT &a, const PointT &b) {
  for (int d = 0; d < a.dim(); ++d)
    if (!(a[d] == b[d])) return false;
  return true;
}

// Lexicographic order on coordinates, comparing with == and <.
template <typename PointT>
bool ExactLexicoLess(const PointT &a, const PointT &b) {
  for (int d = 0; d < a.dim(); ++d) {
    if (a[d] == b[d]) continue;
    return a[d] < b[d];
  }
  return false;
}

}  // namespace bbrcit

#endif
