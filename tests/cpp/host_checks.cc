// Host-side checks of the C++ API's threaded code, built by
// tests/test_host_sanitizers.py with AddressSanitizer + UndefinedBehaviorSanitizer
// and, separately, with ThreadSanitizer.  Without a HIP device the Kdtree takes
// its host paths: the threaded median partition (subtrees on their own threads
// above 2^18 points), the threaded node attributes, KernelDensity's threaded
// weight normalisation and HipDirectKde's threaded flatten.  Each check prints
// one line; the exit status is the number of failed checks.
#include <HipDirectKde.h>
#include <KernelDensity.h>
#include <Kernels/GaussianProductKernel2d.h>

#include <cmath>
#include <cstdio>
#include <random>
#include <vector>

using namespace bbrcit;

namespace {

int failures = 0;

void check(bool ok, const char *what) {
  std::printf("%s %s\n", ok ? "ok  " : "FAIL", what);
  if (!ok) ++failures;
}

// n points of a 2-D mixture, every tenth one a copy of an earlier point
// (duplicates are merged by the tree build)
template <class KDE>
std::vector<typename KDE::DataPointType> mixture(size_t n, unsigned seed) {
  using P = typename KDE::DataPointType;
  std::mt19937_64 e(seed);
  std::normal_distribution<double> g(0.0, 1.0);
  std::uniform_real_distribution<double> u(0.5, 2.0);
  std::vector<P> pts;
  pts.reserve(n);
  for (size_t i = 0; i < n; ++i) {
    typename P::PointType x;
    if (i % 10 == 9) {
      const auto &c = pts[(i * 7919) % i];
      x[0] = c[0];
      x[1] = c[1];
    } else {
      const double c = (i % 3) - 1.0;
      x[0] = c + 0.4 * g(e);
      x[1] = -c + 0.6 * g(e);
    }
    pts.emplace_back(x, typename KDE::AttributeType(u(e)));
  }
  return pts;
}

template <class Tree>
bool tree_invariants(const Tree &t, double total_weight) {
  const auto &pts = t.points();
  bool ok = true;
  size_t leaf_points = 0;
  t.visit_nodes([&](size_t i, size_t j, double, const auto &cell, const auto &a, bool leaf) {
    if (j < i || j >= pts.size()) {
      ok = false;
      return;
    }
    double w = 0.0;
    for (size_t k = i; k <= j; ++k) {
      w += pts[k].attributes().weight();
      if (!cell.contains(pts[k])) ok = false;
    }
    if (std::fabs(w - a.weight()) > 1e-9 * std::fabs(w)) ok = false;
    if (leaf) leaf_points += j - i + 1;
  });
  double w = 0.0;
  for (const auto &p : pts) w += p.attributes().weight();
  return ok && leaf_points == pts.size() && std::fabs(w - total_weight) <= 1e-9 * total_weight;
}

template <class Tree>
bool same_points(const Tree &a, const Tree &b) {
  const auto &p = a.points(), &q = b.points();
  if (p.size() != q.size()) return false;
  for (size_t i = 0; i < p.size(); ++i)
    if (p[i][0] != q[i][0] || p[i][1] != q[i][1] || p[i].attributes().weight() != q[i].attributes().weight())
      return false;
  return true;
}

}  // namespace

int main() {
  const size_t n = 300000;  // above the 2^18 threshold of every threaded path

  {
    using Tree = Kdtree<2, PointWeights<double>, double>;
    using P = Tree::DataPointType;
    std::vector<P> pts;
    std::mt19937_64 e(5);
    std::uniform_real_distribution<double> u(-1.0, 1.0);
    double total = 0.0;
    for (size_t i = 0; i < n; ++i) {
      P::PointType x;
      x[0] = i % 10 == 9 ? pts[i / 2][0] : u(e);
      x[1] = i % 10 == 9 ? pts[i / 2][1] : u(e);
      pts.emplace_back(x, PointWeights<double>(1.0));
      total += 1.0;
    }
    Tree a(pts, 32), b(pts, 32);
    check(tree_invariants(a, total), "Kdtree<2, PointWeights<double>>: cells hold their points, weights add up");
    check(same_points(a, b), "Kdtree: two threaded builds give the same order");
    check(a.points().size() < n, "Kdtree: duplicates merged");
  }

  {
    using KDE = KernelDensity<2, GaussianProductKernel2d<double>, double>;
    auto pts = mixture<KDE>(n, 11);
    KDE kde(pts, 64), again(pts, 64);
    check(tree_invariants(kde.data_tree(), 1.0), "KernelDensity: normalised weights, cells hold their points");
    check(same_points(kde.data_tree(), again.data_tree()), "KernelDensity: two builds give the same order");
    bool masses = true;
    for (const auto &p : kde.points())
      if (p.attributes().mass() != p.attributes().weight()) masses = false;
    check(masses, "KernelDensity: mass = weight with abw 1");

    // flatten on threads against an element-wise copy
    std::vector<double> xy, m, a;
    hip_detail::flatten<2>(kde.points(), xy, m, a);
    bool same = xy.size() == 2 * kde.points().size();
    for (size_t i = 0; same && i < kde.points().size(); ++i) {
      const auto &p = kde.points()[i];
      same = xy[2 * i] == p[0] && xy[2 * i + 1] == p[1] && m[i] == p.attributes().mass() &&
             a[i] == p.attributes().abw();
    }
    check(same, "flatten: threaded copy equals the points");
  }

  std::printf("%d failed\n", failures);
  return failures;
}
