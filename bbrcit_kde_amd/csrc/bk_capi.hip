// C-ABI of the direct kernel-density library (declared in include/bbrcit_hip.h).
//
// A context owns the device copies of the reference and query sets, the
// cached per-reference features of the last kernel descriptor, and staging
// buffers that grow on demand and are reused across calls.  Every call
// validates its arguments the way CudaDirectKde::unnormalized_eval does
// (include/CudaDirectKde.h:181-190), checks every HIP return code, and turns
// failures into status codes plus a thread-local message.
#include "../../include/bbrcit_hip.h"
#include "bk_common.h"
#include "bk_internal.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

namespace {

using bk::DevBuf;
using bk::DeviceGuard;
using bk::fail;
using bk::guarded;
using bk::hip_check;

struct PointSet {
  size_t n = 0;
  int dim = 0;
  bool uniform_abw = true;  // every abw == 1
  double m_max = 0.0;       // largest mass (as stored on the device)
  double m_sum = 0.0;       // sum of the masses (as stored on the device)
  float box_lo = 0.0f, box_hi = 1.0f;  // square holding every FP32 point (D <= 2), for spatial sorting
  // FP32 storage: x, y (D <= 2; y == 0 for D == 1) or row-major coords (D > 2)
  DevBuf x, y, coords, m, abw;
  // FP64 storage (D <= 2)
  DevBuf xd, yd, md, abwd;
  unsigned long long version = 0;
  void release() {
    for (DevBuf *b : {&x, &y, &coords, &m, &abw, &xd, &yd, &md, &abwd}) b->release();
    n = 0;
    dim = 0;
  }
};

}  // namespace

struct bk_ctx {
  int device = 0;
  int precision = BK_PREC_F32;
  hipStream_t own_stream = nullptr;
  hipStream_t stream = nullptr;
  PointSet refs, queries;
  bool alias = false;
  unsigned long long version_counter = 0;

  // feature cache (per reference set and kernel descriptor)
  bool feat_valid = false;
  bk_kernel_desc feat_desc{};
  unsigned long long feat_version = 0;
  DevBuf f0, f1, tiles;

  // options (bk_set_option)
  int gauss_form = BK_GAUSS_FORM_AUTO;
  int high_dim_form = BK_HIGH_DIM_AUTO;
  int feat_path = -1;  // Path of the cached features
  int query_order = BK_QUERY_ORDER_GIVEN;
  int count_terms = 0;  // BK_OPT_COUNT_TERMS
  DevBuf term_count;    // two counters of the counting FP32 Gaussian / Epanechnikov launches
  DevBuf ord_list, ord_keys_a, ord_keys_b, ord_tmp, ord_x, ord_y, ord_out;  // spatial query order
  // bounded evaluation: tile summaries of the reference set (per version),
  // per-block thresholds, pruning counters, bounds planes
  DevBuf tsum, bnd_mask, bnd_lo, bnd_gap, bnd_counts, bnd_exact, bnd_hi, bnd_vals, bnd_order, live_boxes;
  unsigned long long tsum_version = 0;

  // staging
  DevBuf partial, out, items, work_a, work_b, weights, scratch, scalar;
  std::vector<double> host_tmp;
};

namespace bk {
std::string &last_error() {
  thread_local std::string msg;
  return msg;
}
int ctx_device(const bk_ctx *ctx) { return ctx->device; }
hipStream_t ctx_stream(const bk_ctx *ctx) { return ctx->stream; }
}  // namespace bk

namespace {

using bk::EvalParams;
using bk::WorkItem;

constexpr double kLog2e = 1.4426950408889634074;
constexpr double kPi = 3.14159265358979323846;

void require_ctx(const bk_ctx *ctx) {
  if (!ctx) fail(BK_ERR_INVALID_ARG, "null context");
}

// Kernel descriptor -> device family and constants.  The constants restate
// the reference's KernelTraits (include/Kernels/KernelTraits.h): the
// normalisations C(h) and the exponent constants of the Gaussian kinds,
// including GaussianProductConv2dTraits<float|double>, whose kernel is
// exp(-0.5 arg) (KernelTraits.h:432, 445).
EvalParams params_of(const bk_kernel_desc *d) {
  if (!d) fail(BK_ERR_INVALID_ARG, "null kernel descriptor");
  EvalParams p{};
  p.dim = d->dim;
  p.mass_fold = 1.0;
  const double hx = d->hx, hy = d->hy;
  auto need_pos = [](double h, const char *name) {
    if (!(h > 0.0) || !std::isfinite(h))
      fail(BK_ERR_INVALID_ARG, std::string("bandwidth ") + name + " must be positive and finite");
  };
  auto need_dim = [&](int lo, int hi) {
    if (d->dim < lo || d->dim > hi)
      fail(BK_ERR_UNSUPPORTED, "kernel kind " + std::to_string(d->kind) +
                                   " does not support dimension " + std::to_string(d->dim));
  };
  need_pos(hx, "hx");
  switch (d->kind) {
    case BK_GAUSSIAN: {
      need_dim(1, 16);
      p.family = bk::FAM_GAUSS;
      p.c = 0.5;
      p.gx = p.gy = p.c * kLog2e / (hx * hx);
      p.iso = 1;
      p.norm = std::pow(2.0 * kPi, -0.5 * d->dim) / std::pow(hx, d->dim);
      break;
    }
    case BK_EPANECHNIKOV: {
      need_dim(1, 16);
      p.family = bk::FAM_EPAN;
      p.gx = p.gy = 1.0 / (hx * hx);
      p.ix = p.iy = 1.0 / hx;
      p.iso = 1;
      const double D = d->dim;
      p.norm = 0.5 * (D + 2.0) / (std::pow(kPi, 0.5 * D) / std::tgamma(1.0 + 0.5 * D)) /
               std::pow(hx, D);
      break;
    }
    case BK_GAUSSIAN_PRODUCT_2D:
    case BK_GAUSSIAN_PRODUCT_CONV_2D: {
      need_dim(2, 2);
      need_pos(hy, "hy");
      p.family = bk::FAM_GAUSS;
      p.c = 0.5;
      p.gx = p.c * kLog2e / (hx * hx);
      p.gy = p.c * kLog2e / (hy * hy);
      p.iso = hx == hy;
      p.norm = d->kind == BK_GAUSSIAN_PRODUCT_2D ? 1.0 / (2.0 * kPi * hx * hy)
                                                 : 1.0 / (4.0 * kPi * hx * hy);
      break;
    }
    case BK_GAUSSIAN_CONV_1D: {
      need_dim(1, 1);
      p.family = bk::FAM_GAUSS;
      p.c = 0.25;
      p.gx = p.gy = p.c * kLog2e / (hx * hx);
      p.iso = 1;
      p.norm = 1.0 / (std::sqrt(4.0 * kPi) * hx);
      break;
    }
    case BK_EPANECHNIKOV_PRODUCT_2D: {
      need_dim(2, 2);
      need_pos(hy, "hy");
      p.family = bk::FAM_EPANPROD;
      p.gx = 1.0 / (hx * hx);
      p.gy = 1.0 / (hy * hy);
      p.ix = 1.0 / hx;
      p.iy = 1.0 / hy;
      p.iso = 0;
      p.norm = 0.5625 / (hx * hy);
      break;
    }
    case BK_EPANECHNIKOV_CONV_1D: {
      need_dim(1, 1);
      p.family = bk::FAM_EPANCONV;
      p.ix = 1.0 / hx;
      p.iy = 1.0;  // y == 0 for every 1-D point: the y factor is k(0) = 1
      p.gx = p.gy = 1.0;
      p.mass_fold = 1.0 / 1024.0;
      p.norm = 0.6 / hx;
      break;
    }
    case BK_EPANECHNIKOV_PRODUCT_CONV_2D: {
      need_dim(2, 2);
      need_pos(hy, "hy");
      p.family = bk::FAM_EPANCONV;
      p.ix = 1.0 / hx;
      p.iy = 1.0 / hy;
      p.gx = p.gy = 1.0;
      p.mass_fold = 1.0 / 1024.0;
      p.norm = 0.36 / (hx * hy);
      break;
    }
    default:
      fail(BK_ERR_UNSUPPORTED, "unknown kernel kind " + std::to_string(d->kind));
  }
  return p;
}

PointSet &set_of(bk_ctx *ctx, int set) {
  if (set == BK_SET_REF) return ctx->refs;
  if (set == BK_SET_QUERY) {
    if (ctx->alias) fail(BK_ERR_STATE, "the query set is an alias of the reference set");
    return ctx->queries;
  }
  fail(BK_ERR_INVALID_ARG, "set must be BK_SET_REF or BK_SET_QUERY");
}

const PointSet &query_set(const bk_ctx *ctx) { return ctx->alias ? ctx->refs : ctx->queries; }

// body(lo, hi) over disjoint slices of [0, n), on up to 16 host threads for
// large n; body must not throw.
template <class F>
void host_parallel(size_t n, F &&body, size_t min_n = size_t(1) << 18) {
  const unsigned hw = std::thread::hardware_concurrency();
  const size_t t = n < min_n || hw < 2 ? 1 : std::min<size_t>({(size_t)hw, (size_t)16, n});
  if (t == 1) {
    body(size_t(0), n);
    return;
  }
  std::vector<std::thread> workers;
  for (size_t c = 0; c < t; ++c) workers.emplace_back([&body, n, c, t] { body(n * c / t, n * (c + 1) / t); });
  for (auto &w : workers) w.join();
}

// Copy host data into a point set; masses/abw default to 1.
void upload_set(bk_ctx *ctx, PointSet &ps, int dim, size_t n, const double *coords,
                const double *mass, const double *abw) {
  if (dim < 1 || dim > 16) fail(BK_ERR_UNSUPPORTED, "dimension must be between 1 and 16");
  if (n > 0 && !coords) fail(BK_ERR_INVALID_ARG, "null coordinate array");
  if (ctx->precision == BK_PREC_F64 && dim > 2)
    fail(BK_ERR_UNSUPPORTED, "FP64 contexts support dimensions 1 and 2");
  ps.dim = dim;
  ps.n = n;
  ps.uniform_abw = true;
  ps.m_max = 0.0;
  ps.m_sum = 0.0;
  ps.version = ++ctx->version_counter;
  if (n == 0) return;
  const hipStream_t s = ctx->stream;
  if (ctx->precision == BK_PREC_F32) {
    // host buffers left uninitialised (the threads write every element, so
    // their pages are first touched in parallel)
    std::unique_ptr<float[]> m(new float[n]), a(new float[n]);
    // conversion and validation on threads; the mass sum over fixed chunks
    // added in chunk order, so it does not depend on the thread count
    constexpr size_t kChunk = size_t(1) << 16;
    const size_t n_chunks = (n + kChunk - 1) / kChunk;
    std::vector<double> chunk_sum(n_chunks, 0.0), chunk_max(n_chunks, 0.0);
    std::vector<char> chunk_uniform(n_chunks, 1), chunk_bad_abw(n_chunks, 0), chunk_bad_m(n_chunks, 0);
    host_parallel(n_chunks, [&](size_t c0, size_t c1) {
      for (size_t c = c0; c < c1; ++c) {
        double sum = 0.0, mx = 0.0;
        char uniform = 1, bad_a = 0, bad_m = 0;
        for (size_t i = c * kChunk, e = std::min(n, (c + 1) * kChunk); i < e; ++i) {
          const float mi = mass ? (float)mass[i] : 1.0f;
          const float ai = abw ? (float)abw[i] : 1.0f;
          m[i] = mi;
          a[i] = ai;
          uniform &= ai == 1.0f;
          bad_a |= !(ai > 0.0f) || !std::isfinite(ai);
          bad_m |= !(mi >= 0.0f) || !std::isfinite(mi);
          mx = std::max(mx, (double)mi);
          sum += (double)mi;
        }
        chunk_sum[c] = sum;
        chunk_max[c] = mx;
        chunk_uniform[c] = uniform;
        chunk_bad_abw[c] = bad_a;
        chunk_bad_m[c] = bad_m;
      }
    }, size_t(4));
    for (size_t c = 0; c < n_chunks; ++c) {
      if (chunk_bad_abw[c]) fail(BK_ERR_INVALID_ARG, "abw must be positive and finite");
      if (chunk_bad_m[c]) fail(BK_ERR_INVALID_ARG, "masses must be non-negative and finite");
      if (!chunk_uniform[c]) ps.uniform_abw = false;
      ps.m_max = std::max(ps.m_max, chunk_max[c]);
      ps.m_sum += chunk_sum[c];
    }
    ps.m.reserve(n * sizeof(float));
    ps.abw.reserve(n * sizeof(float));
    hip_check(hipMemcpyAsync(ps.m.ptr, m.get(), n * sizeof(float), hipMemcpyHostToDevice, s), "upload masses");
    hip_check(hipMemcpyAsync(ps.abw.ptr, a.get(), n * sizeof(float), hipMemcpyHostToDevice, s), "upload abw");
    if (dim <= 2) {
      std::unique_ptr<float[]> x(new float[n]), y(new float[n]);
      std::mutex part_mu;
      float lo = (float)coords[0], hi = lo;
      host_parallel(n, [&](size_t b, size_t e) {
        float l = (float)coords[b * dim], h = l;
        for (size_t i = b; i < e; ++i) {
          x[i] = (float)coords[i * dim];
          y[i] = dim == 2 ? (float)coords[i * dim + 1] : 0.0f;
          l = std::min(l, std::min(x[i], y[i]));
          h = std::max(h, std::max(x[i], y[i]));
        }
        std::lock_guard<std::mutex> g(part_mu);  // min and max do not depend on the order
        lo = std::min(lo, l);
        hi = std::max(hi, h);
      });
      ps.box_lo = std::isfinite(lo) ? lo : -1.0f;
      ps.box_hi = std::isfinite(hi) && hi > ps.box_lo ? hi : ps.box_lo + 1.0f;
      ps.x.reserve(n * sizeof(float));
      ps.y.reserve(n * sizeof(float));
      hip_check(hipMemcpyAsync(ps.x.ptr, x.get(), n * sizeof(float), hipMemcpyHostToDevice, s), "upload x");
      hip_check(hipMemcpyAsync(ps.y.ptr, y.get(), n * sizeof(float), hipMemcpyHostToDevice, s), "upload y");
      hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    } else {
      std::vector<float> c(n * dim);
      for (size_t i = 0; i < n * dim; ++i) c[i] = (float)coords[i];
      ps.coords.reserve(n * dim * sizeof(float));
      hip_check(hipMemcpyAsync(ps.coords.ptr, c.data(), n * dim * sizeof(float), hipMemcpyHostToDevice, s),
                "upload coords");
      hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    }
  } else {
    std::vector<double> x(n), y(n, 0.0), m(n), a(n);
    for (size_t i = 0; i < n; ++i) {
      x[i] = coords[i * dim];
      if (dim == 2) y[i] = coords[i * dim + 1];
      m[i] = mass ? mass[i] : 1.0;
      a[i] = abw ? abw[i] : 1.0;
      if (a[i] != 1.0) ps.uniform_abw = false;
      if (!(a[i] > 0.0) || !std::isfinite(a[i])) fail(BK_ERR_INVALID_ARG, "abw must be positive and finite");
      if (!(m[i] >= 0.0) || !std::isfinite(m[i]))
        fail(BK_ERR_INVALID_ARG, "masses must be non-negative and finite");
      ps.m_max = std::max(ps.m_max, m[i]);
      ps.m_sum += m[i];
    }
    if (dim <= 2) {
      // square for spatial sorting, in FP32 (the Morton cells only group points)
      float lo = (float)x[0], hi = lo;
      for (size_t i = 0; i < n; ++i) {
        lo = std::min(lo, std::min((float)x[i], (float)y[i]));
        hi = std::max(hi, std::max((float)x[i], (float)y[i]));
      }
      ps.box_lo = std::isfinite(lo) ? lo : -1.0f;
      ps.box_hi = std::isfinite(hi) && hi > ps.box_lo ? hi : ps.box_lo + 1.0f;
    }
    for (DevBuf *b : {&ps.xd, &ps.yd, &ps.md, &ps.abwd}) b->reserve(n * sizeof(double));
    hip_check(hipMemcpyAsync(ps.xd.ptr, x.data(), n * 8, hipMemcpyHostToDevice, s), "upload x");
    hip_check(hipMemcpyAsync(ps.yd.ptr, y.data(), n * 8, hipMemcpyHostToDevice, s), "upload y");
    hip_check(hipMemcpyAsync(ps.md.ptr, m.data(), n * 8, hipMemcpyHostToDevice, s), "upload masses");
    hip_check(hipMemcpyAsync(ps.abwd.ptr, a.data(), n * 8, hipMemcpyHostToDevice, s), "upload abw");
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
  }
}

// Replace masses and/or abw of the reference set (same n, same dim).
void update_refs(bk_ctx *ctx, const double *mass, const double *abw) {
  PointSet &ps = ctx->refs;
  const size_t n = ps.n;
  if (n == 0) return;
  const hipStream_t s = ctx->stream;
  ps.version = ++ctx->version_counter;
  if (ctx->precision == BK_PREC_F32) {
    std::vector<float> buf(n);
    if (mass) {
      ps.m_max = 0.0;
      ps.m_sum = 0.0;
      for (size_t i = 0; i < n; ++i) {
        buf[i] = (float)mass[i];
        if (!(buf[i] >= 0.0f) || !std::isfinite(buf[i]))
          fail(BK_ERR_INVALID_ARG, "masses must be non-negative and finite");
        ps.m_max = std::max(ps.m_max, (double)buf[i]);
        ps.m_sum += (double)buf[i];
      }
      hip_check(hipMemcpyAsync(ps.m.ptr, buf.data(), n * 4, hipMemcpyHostToDevice, s), "update masses");
      hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    }
    if (abw) {
      ps.uniform_abw = true;
      for (size_t i = 0; i < n; ++i) {
        buf[i] = (float)abw[i];
        if (!(buf[i] > 0.0f) || !std::isfinite(buf[i])) fail(BK_ERR_INVALID_ARG, "abw must be positive and finite");
        if (buf[i] != 1.0f) ps.uniform_abw = false;
      }
      hip_check(hipMemcpyAsync(ps.abw.ptr, buf.data(), n * 4, hipMemcpyHostToDevice, s), "update abw");
      hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    }
  } else {
    if (mass) {
      ps.m_max = 0.0;
      ps.m_sum = 0.0;
      for (size_t i = 0; i < n; ++i) {
        if (!(mass[i] >= 0.0) || !std::isfinite(mass[i]))
          fail(BK_ERR_INVALID_ARG, "masses must be non-negative and finite");
        ps.m_max = std::max(ps.m_max, mass[i]);
        ps.m_sum += mass[i];
      }
      hip_check(hipMemcpyAsync(ps.md.ptr, mass, n * 8, hipMemcpyHostToDevice, s), "update masses");
    }
    if (abw) {
      ps.uniform_abw = true;
      for (size_t i = 0; i < n; ++i) {
        if (!(abw[i] > 0.0) || !std::isfinite(abw[i])) fail(BK_ERR_INVALID_ARG, "abw must be positive and finite");
        if (abw[i] != 1.0) ps.uniform_abw = false;
      }
      hip_check(hipMemcpyAsync(ps.abwd.ptr, abw, n * 8, hipMemcpyHostToDevice, s), "update abw");
    }
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
  }
}

double mass_scale(const PointSet &ps) { return ps.m_max > 0.0 ? ps.m_max : 1.0; }

bool same_desc(const bk_kernel_desc &a, const bk_kernel_desc &b) {
  return a.kind == b.kind && a.dim == b.dim && a.hx == b.hx && a.hy == b.hy;
}

// Which kernel family evaluates a descriptor in this context.
enum class Path { F64, GAUSS32, EPAN64, GENERAL, GRAM };

Path path_of(const bk_ctx *ctx, const EvalParams &p) {
  if (ctx->precision == BK_PREC_F64) return Path::F64;
  if (p.dim > 2) return (p.dim >= 8 && ctx->high_dim_form == BK_HIGH_DIM_AUTO) ? Path::GRAM : Path::GENERAL;
  return p.family == bk::FAM_GAUSS ? Path::GAUSS32 : Path::EPAN64;
}

// Queries handled by one workgroup / work item on a path.
int queries_per_item(Path path) {
  if (path == Path::GAUSS32) return bk::kGaussQueriesPerBlock;
  if (path == Path::EPAN64) return bk::kEpanQueriesPerBlock;
  return path == Path::F64 ? bk::kQueriesPerBlock : 256;
}

int tile_of(Path path) {
  if (path == Path::GAUSS32) return bk::kGaussTile;
  return (path == Path::GENERAL || path == Path::GRAM) ? 128 : 256;
}

// Device addresses of the per-reference features of one kernel descriptor
// (f1 and tiles may be unused by a path).
struct FeatPtrs {
  void *f0 = nullptr, *f1 = nullptr, *tiles = nullptr;
};

FeatPtrs ctx_feats(const bk_ctx *ctx) { return FeatPtrs{ctx->f0.ptr, ctx->f1.ptr, ctx->tiles.ptr}; }

// Bytes of the three feature arrays of a path for n references.
struct FeatBytes {
  size_t f0 = 0, f1 = 0, tiles = 0;
};

FeatBytes feature_bytes(Path path, size_t n) {
  FeatBytes b;
  switch (path) {
    case Path::F64:
      b.f0 = b.f1 = n * sizeof(double4);
      b.tiles = (size_t)bk::epan_meta_tiles((long long)n) * bk::kF64TileMeta * sizeof(double4);
      break;
    case Path::GAUSS32:
      // f1: whole tiles (a factored tile's packed features fill the first
      // three quarters of its block, masked past n)
      b.f0 = n * sizeof(float4);
      b.f1 = (n + bk::kGaussTile - 1) / bk::kGaussTile * bk::kGaussTile * sizeof(float4);
      b.tiles = (n + bk::kGaussTile - 1) / bk::kGaussTile * bk::kGaussTileMeta * sizeof(float4);
      break;
    case Path::EPAN64:
      b.f0 = b.f1 = n * sizeof(double4);
      b.tiles = (size_t)bk::epan_meta_tiles((long long)n) * bk::kEpanTileMeta * sizeof(double4);
      break;
    case Path::GENERAL:
      b.f0 = n * sizeof(double2);
      break;
    case Path::GRAM:
      b.f0 = n * sizeof(double4);
      break;
  }
  return b;
}

// Prepare the features of `p` for the context's reference set into fp.
void launch_prep(bk_ctx *ctx, Path path, const EvalParams &p, const FeatPtrs &fp) {
  PointSet &r = ctx->refs;
  const size_t n = r.n;
  const hipStream_t s = ctx->stream;
  switch (path) {
    case Path::F64:
      hip_check(bk::launch_prep_f64(p, r.xd.as<double>(), r.yd.as<double>(), r.md.as<double>(),
                                    r.abwd.as<double>(), n, static_cast<double4 *>(fp.f0),
                                    static_cast<double4 *>(fp.f1), static_cast<double4 *>(fp.tiles), s),
                "prep_f64 launch");
      break;
    case Path::GAUSS32:
      hip_check(bk::launch_prep_gauss(p, r.x.as<float>(), r.y.as<float>(), r.m.as<float>(), r.abw.as<float>(), n,
                                      mass_scale(r), static_cast<float4 *>(fp.f0), static_cast<float4 *>(fp.f1),
                                      static_cast<float4 *>(fp.tiles), s),
                "prep_gauss launch");
      break;
    case Path::EPAN64:
      hip_check(bk::launch_prep_epan(p, r.x.as<float>(), r.y.as<float>(), r.m.as<float>(), r.abw.as<float>(), n,
                                     static_cast<double4 *>(fp.f0), static_cast<double4 *>(fp.f1),
                                     static_cast<double4 *>(fp.tiles), s),
                "prep_epan launch");
      break;
    case Path::GENERAL:
      hip_check(bk::launch_prep_general(p, r.abw.as<float>(), r.m.as<float>(), n, static_cast<double2 *>(fp.f0), s),
                "prep_general launch");
      break;
    case Path::GRAM:
      hip_check(bk::launch_prep_gram(p, r.coords.as<float>(), r.abw.as<float>(), r.m.as<float>(), n,
                                     static_cast<double4 *>(fp.f0), s),
                "prep_gram launch");
      break;
  }
}

// Make sure the per-reference features of `desc` are on the device.
void ensure_features(bk_ctx *ctx, const bk_kernel_desc *desc, const EvalParams &p) {
  PointSet &r = ctx->refs;
  const Path want = path_of(ctx, p);
  if (ctx->feat_valid && same_desc(ctx->feat_desc, *desc) && ctx->feat_version == r.version &&
      ctx->feat_path == (int)want)
    return;
  ctx->feat_valid = false;
  const size_t n = r.n;
  if (n == 0) return;
  const FeatBytes fb = feature_bytes(want, n);
  ctx->f0.reserve(fb.f0);
  if (fb.f1) ctx->f1.reserve(fb.f1);
  if (fb.tiles) ctx->tiles.reserve(fb.tiles);
  launch_prep(ctx, want, p, ctx_feats(ctx));
  ctx->feat_path = (int)want;
  ctx->feat_desc = *desc;
  ctx->feat_version = r.version;
  ctx->feat_valid = true;
}

void check_compatible(const bk_ctx *ctx, const EvalParams &p) {
  const PointSet &q = query_set(ctx);
  if (ctx->refs.n == 0 && q.n == 0) return;
  if (ctx->refs.dim != p.dim || (q.n > 0 && q.dim != p.dim))
    fail(BK_ERR_INVALID_ARG, "kernel dimension " + std::to_string(p.dim) +
                                 " does not match the uploaded point dimension");
  if (p.dim > 2 && p.family != bk::FAM_GAUSS && p.family != bk::FAM_EPAN)
    fail(BK_ERR_UNSUPPORTED, "dimensions above 2 support the Gaussian and Epanechnikov kernels only");
}

void check_range(const bk_ctx *ctx, size_t r_i, size_t r_j, size_t q_i, size_t q_j) {
  const PointSet &q = query_set(ctx);
  if (r_j >= ctx->refs.n || q_j >= q.n)
    fail(BK_ERR_OUT_OF_RANGE, "r_j and q_j must be less than the number of reference and query points");
  if (r_i > r_j || q_i > q_j) fail(BK_ERR_INVALID_ARG, "must have r_i <= r_j and q_i <= q_j");
}

// Reference splits of a range evaluation: about kSplitTarget workgroups per
// CU.  With `slots` resident workgroups known, the count is moved (within
// +-25%) to the value whose last round of equal workgroups is fullest.
int choose_splits(long long nq, long long nr, int queries_per_block, int tile, long long slots = 0) {
  const long long qblocks = (nq + queries_per_block - 1) / queries_per_block;
  const long long target = (long long)bk::kSplitTarget * bk::device_cu_count();
  long long splits = (target + qblocks - 1) / qblocks;
  const long long max_splits = std::min(65535LL, std::max(1LL, nr / (4LL * tile)));
  splits = std::max(1LL, std::min(splits, max_splits));
  if (slots > 0 && splits > 4) {
    long long best = splits;
    double best_eff = 0.0;
    for (long long s = splits - splits / 4; s <= std::min(max_splits, splits + splits / 4); ++s) {
      // non-empty splits after tile alignment
      long long len = (nr + s - 1) / s;
      len = (len + tile - 1) / tile * tile;
      const long long eff_s = (nr + len - 1) / len;
      const double items = (double)qblocks * (double)eff_s;
      const double rounds = std::ceil(items / (double)slots);
      const double eff = items / (rounds * (double)slots);
      if (eff > best_eff + 1e-9) {
        best_eff = eff;
        best = s;
      }
    }
    splits = best;
  }
  return (int)splits;
}

// Result scale of a path: FP32 Gaussian features carry masses divided by
// the largest mass.
double path_scale(const bk_ctx *ctx, Path path, const EvalParams &p, int normalize) {
  double scale = normalize ? p.norm : 1.0;
  if (path == Path::GAUSS32) scale *= mass_scale(ctx->refs);
  return scale;
}

// Launch the evaluation kernels of one call: range mode (items == nullptr)
// or work-item mode, with the features at fp.
hipError_t launch_path(bk_ctx *ctx, Path path, const EvalParams &p, const FeatPtrs &fp, const WorkItem *items,
                       int n_items, long long r_i, long long r_end, long long q_i, long long nq, int splits,
                       double *dst, double scale, int self_excl, const bk::BoundArgs *order = nullptr) {
  const hipStream_t s = ctx->stream;
  const PointSet &r = ctx->refs;
  const PointSet &q = query_set(ctx);
  switch (path) {
    case Path::F64:
      return bk::launch_direct_f64(p, q.xd.as<double>(), q.yd.as<double>(), static_cast<const double4 *>(fp.f0),
                                   static_cast<const double4 *>(fp.f1), static_cast<const double4 *>(fp.tiles),
                                   items, n_items, r_i, r_end, q_i, nq, splits, dst, scale, self_excl,
                                   r.uniform_abw, s, order);
    case Path::GAUSS32:
      return bk::launch_gauss(p, q.x.as<float>(), q.y.as<float>(), static_cast<const float4 *>(fp.f0),
                              static_cast<const float4 *>(fp.f1), static_cast<const float4 *>(fp.tiles), items,
                              n_items, r_i, r_end, q_i, nq, splits, dst, scale, self_excl,
                              ctx->gauss_form == BK_GAUSS_FORM_DIRECT ? bk::GAUSS_DIRECT : bk::GAUSS_AUTO, s, order);
    case Path::EPAN64:
      return bk::launch_epan(p, q.x.as<float>(), q.y.as<float>(), static_cast<const double4 *>(fp.f0),
                             static_cast<const double4 *>(fp.f1), static_cast<const double4 *>(fp.tiles), items,
                             n_items, r_i, r_end, q_i, nq, splits, dst, scale, self_excl, r.uniform_abw, s, order);
    case Path::GRAM:
      return bk::launch_gram(p, q.coords.as<float>(), r.coords.as<float>(), static_cast<const double4 *>(fp.f0),
                             items, n_items, r_i, r_end, q_i, nq, splits, dst, scale, self_excl, s);
    case Path::GENERAL:
    default:
      return bk::launch_direct_general(p, q.coords.as<float>(), r.coords.as<float>(),
                                       static_cast<const double2 *>(fp.f0), items, n_items, r_i, r_end, q_i, nq,
                                       splits, dst, scale, self_excl, s);
  }
}

// Dispatch order of a culled direct launch (queries in Z-curve order, or self
// sums over kd-ordered points): the (block, split) pairs sorted by the tiles
// they keep, most first, so the launch does not end on a few long
// workgroups.  Results do not depend on it.  Returns an order-free
// BoundArgs where it does not apply.  The FP32 Gaussian path is left in grid
// order: its blocks are four times smaller, so the count pass (one test per
// block and tile) costs more than the balance gains (measured: the sorted
// 6.9M-query surrogate fallback 43 -> 46 ms); the FP64 kernels gain (1M x 1M
// Epanechnikov product in Z-curve order 4.3 -> 3.7 ms).
bk::BoundArgs live_order(bk_ctx *ctx, Path path, const EvalParams &p, const float *qx32, const float *qy32,
                         const double *qx64, const double *qy64, long long nq, long long r_i, long long r_end,
                         int splits) {
  bk::BoundArgs ba{nullptr, 0};
  if (path != Path::EPAN64 && path != Path::F64) return ba;
  const int qpb = queries_per_item(path);
  const long long nblocks = (nq + qpb - 1) / qpb;
  const long long n = nblocks * splits;
  if (n <= 1 || n > 0x7fffffffLL) return ba;
  const hipStream_t s = ctx->stream;
  const int kind = path == Path::GAUSS32 ? 0 : path == Path::EPAN64 ? 1 : 2;
  ctx->live_boxes.reserve((size_t)nblocks * sizeof(double4));
  ctx->ord_keys_a.reserve((size_t)n * sizeof(unsigned int));
  ctx->bnd_vals.reserve((size_t)n * sizeof(int));
  ctx->bnd_order.reserve((size_t)n * sizeof(int));
  hip_check(bk::launch_live_counts(p, kind, qx32, qy32, qx64, qy64, nq, qpb, ctx->tiles.ptr, r_i, r_end, splits,
                                   ctx->live_boxes.as<double4>(), ctx->ord_keys_a.as<unsigned int>(),
                                   ctx->bnd_vals.as<int>(), s),
            "live count launch");
  bk::order_by_keys(ctx->ord_keys_a.as<unsigned int>(), ctx->bnd_vals.as<int>(), n, ctx->bnd_order.as<int>(),
                    ctx->ord_keys_b, ctx->ord_tmp, s);
  ba.order = ctx->bnd_order.as<int>();
  ba.nblocks = (int)nblocks;
  return ba;
}

// Direct sums over one range pair into device memory `out` (nq doubles).
void eval_range_dev(bk_ctx *ctx, const bk_kernel_desc *desc, size_t r_i, size_t r_j, size_t q_i,
                    size_t q_j, double *out, int normalize, int self_excl) {
  const EvalParams p = params_of(desc);
  check_compatible(ctx, p);
  check_range(ctx, r_i, r_j, q_i, q_j);
  ensure_features(ctx, desc, p);
  const Path path = path_of(ctx, p);
  const long long nq = (long long)(q_j - q_i + 1);
  const long long nr = (long long)(r_j - r_i + 1);
  const long long r_end = (long long)r_j + 1;
  const double scale = path_scale(ctx, path, p, normalize);
  const long long slots =
      path == Path::GAUSS32 ? (long long)bk::gauss_resident_blocks() * bk::device_cu_count() : 0;
  const int splits = choose_splits(nq, nr, queries_per_item(path), tile_of(path), slots);
  const bool spatial = ctx->query_order == BK_QUERY_ORDER_SPATIAL && !self_excl &&
                       (path == Path::GAUSS32 || path == Path::EPAN64 || path == Path::F64) &&
                       nq > queries_per_item(path) && nq < (1LL << 32);
  if (spatial) {
    // evaluate the queries in Z-curve order, then scatter back: each workgroup
    // holds nearby queries and skips the reference tiles out of their reach
    const hipStream_t s = ctx->stream;
    const PointSet &qs = query_set(ctx);
    const size_t csize = path == Path::F64 ? sizeof(double) : sizeof(float);
    ctx->ord_list.reserve((size_t)nq * sizeof(long long));
    ctx->ord_x.reserve((size_t)nq * csize);
    ctx->ord_y.reserve((size_t)nq * csize);
    ctx->ord_out.reserve((size_t)nq * sizeof(double));
    long long *list = ctx->ord_list.as<long long>();
    double *res = ctx->ord_out.as<double>();
    double *dst = res;
    if (splits > 1) {
      ctx->partial.reserve((size_t)splits * nq * sizeof(double));
      dst = ctx->partial.as<double>();
    }
    hipError_t e;
    if (path == Path::F64) {
      const double *qx = qs.xd.as<double>() + q_i, *qy = qs.yd.as<double>() + q_i;
      bk::spatial_sort(qx, qy, nullptr, nq, nq, qs.box_lo, qs.box_hi, list, ctx->ord_keys_a, ctx->ord_keys_b,
                       ctx->ord_tmp, s);
      double *ox = ctx->ord_x.as<double>(), *oy = ctx->ord_y.as<double>();
      hip_check(bk::launch_gather_xy(qx, qy, list, nq, ox, oy, s), "gather launch");
      const bk::BoundArgs ord = live_order(ctx, path, p, nullptr, nullptr, ox, oy, nq, (long long)r_i, r_end, splits);
      e = bk::launch_direct_f64(p, ox, oy, ctx->f0.as<double4>(), ctx->f1.as<double4>(), ctx->tiles.as<double4>(),
                                nullptr, 0, (long long)r_i, r_end, 0, nq, splits, dst, scale, 0,
                                ctx->refs.uniform_abw, s, ord.order ? &ord : nullptr);
      hip_check(e, "direct kernel launch");
      if (splits > 1) hip_check(bk::launch_reduce_splits(dst, splits, nq, res, s), "reduce launch");
      hip_check(bk::launch_scatter_f64(list, nq, res, out, s), "scatter launch");
      return;
    }
    const float *qx = qs.x.as<float>() + q_i, *qy = qs.y.as<float>() + q_i;
    bk::spatial_sort(qx, qy, nullptr, nq, nq, qs.box_lo, qs.box_hi, list, ctx->ord_keys_a, ctx->ord_keys_b,
                     ctx->ord_tmp, s);
    float *ox = ctx->ord_x.as<float>(), *oy = ctx->ord_y.as<float>();
    hip_check(bk::launch_gather_xy(qx, qy, list, nq, ox, oy, s), "gather launch");
    bk::BoundArgs ord = live_order(ctx, path, p, ox, oy, nullptr, nullptr, nq, (long long)r_i, r_end, splits);
    if (ctx->count_terms && (path == Path::GAUSS32 || path == Path::EPAN64))
      ord.count = ctx->term_count.as<unsigned long long>();
    const bk::BoundArgs *op = (ord.order || ord.count) ? &ord : nullptr;
    if (path == Path::GAUSS32)
      e = bk::launch_gauss(p, ox, oy, ctx->f0.as<float4>(), ctx->f1.as<float4>(), ctx->tiles.as<float4>(), nullptr,
                           0, (long long)r_i, r_end, 0, nq, splits, dst, scale, 0,
                           ctx->gauss_form == BK_GAUSS_FORM_DIRECT ? bk::GAUSS_DIRECT : bk::GAUSS_AUTO, s, op);
    else
      e = bk::launch_epan(p, ox, oy, ctx->f0.as<double4>(), ctx->f1.as<double4>(), ctx->tiles.as<double4>(),
                          nullptr, 0, (long long)r_i, r_end, 0, nq, splits, dst, scale, 0, ctx->refs.uniform_abw, s,
                          op);
    hip_check(e, "direct kernel launch");
    if (splits > 1) hip_check(bk::launch_reduce_splits(dst, splits, nq, res, s), "reduce launch");
    hip_check(bk::launch_scatter_f64(list, nq, res, out, s), "scatter launch");
    return;
  }
  double *dst = out;
  if (splits > 1) {
    ctx->partial.reserve((size_t)splits * nq * sizeof(double));
    dst = ctx->partial.as<double>();
  }
  bk::BoundArgs ord{nullptr, 0};
  if (self_excl) {
    // self sums over kd-ordered points cull without any option
    const PointSet &qs = query_set(ctx);
    const bool f64 = path == Path::F64;
    ord = live_order(ctx, path, p, f64 ? nullptr : qs.x.as<float>() + q_i, f64 ? nullptr : qs.y.as<float>() + q_i,
                     f64 ? qs.xd.as<double>() + q_i : nullptr, f64 ? qs.yd.as<double>() + q_i : nullptr, nq,
                     (long long)r_i, r_end, splits);
  }
  if (ctx->count_terms && (path == Path::GAUSS32 || path == Path::EPAN64))
    ord.count = ctx->term_count.as<unsigned long long>();
  hip_check(launch_path(ctx, path, p, ctx_feats(ctx), nullptr, 0, (long long)r_i, r_end, (long long)q_i, nq, splits,
                        dst, scale, self_excl, (ord.order || ord.count) ? &ord : nullptr),
            "direct kernel launch");
  if (splits > 1) hip_check(bk::launch_reduce_splits(dst, splits, nq, out, ctx->stream), "reduce launch");
}

// Tile summaries (box, mass, abw range per aligned tile of kTile
// references) of the current reference set.
void ensure_tile_summary(bk_ctx *ctx) {
  PointSet &r = ctx->refs;
  if (ctx->tsum_version == r.version) return;
  const size_t ntiles = (r.n + bk::kTile - 1) / bk::kTile;
  ctx->tsum.reserve(std::max<size_t>(ntiles, 1) * bk::kTileSummary * sizeof(double4));
  if (ctx->precision == BK_PREC_F64)
    hip_check(bk::launch_tile_summary_f64(r.xd.as<double>(), r.yd.as<double>(), r.md.as<double>(),
                                          r.abwd.as<double>(), r.n, ctx->tsum.as<double4>(), ctx->stream),
              "tile summary launch");
  else
    hip_check(bk::launch_tile_summary_f32(r.x.as<float>(), r.y.as<float>(), r.m.as<float>(), r.abw.as<float>(), r.n,
                                          ctx->tsum.as<double4>(), ctx->stream),
              "tile summary launch");
  ctx->tsum_version = r.version;
}

// Bounded (dual-tree) evaluation of queries q_i..q_j against every
// reference into device arrays lower / upper.
void eval_bounded_dev(bk_ctx *ctx, const bk_kernel_desc *desc, size_t q_i, size_t q_j, double rel_err,
                      double abs_err, double *lower, double *upper, int normalize, bk_bound_stats *stats) {
  const EvalParams p = params_of(desc);
  check_compatible(ctx, p);
  if (ctx->refs.n == 0) fail(BK_ERR_STATE, "no reference points uploaded");
  check_range(ctx, 0, ctx->refs.n - 1, q_i, q_j);
  if (!(rel_err >= 0.0) || !(abs_err >= 0.0) || !std::isfinite(rel_err) || !std::isfinite(abs_err))
    fail(BK_ERR_INVALID_ARG, "rel_err and abs_err must be finite and non-negative");
  ensure_features(ctx, desc, p);
  const Path path = path_of(ctx, p);
  const hipStream_t s = ctx->stream;
  const long long nq = (long long)(q_j - q_i + 1);
  const long long nr = (long long)ctx->refs.n;
  bk_bound_stats st{};
  if (path == Path::GENERAL || path == Path::GRAM) {
    // D > 2: exact sums, the bounds coincide
    eval_range_dev(ctx, desc, 0, ctx->refs.n - 1, q_i, q_j, lower, normalize, 0);
    hip_check(hipMemcpyAsync(upper, lower, (size_t)nq * sizeof(double), hipMemcpyDeviceToDevice, s), "copy bounds");
    const long long tile = tile_of(path);
    st.evaluated = (unsigned long long)(((nq + 255) / 256) * ((nr + tile - 1) / tile));
    if (stats) *stats = st;
    return;
  }
  ensure_tile_summary(ctx);
  const long long ntiles = (nr + bk::kTile - 1) / bk::kTile;
  const long long nwords = (ntiles + 63) / 64;
  const double scale = path_scale(ctx, path, p, normalize);
  const double factor = normalize ? p.norm : 1.0;  // mass units -> output units
  const double abs_u = normalize ? abs_err / p.norm : abs_err;
  const double m_total = ctx->refs.m_sum > 0.0 ? ctx->refs.m_sum : 1.0;
  // tiles are listed per block of qpb queries (kQueriesPerBlock on the
  // Epanechnikov path: its launch gives each listed block several workgroups)
  const int qpb = path == Path::EPAN64 ? bk::kQueriesPerBlock : queries_per_item(path);
  // queries are processed in chunks whose tile masks stay within 2^29 words (4 GiB)
  const long long chunk_blocks = std::max<long long>(1, (1LL << 29) / nwords);
  const long long chunk_q = chunk_blocks * qpb;
  const long long mask_blocks = std::min<long long>(chunk_blocks, (nq + qpb - 1) / qpb);
  ctx->bnd_mask.reserve((size_t)(mask_blocks * nwords) * sizeof(unsigned long long));
  ctx->bnd_lo.reserve((size_t)mask_blocks * sizeof(double));
  ctx->bnd_gap.reserve((size_t)mask_blocks * sizeof(double));
  ctx->bnd_counts.reserve(4 * sizeof(unsigned long long));
  hip_check(hipMemsetAsync(ctx->bnd_counts.ptr, 0, 4 * sizeof(unsigned long long), s), "clear counters");
  const PointSet &qs = query_set(ctx);
  // queries: as stored, or gathered in Z-curve order (FP32 contexts)
  const bool spatial = ctx->query_order == BK_QUERY_ORDER_SPATIAL && nq > qpb && nq < (1LL << 32);
  const float *qx32 = nullptr, *qy32 = nullptr;
  const double *qx64 = nullptr, *qy64 = nullptr;
  long long *list = nullptr;
  long long q_base = (long long)q_i;
  if (path == Path::F64 && spatial) {
    const double *bx = qs.xd.as<double>() + q_i, *by = qs.yd.as<double>() + q_i;
    ctx->ord_list.reserve((size_t)nq * sizeof(long long));
    ctx->ord_x.reserve((size_t)nq * sizeof(double));
    ctx->ord_y.reserve((size_t)nq * sizeof(double));
    list = ctx->ord_list.as<long long>();
    bk::spatial_sort(bx, by, nullptr, nq, nq, qs.box_lo, qs.box_hi, list, ctx->ord_keys_a, ctx->ord_keys_b,
                     ctx->ord_tmp, s);
    hip_check(bk::launch_gather_xy(bx, by, list, nq, ctx->ord_x.as<double>(), ctx->ord_y.as<double>(), s),
              "gather launch");
    qx64 = ctx->ord_x.as<double>();
    qy64 = ctx->ord_y.as<double>();
    q_base = 0;
  } else if (path == Path::F64) {
    qx64 = qs.xd.as<double>();
    qy64 = qs.yd.as<double>();
  } else if (spatial) {
    const float *bx = qs.x.as<float>() + q_i, *by = qs.y.as<float>() + q_i;
    ctx->ord_list.reserve((size_t)nq * sizeof(long long));
    ctx->ord_x.reserve((size_t)nq * sizeof(float));
    ctx->ord_y.reserve((size_t)nq * sizeof(float));
    list = ctx->ord_list.as<long long>();
    bk::spatial_sort(bx, by, nullptr, nq, nq, qs.box_lo, qs.box_hi, list, ctx->ord_keys_a, ctx->ord_keys_b,
                     ctx->ord_tmp, s);
    hip_check(bk::launch_gather_xy(bx, by, list, nq, ctx->ord_x.as<float>(), ctx->ord_y.as<float>(), s),
              "gather launch");
    qx32 = ctx->ord_x.as<float>();
    qy32 = ctx->ord_y.as<float>();
    q_base = 0;
  } else {
    qx32 = qs.x.as<float>();
    qy32 = qs.y.as<float>();
  }
  unsigned long long *mask = ctx->bnd_mask.as<unsigned long long>();
  unsigned long long *counts = ctx->bnd_counts.as<unsigned long long>();
  double *lo_b = ctx->bnd_lo.as<double>(), *gap_b = ctx->bnd_gap.as<double>();
  ctx->bnd_exact.reserve((size_t)nq * sizeof(double));
  double *exact = ctx->bnd_exact.as<double>();
  double *lo_all = lower, *hi_all = upper;
  if (spatial) {
    // bounds in Z-curve order first, back to the caller's order at the end
    ctx->ord_out.reserve((size_t)nq * sizeof(double));
    ctx->bnd_hi.reserve((size_t)nq * sizeof(double));
    lo_all = ctx->ord_out.as<double>();
    hi_all = ctx->bnd_hi.as<double>();
  }
  const long long slots =
      path == Path::GAUSS32 ? (long long)bk::gauss_resident_blocks() * bk::device_cu_count() : 0;
  for (long long c0 = 0; c0 < nq; c0 += chunk_q) {
    const long long cn = std::min(chunk_q, nq - c0);
    const long long q_off = q_base + c0;
    hip_check(bk::launch_bound_pass(p, qx32 ? qx32 + q_off : nullptr, qy32 ? qy32 + q_off : nullptr,
                                    qx64 ? qx64 + q_off : nullptr, qy64 ? qy64 + q_off : nullptr, cn, qpb,
                                    ctx->tsum.as<double4>(), ntiles, rel_err, abs_u, m_total, nwords, mask, lo_b,
                                    gap_b, counts, s),
              "bound pass launch");
    // the summed tiles' exact part; the split partial sums are held within
    // 256 MiB (the pruned evaluation sums few tiles per block, so a few
    // splits keep the device busy and the buffer small)
    const long long split_cap = std::max<long long>(1, (256LL << 20) / (cn * (long long)sizeof(double)));
    const int splits = (int)std::min<long long>(choose_splits(cn, nr, qpb, tile_of(path), slots), split_cap);
    double *dst = exact + c0;
    if (splits > 1) {
      ctx->partial.reserve((size_t)splits * cn * sizeof(double));
      dst = ctx->partial.as<double>();
    }
    bk::BoundArgs ba{mask, nwords};
    // (block, split) pairs with the most listed tiles start first, so the
    // launch does not end on a few long workgroups
    const long long nblocks = (cn + qpb - 1) / qpb;
    if (nblocks * splits > 1 && nblocks * splits <= 0x7fffffffLL) {
      ctx->bnd_order.reserve((size_t)(nblocks * splits) * sizeof(int));
      const long long split_tiles = bk::split_length(0, nr, splits, tile_of(path)) / bk::kTile;
      bk::listed_order(mask, nwords, (int)nblocks, splits, split_tiles, ctx->bnd_order.as<int>(), ctx->ord_keys_a,
                       ctx->ord_keys_b, ctx->bnd_vals, ctx->ord_tmp, s);
      ba.order = ctx->bnd_order.as<int>();
      ba.nblocks = (int)nblocks;
    }
    hipError_t e;
    if (path == Path::GAUSS32)
      e = bk::launch_gauss(p, qx32, qy32, ctx->f0.as<float4>(), ctx->f1.as<float4>(), ctx->tiles.as<float4>(),
                           nullptr, 0, 0, nr, q_off, cn, splits, dst, scale, 0,
                           ctx->gauss_form == BK_GAUSS_FORM_DIRECT ? bk::GAUSS_DIRECT : bk::GAUSS_AUTO, s, &ba);
    else if (path == Path::EPAN64)
      e = bk::launch_epan(p, qx32, qy32, ctx->f0.as<double4>(), ctx->f1.as<double4>(), ctx->tiles.as<double4>(),
                          nullptr, 0, 0, nr, q_off, cn, splits, dst, scale, 0, ctx->refs.uniform_abw, s, &ba);
    else
      e = bk::launch_direct_f64(p, qx64, qy64, ctx->f0.as<double4>(), ctx->f1.as<double4>(),
                                ctx->tiles.as<double4>(), nullptr, 0, 0, nr, q_off, cn, splits, dst, scale, 0,
                                ctx->refs.uniform_abw, s, &ba);
    hip_check(e, "bounded kernel launch");
    if (splits > 1) hip_check(bk::launch_reduce_splits(dst, splits, cn, exact + c0, s), "reduce launch");
    hip_check(bk::launch_bounds_combine(exact + c0, lo_b, gap_b, cn, qpb, factor, lo_all + c0, hi_all + c0, s),
              "bounds launch");
  }
  if (spatial) {
    hip_check(bk::launch_scatter_f64(list, nq, lo_all, lower, s), "scatter launch");
    hip_check(bk::launch_scatter_f64(list, nq, hi_all, upper, s), "scatter launch");
  }
  if (stats) {
    unsigned long long c[4] = {0, 0, 0, 0};
    hip_check(hipMemcpyAsync(c, counts, 4 * sizeof(unsigned long long), hipMemcpyDeviceToHost, s), "read counts");
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    st.excluded = c[0];
    st.approximated = c[1];
    st.evaluated = c[2];
    *stats = st;
  }
}

void copy_to_host(bk_ctx *ctx, double *host, const double *dev, size_t n) {
  if (n == 0) return;
  hip_check(hipMemcpyAsync(host, dev, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream), "copy results");
  hip_check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
}

double *staging_out(bk_ctx *ctx, size_t n) {
  ctx->out.reserve(std::max<size_t>(n, 1) * sizeof(double));
  return ctx->out.as<double>();
}

void require_alias(const bk_ctx *ctx) {
  if (!ctx->alias) fail(BK_ERR_STATE, "this call needs bk_alias_queries_to_refs()");
  if (ctx->refs.n == 0) fail(BK_ERR_STATE, "no reference points uploaded");
}

// FP32 Gaussian sums below 2^-40 of the largest mass are dominated by terms
// whose base-2 exponent is below -40; they are recomputed in FP64 (only
// where a leave-one-out value feeds a logarithm or is returned as such).
constexpr double kRescueFraction = 9.094947017729282e-13;  // 2^-40

void rescue_small(bk_ctx *ctx, const EvalParams &p, double *S_unnormalised, int self_excl) {
  const PointSet &r = ctx->refs;
  const PointSet &q = query_set(ctx);
  const long long n = (long long)q.n;
  ctx->scratch.reserve((size_t)(n + 1) * sizeof(int));
  hip_check(bk::launch_rescue_gauss(p, r.x.as<float>(), r.y.as<float>(), r.m.as<float>(), r.abw.as<float>(),
                                    (long long)r.n, q.x.as<float>(), q.y.as<float>(), 0, n, S_unnormalised,
                                    kRescueFraction * mass_scale(r), self_excl, ctx->scratch.as<int>(),
                                    ctx->stream),
            "rescue launch");
}

// Self evaluation into device buffer `dst` (n_ref doubles, unnormalised
// unless normalize).  Leave-one-out FP32 Gaussian sums are rescued.
void self_eval_dev(bk_ctx *ctx, const bk_kernel_desc *desc, int exclude_self, double *dst, int normalize) {
  require_alias(ctx);
  const EvalParams p = params_of(desc);
  const size_t n = ctx->refs.n;
  if (exclude_self && ctx->precision == BK_PREC_F32 && p.dim <= 2 && p.family == bk::FAM_GAUSS) {
    eval_range_dev(ctx, desc, 0, n - 1, 0, n - 1, dst, 0, 1);
    rescue_small(ctx, p, dst, 1);
    if (normalize) {
      // scale in place through the split reducer (one split = copy with scale)
      hip_check(bk::launch_scale(dst, (long long)n, p.norm, ctx->stream), "scale launch");
    }
    return;
  }
  eval_range_dev(ctx, desc, 0, n - 1, 0, n - 1, dst, normalize, exclude_self ? 1 : 0);
}

double *upload_weights(bk_ctx *ctx, const double *w) {
  if (!w) fail(BK_ERR_INVALID_ARG, "null weight array");
  const size_t n = ctx->refs.n;
  ctx->weights.reserve(n * sizeof(double));
  hip_check(hipMemcpyAsync(ctx->weights.ptr, w, n * sizeof(double), hipMemcpyHostToDevice, ctx->stream),
            "upload weights");
  return ctx->weights.as<double>();
}

double reduce_to_host(bk_ctx *ctx, int mode, const double *w, const double *u, const double *v, double a,
                      double b) {
  ctx->scratch.reserve(bk::kReduceBlocks * sizeof(double));
  ctx->scalar.reserve(sizeof(double));
  hip_check(bk::launch_weighted_reduce(mode, w, u, v, a, b, (long long)ctx->refs.n, ctx->scratch.as<double>(),
                                       ctx->scalar.as<double>(), ctx->stream),
            "reduction launch");
  double r = 0.0;
  copy_to_host(ctx, &r, ctx->scalar.as<double>(), 1);
  return r;
}

}  // namespace

namespace bk {
// Normalised direct sums over every reference of an FP32 2-D context for
// nq device-resident FP32 queries (qx, qy), written to out (device), all on
// stream s (used by the surrogate's sampled check and tile fallback).
void eval_queries_dev(bk_ctx *ctx, const bk_kernel_desc *desc, const float *qx, const float *qy, size_t nq,
                      double *out, hipStream_t s) {
  if (nq == 0) return;
  const EvalParams p = params_of(desc);
  if (ctx->precision != BK_PREC_F32 || ctx->refs.dim != 2 || p.dim != 2)
    fail(BK_ERR_UNSUPPORTED, "device queries need an FP32 context with 2-D references");
  if (ctx->refs.n == 0) fail(BK_ERR_STATE, "no reference points uploaded");
  const hipStream_t saved = ctx->stream;
  ctx->stream = s;
  try {
    ensure_features(ctx, desc, p);
    const Path path = path_of(ctx, p);
    const long long nr = (long long)ctx->refs.n;
    const long long slots =
        path == Path::GAUSS32 ? (long long)bk::gauss_resident_blocks() * bk::device_cu_count() : 0;
    const int splits = choose_splits((long long)nq, nr, queries_per_item(path), tile_of(path), slots);
    double *dst = out;
    if (splits > 1) {
      ctx->partial.reserve((size_t)splits * nq * sizeof(double));
      dst = ctx->partial.as<double>();
    }
    const double scale = path_scale(ctx, path, p, 1);
    // the callers pass queries in Z-curve order: longest workgroups first
    BoundArgs ord = live_order(ctx, path, p, qx, qy, nullptr, nullptr, (long long)nq, 0, nr, splits);
    if (ctx->count_terms && (path == Path::GAUSS32 || path == Path::EPAN64))
      ord.count = ctx->term_count.as<unsigned long long>();
    const BoundArgs *op = (ord.order || ord.count) ? &ord : nullptr;
    hipError_t e;
    if (path == Path::GAUSS32)
      e = launch_gauss(p, qx, qy, ctx->f0.as<float4>(), ctx->f1.as<float4>(), ctx->tiles.as<float4>(), nullptr, 0, 0,
                       nr, 0, (long long)nq, splits, dst, scale, 0,
                       ctx->gauss_form == BK_GAUSS_FORM_DIRECT ? GAUSS_DIRECT : GAUSS_AUTO, s, op);
    else
      e = launch_epan(p, qx, qy, ctx->f0.as<double4>(), ctx->f1.as<double4>(), ctx->tiles.as<double4>(), nullptr, 0,
                      0, nr, 0, (long long)nq, splits, dst, scale, 0, ctx->refs.uniform_abw, s, op);
    hip_check(e, "direct kernel launch");
    if (splits > 1) hip_check(launch_reduce_splits(dst, splits, (long long)nq, out, s), "reduce launch");
  } catch (...) {
    ctx->stream = saved;
    throw;
  }
  ctx->stream = saved;
}
}  // namespace bk

// ---- allocation-free range evaluation ------------------------------------------
// bk_eval_range_device_noalloc keeps the reference features and the split
// partial sums of a call in one file-scope device array per GPU instead of
// buffers the context allocates.  Host code never passes the array's name to
// a kernel: its device address is looked up once per device (in bk_create,
// with that device current) and kept in scratch_ptr[device].
namespace {

constexpr size_t kScratchBytes = size_t(64) << 20;  // 64 MiB per device
constexpr int kScratchDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex scratch_mu;
void *scratch_ptr[kScratchDevices] = {};

// Device address of `device`'s scratch array; `device` must be current.
void *scratch_for(int device) {
  if (device < 0 || device >= kScratchDevices)
    fail(BK_ERR_UNSUPPORTED, "device index " + std::to_string(device) + " beyond the scratch table");
  std::lock_guard<std::mutex> g(scratch_mu);
  if (!scratch_ptr[device]) {
    void *ptr = nullptr;
    hip_check(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    if (!ptr) fail(BK_ERR_HIP, "hipGetSymbolAddress(scratch) returned a null address");
    scratch_ptr[device] = ptr;
  }
  return scratch_ptr[device];
}

size_t round_up_256(size_t b) { return (b + 255) / 256 * 256; }

// As eval_range_dev (queries in their stored order, no self exclusion), with
// every buffer beyond the context's points and `out` inside the scratch: the
// features of the whole reference set first, the split partial sums after
// them (the split count is lowered until they fit; one split writes `out`
// directly).  Nothing here allocates, and nothing needs clearing: the
// features and every partial sum are written before they are read.
void eval_range_noalloc(bk_ctx *ctx, const bk_kernel_desc *desc, size_t r_i, size_t r_j, size_t q_i, size_t q_j,
                        double *out, int normalize) {
  const EvalParams p = params_of(desc);
  check_compatible(ctx, p);
  check_range(ctx, r_i, r_j, q_i, q_j);
  const Path path = path_of(ctx, p);
  const FeatBytes fb = feature_bytes(path, ctx->refs.n);
  const size_t feat = round_up_256(fb.f0) + round_up_256(fb.f1) + round_up_256(fb.tiles);
  if (feat > kScratchBytes)
    fail(BK_ERR_INVALID_ARG, "bk_eval_range_device_noalloc: the features of " + std::to_string(ctx->refs.n) +
                                 " references need " + std::to_string(feat) + " bytes, more than the " +
                                 std::to_string(kScratchBytes) + "-byte scratch (use bk_eval_range_device)");
  char *base = static_cast<char *>(scratch_for(ctx->device));
  FeatPtrs fp;
  fp.f0 = base;
  fp.f1 = base + round_up_256(fb.f0);
  fp.tiles = base + round_up_256(fb.f0) + round_up_256(fb.f1);
  launch_prep(ctx, path, p, fp);
  const long long nq = (long long)(q_j - q_i + 1);
  const long long nr = (long long)(r_j - r_i + 1);
  const long long r_end = (long long)r_j + 1;
  const double scale = path_scale(ctx, path, p, normalize);
  const long long slots =
      path == Path::GAUSS32 ? (long long)bk::gauss_resident_blocks() * bk::device_cu_count() : 0;
  long long splits = choose_splits(nq, nr, queries_per_item(path), tile_of(path), slots);
  const long long fit = (long long)((kScratchBytes - feat) / ((size_t)nq * sizeof(double)));
  if (splits > fit) splits = fit >= 2 ? fit : 1;
  double *dst = splits > 1 ? reinterpret_cast<double *>(base + feat) : out;
  hip_check(launch_path(ctx, path, p, fp, nullptr, 0, (long long)r_i, r_end, (long long)q_i, nq, (int)splits, dst,
                        scale, 0),
            "direct kernel launch");
  if (splits > 1) hip_check(bk::launch_reduce_splits(dst, (int)splits, nq, out, ctx->stream), "reduce launch");
}

}  // namespace

// ---- C ABI -----------------------------------------------------------------

extern "C" {

const char *bk_last_error(void) { return bk::last_error().c_str(); }

int bk_create(bk_ctx **ctx_out, int device, int precision) {
  return guarded([&] {
    if (!ctx_out) fail(BK_ERR_INVALID_ARG, "null output pointer");
    *ctx_out = nullptr;
    if (precision != BK_PREC_F32 && precision != BK_PREC_F64)
      fail(BK_ERR_INVALID_ARG, "precision must be BK_PREC_F32 or BK_PREC_F64");
    int count = 0;
    hip_check(hipGetDeviceCount(&count), "hipGetDeviceCount");
    if (device < 0 || device >= count)
      fail(BK_ERR_INVALID_ARG, "device " + std::to_string(device) + " not present (" +
                                   std::to_string(count) + " devices)");
    DeviceGuard g(device);
    bk_ctx *c = new bk_ctx();
    c->device = device;
    c->precision = precision;
    if (hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking) != hipSuccess) {
      delete c;
      fail(BK_ERR_HIP, "hipStreamCreateWithFlags failed");
    }
    c->stream = c->own_stream;
    try {
      (void)scratch_for(device);  // the device address of this GPU's scratch array
    } catch (...) {
      (void)hipStreamDestroy(c->own_stream);
      delete c;
      throw;
    }
    *ctx_out = c;
  });
}

int bk_destroy(bk_ctx *ctx) {
  return guarded([&] {
    if (!ctx) return;
    {
      DeviceGuard g(ctx->device);
      (void)hipStreamSynchronize(ctx->stream);
      ctx->refs.release();
      ctx->queries.release();
      for (DevBuf *b : {&ctx->f0, &ctx->f1, &ctx->tiles, &ctx->partial, &ctx->out, &ctx->items, &ctx->work_a, &ctx->work_b,
                        &ctx->weights, &ctx->scratch, &ctx->scalar, &ctx->ord_list, &ctx->ord_keys_a,
                        &ctx->ord_keys_b, &ctx->ord_tmp, &ctx->ord_x, &ctx->ord_y, &ctx->ord_out, &ctx->tsum,
                        &ctx->bnd_mask, &ctx->bnd_lo, &ctx->bnd_gap, &ctx->bnd_counts, &ctx->bnd_exact,
                        &ctx->bnd_hi, &ctx->bnd_vals, &ctx->bnd_order, &ctx->live_boxes, &ctx->term_count})
        b->release();
      if (ctx->own_stream) (void)hipStreamDestroy(ctx->own_stream);
    }
    delete ctx;
  });
}

int bk_set_stream(bk_ctx *ctx, void *stream) {
  return guarded([&] {
    require_ctx(ctx);
    ctx->stream = stream ? static_cast<hipStream_t>(stream) : ctx->own_stream;
  });
}

int bk_set_option(bk_ctx *ctx, int option, int value) {
  return guarded([&] {
    require_ctx(ctx);
    switch (option) {
      case BK_OPT_GAUSS_FORM:
        if (value != BK_GAUSS_FORM_AUTO && value != BK_GAUSS_FORM_DIRECT)
          fail(BK_ERR_INVALID_ARG, "BK_OPT_GAUSS_FORM takes BK_GAUSS_FORM_AUTO or BK_GAUSS_FORM_DIRECT");
        ctx->gauss_form = value;
        break;
      case BK_OPT_HIGH_DIM_FORM:
        if (value != BK_HIGH_DIM_AUTO && value != BK_HIGH_DIM_DIRECT)
          fail(BK_ERR_INVALID_ARG, "BK_OPT_HIGH_DIM_FORM takes BK_HIGH_DIM_AUTO or BK_HIGH_DIM_DIRECT");
        ctx->high_dim_form = value;
        break;
      case BK_OPT_QUERY_ORDER:
        if (value != BK_QUERY_ORDER_GIVEN && value != BK_QUERY_ORDER_SPATIAL)
          fail(BK_ERR_INVALID_ARG, "BK_OPT_QUERY_ORDER takes BK_QUERY_ORDER_GIVEN or BK_QUERY_ORDER_SPATIAL");
        ctx->query_order = value;
        break;
      case BK_OPT_COUNT_TERMS:
        if (value != 0 && value != 1) fail(BK_ERR_INVALID_ARG, "BK_OPT_COUNT_TERMS takes 0 or 1");
        if (value) {
          DeviceGuard g(ctx->device);
          ctx->term_count.reserve(2 * sizeof(unsigned long long));
          hip_check(hipMemsetAsync(ctx->term_count.ptr, 0, 2 * sizeof(unsigned long long), ctx->stream),
                    "clear term counters");
        }
        ctx->count_terms = value;
        break;
      default:
        fail(BK_ERR_INVALID_ARG, "unknown option " + std::to_string(option));
    }
  });
}

int bk_upload(bk_ctx *ctx, int set, int dim, size_t n, const double *coords, const double *mass,
              const double *abw) {
  return guarded([&] {
    require_ctx(ctx);
    DeviceGuard g(ctx->device);
    PointSet &ps = set_of(ctx, set);
    upload_set(ctx, ps, dim, n, coords, mass, abw);
    if (set == BK_SET_REF) ctx->feat_valid = false;
  });
}

int bk_alias_queries_to_refs(bk_ctx *ctx) {
  return guarded([&] {
    require_ctx(ctx);
    DeviceGuard g(ctx->device);
    ctx->queries.release();
    ctx->alias = true;
  });
}

int bk_size(const bk_ctx *ctx, int set, size_t *n_out) {
  return guarded([&] {
    require_ctx(ctx);
    if (!n_out) fail(BK_ERR_INVALID_ARG, "null output pointer");
    if (set == BK_SET_REF)
      *n_out = ctx->refs.n;
    else if (set == BK_SET_QUERY)
      *n_out = query_set(ctx).n;
    else
      fail(BK_ERR_INVALID_ARG, "set must be BK_SET_REF or BK_SET_QUERY");
  });
}

int bk_update_refs(bk_ctx *ctx, const double *mass, const double *abw) {
  return guarded([&] {
    require_ctx(ctx);
    DeviceGuard g(ctx->device);
    update_refs(ctx, mass, abw);
  });
}

int bk_term_count(bk_ctx *ctx, unsigned long long *terms_out, unsigned long long *query_terms_out) {
  return guarded([&] {
    require_ctx(ctx);
    if (!ctx->count_terms) fail(BK_ERR_STATE, "term counting is off (BK_OPT_COUNT_TERMS)");
    DeviceGuard g(ctx->device);
    unsigned long long c[2] = {0, 0};
    hip_check(hipMemcpyAsync(c, ctx->term_count.ptr, sizeof(c), hipMemcpyDeviceToHost, ctx->stream), "read counters");
    hip_check(hipMemsetAsync(ctx->term_count.ptr, 0, sizeof(c), ctx->stream), "clear term counters");
    hip_check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    if (terms_out) *terms_out = c[0];
    if (query_terms_out) *query_terms_out = c[1];
  });
}

int bk_eval_range(bk_ctx *ctx, const bk_kernel_desc *desc, size_t r_i, size_t r_j, size_t q_i, size_t q_j,
                  double *out_host, int normalize) {
  return guarded([&] {
    require_ctx(ctx);
    if (!out_host) fail(BK_ERR_INVALID_ARG, "null output array");
    DeviceGuard g(ctx->device);
    check_range(ctx, r_i, r_j, q_i, q_j);
    const size_t nq = q_j - q_i + 1;
    double *dev = staging_out(ctx, nq);
    eval_range_dev(ctx, desc, r_i, r_j, q_i, q_j, dev, normalize, 0);
    copy_to_host(ctx, out_host, dev, nq);
  });
}

int bk_eval_range_device(bk_ctx *ctx, const bk_kernel_desc *desc, size_t r_i, size_t r_j, size_t q_i,
                         size_t q_j, double *out_dev, int normalize) {
  return guarded([&] {
    require_ctx(ctx);
    if (!out_dev) fail(BK_ERR_INVALID_ARG, "null output array");
    DeviceGuard g(ctx->device);
    eval_range_dev(ctx, desc, r_i, r_j, q_i, q_j, out_dev, normalize, 0);
  });
}

int bk_eval_range_device_noalloc(bk_ctx *ctx, const bk_kernel_desc *desc, size_t r_i, size_t r_j, size_t q_i,
                                 size_t q_j, double *out_dev, int normalize) {
  return guarded([&] {
    require_ctx(ctx);
    if (!out_dev) fail(BK_ERR_INVALID_ARG, "null output array");
    DeviceGuard g(ctx->device);
    eval_range_noalloc(ctx, desc, r_i, r_j, q_i, q_j, out_dev, normalize);
  });
}

int bk_eval_bounded_device(bk_ctx *ctx, const bk_kernel_desc *desc, size_t q_i, size_t q_j, double rel_err,
                           double abs_err, double *lower_dev, double *upper_dev, int normalize,
                           bk_bound_stats *stats) {
  return guarded([&] {
    require_ctx(ctx);
    if (!lower_dev || !upper_dev) fail(BK_ERR_INVALID_ARG, "null output array");
    DeviceGuard g(ctx->device);
    eval_bounded_dev(ctx, desc, q_i, q_j, rel_err, abs_err, lower_dev, upper_dev, normalize, stats);
  });
}

int bk_eval_bounded(bk_ctx *ctx, const bk_kernel_desc *desc, size_t q_i, size_t q_j, double rel_err,
                    double abs_err, double *lower_host, double *upper_host, int normalize, bk_bound_stats *stats) {
  return guarded([&] {
    require_ctx(ctx);
    if (!lower_host || !upper_host) fail(BK_ERR_INVALID_ARG, "null output array");
    DeviceGuard g(ctx->device);
    if (ctx->refs.n > 0) check_range(ctx, 0, ctx->refs.n - 1, q_i, q_j);
    const size_t nq = q_j >= q_i ? q_j - q_i + 1 : 0;
    double *lo = staging_out(ctx, 2 * std::max<size_t>(nq, 1));
    eval_bounded_dev(ctx, desc, q_i, q_j, rel_err, abs_err, lo, lo + nq, normalize, stats);
    copy_to_host(ctx, lower_host, lo, nq);
    copy_to_host(ctx, upper_host, lo + nq, nq);
  });
}

int bk_eval_leaf_pairs(bk_ctx *ctx, const bk_kernel_desc *desc, const bk_range_pair *pairs, size_t n_pairs,
                       double *out_host, int normalize) {
  return guarded([&] {
    require_ctx(ctx);
    if (n_pairs == 0) return;
    if (!pairs || !out_host) fail(BK_ERR_INVALID_ARG, "null pair or output array");
    DeviceGuard g(ctx->device);
    const EvalParams p = params_of(desc);
    check_compatible(ctx, p);
    const Path path = path_of(ctx, p);
    const int qpb = queries_per_item(path);
    std::vector<WorkItem> items;
    long long total = 0;
    for (size_t k = 0; k < n_pairs; ++k) {
      const bk_range_pair &pr = pairs[k];
      check_range(ctx, pr.r_i, pr.r_j, pr.q_i, pr.q_j);
      const long long nq = (long long)(pr.q_j - pr.q_i + 1);
      for (long long qb = 0; qb < nq; qb += qpb) {
        WorkItem it{};
        it.r_begin = (long long)pr.r_i;
        it.r_end = (long long)pr.r_j + 1;
        it.q_begin = (long long)pr.q_i + qb;
        it.q_count = (int)std::min<long long>(qpb, nq - qb);
        it.out_index = total + qb;
        items.push_back(it);
      }
      total += nq;
    }
    if (items.size() > 0x7fffffffULL) fail(BK_ERR_INVALID_ARG, "too many leaf pairs in one call");
    ensure_features(ctx, desc, p);
    const hipStream_t s = ctx->stream;
    ctx->items.reserve(items.size() * sizeof(WorkItem));
    hip_check(hipMemcpyAsync(ctx->items.ptr, items.data(), items.size() * sizeof(WorkItem), hipMemcpyHostToDevice,
                             s),
              "upload work items");
    double *dev = staging_out(ctx, (size_t)total);
    const hipError_t e = launch_path(ctx, path, p, ctx_feats(ctx), ctx->items.as<WorkItem>(), (int)items.size(), 0,
                                     0, 0, 0, 1, dev, path_scale(ctx, path, p, normalize), 0);
    hip_check(e, "leaf-pair kernel launch");
    copy_to_host(ctx, out_host, dev, (size_t)total);
  });
}

int bk_self_eval(bk_ctx *ctx, const bk_kernel_desc *desc, int exclude_self, double *out_host, int normalize) {
  return guarded([&] {
    require_ctx(ctx);
    if (!out_host) fail(BK_ERR_INVALID_ARG, "null output array");
    DeviceGuard g(ctx->device);
    require_alias(ctx);
    const size_t n = ctx->refs.n;
    double *dev = staging_out(ctx, n);
    self_eval_dev(ctx, desc, exclude_self, dev, normalize);
    copy_to_host(ctx, out_host, dev, n);
  });
}

int bk_likelihood_cv(bk_ctx *ctx, const bk_kernel_desc *desc, const double *weights, double *score_out) {
  return guarded([&] {
    require_ctx(ctx);
    if (!score_out) fail(BK_ERR_INVALID_ARG, "null output pointer");
    DeviceGuard g(ctx->device);
    require_alias(ctx);
    const size_t n = ctx->refs.n;
    const EvalParams p = params_of(desc);
    double *w = upload_weights(ctx, weights);
    ctx->work_a.reserve(n * sizeof(double));
    double *S = ctx->work_a.as<double>();
    self_eval_dev(ctx, desc, 1, S, 0);
    *score_out = reduce_to_host(ctx, 0, w, S, nullptr, p.norm, 0.0);
  });
}

namespace {
// Convolution partner of a kernel kind (ConvKernelAssociator,
// include/Kernels/ConvKernelAssociator.h:17-84).
bk_kernel_desc conv_partner(const bk_kernel_desc *d) {
  bk_kernel_desc c = *d;
  switch (d->kind) {
    case BK_EPANECHNIKOV:
      if (d->dim != 1) fail(BK_ERR_UNSUPPORTED, "no convolution kernel for EpanechnikovKernel with D != 1");
      c.kind = BK_EPANECHNIKOV_CONV_1D;
      break;
    case BK_GAUSSIAN:
      if (d->dim != 1) fail(BK_ERR_UNSUPPORTED, "no convolution kernel for GaussianKernel with D != 1");
      c.kind = BK_GAUSSIAN_CONV_1D;
      break;
    case BK_EPANECHNIKOV_PRODUCT_2D:
      c.kind = BK_EPANECHNIKOV_PRODUCT_CONV_2D;
      break;
    case BK_GAUSSIAN_PRODUCT_2D:
      c.kind = BK_GAUSSIAN_PRODUCT_CONV_2D;
      break;
    default:
      fail(BK_ERR_UNSUPPORTED, "kernel kind has no convolution kernel");
  }
  return c;
}
}  // namespace

int bk_lsq_conv_cv(bk_ctx *ctx, const bk_kernel_desc *desc, const double *weights, double *score_out) {
  return guarded([&] {
    require_ctx(ctx);
    if (!score_out) fail(BK_ERR_INVALID_ARG, "null output pointer");
    DeviceGuard g(ctx->device);
    require_alias(ctx);
    const size_t n = ctx->refs.n;
    const bk_kernel_desc cdesc = conv_partner(desc);
    const EvalParams p = params_of(desc);
    const EvalParams pc = params_of(&cdesc);
    double *w = upload_weights(ctx, weights);
    ctx->work_a.reserve(n * sizeof(double));
    ctx->work_b.reserve(n * sizeof(double));
    double *S = ctx->work_a.as<double>();
    double *Sc = ctx->work_b.as<double>();
    self_eval_dev(ctx, desc, 1, S, 0);
    self_eval_dev(ctx, &cdesc, 0, Sc, 0);
    *score_out = reduce_to_host(ctx, 1, w, Sc, S, pc.norm, 2.0 * p.norm);
  });
}

int bk_adapt(bk_ctx *ctx, const bk_kernel_desc *desc, double alpha, const double *weights, double *g_out,
             double *lambda_out) {
  return guarded([&] {
    require_ctx(ctx);
    DeviceGuard g(ctx->device);
    require_alias(ctx);
    if (!(alpha >= 0.0 && alpha <= 1.0)) fail(BK_ERR_INVALID_ARG, "adapt_density: alpha must be between [0, 1]");
    if (!weights) fail(BK_ERR_INVALID_ARG, "null weight array");
    const size_t n = ctx->refs.n;
    const EvalParams p = params_of(desc);
    check_compatible(ctx, p);
    // non-adaptive state: mass = weight, abw = 1
    std::vector<double> ones(n, 1.0);
    update_refs(ctx, weights, ones.data());
    if (alpha == 0.0) {
      if (lambda_out) std::fill(lambda_out, lambda_out + n, 1.0);
      if (g_out) *g_out = std::nan("");
      return;
    }
    double *w = upload_weights(ctx, weights);
    ctx->work_a.reserve(n * sizeof(double));
    ctx->work_b.reserve(n * sizeof(double));
    double *S = ctx->work_a.as<double>();
    double *lam = ctx->work_b.as<double>();
    self_eval_dev(ctx, desc, 0, S, 0);
    ctx->scratch.reserve(bk::kReduceBlocks * sizeof(double));
    ctx->scalar.reserve(sizeof(double));
    hip_check(bk::launch_weighted_reduce(0, w, S, nullptr, p.norm, 0.0, (long long)n, ctx->scratch.as<double>(),
                                         ctx->scalar.as<double>(), ctx->stream),
              "reduction launch");
    hip_check(bk::launch_adapt_lambda(S, (long long)n, p.norm, ctx->scalar.as<double>(), alpha, lam, ctx->stream),
              "lambda launch");
    std::vector<double> lambda(n);
    double log_g = 0.0;
    copy_to_host(ctx, lambda.data(), lam, n);
    copy_to_host(ctx, &log_g, ctx->scalar.as<double>(), 1);
    std::vector<double> mass(n);
    for (size_t i = 0; i < n; ++i) mass[i] = weights[i] * std::pow(lambda[i], -(double)p.dim);
    update_refs(ctx, mass.data(), lambda.data());
    if (lambda_out) std::copy(lambda.begin(), lambda.end(), lambda_out);
    if (g_out) *g_out = std::exp(log_g);
  });
}

}  // extern "C"
