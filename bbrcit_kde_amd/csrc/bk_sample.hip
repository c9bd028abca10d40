// Batched sampling from a kernel density (the device part of
// KernelDensity::simulate, include/KernelDensityImpl.h:99-138 in the reference).
//
// A sample k picks the reference i_k, the first index whose cumulative
// weight exceeds a uniform draw u_k (std::upper_bound over the cumulative
// weights, the last index if none does), and moves it by the kernel's draw
// scaled by the reference's bandwidth factor:
//     out[k][d] = x[i_k][d] + (FT) ((a_{i_k} h_d) z[k][d])
// with the product in the kernel's float type (float kernels: the FP32
// product of the FP32-rounded factor, bandwidth and draw, as the reference's
// kernels compute a * h * draw in their own type) and the sum in FP64.  The
// draws u and z come from the caller, so the random stream is the caller's
// engine; the device does the index search (a binary search per sample over
// the resident cumulative weights) and the combination, one thread per
// sample, with explicitly rounded operations (no contraction into FMAs).
#include "../../include/bbrcit_hip.h"
#include "bk_common.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <string>

struct bk_sampler {
  int device = 0;
  int dim = 0;
  size_t n = 0;
  hipStream_t stream = nullptr;
  bk::DevBuf coords, cum, abw, u, z, out, index;
};

namespace {

using bk::DevBuf;
using bk::DeviceGuard;
using bk::fail;
using bk::guarded;
using bk::hip_check;

constexpr int kSampleThreads = 256;

// First index in [0, n) with cum[i] > u, or n - 1 when there is none.
__device__ __forceinline__ long long upper_bound_index(const double *__restrict__ cum, long long n, double u) {
  long long lo = 0, len = n;
  while (len > 0) {
    const long long half = len >> 1;
    const long long mid = lo + half;
    if (!(u < cum[mid])) {
      lo = mid + 1;
      len -= half + 1;
    } else {
      len = half;
    }
  }
  return lo < n ? lo : n - 1;
}

template <bool KF32>
__global__ __launch_bounds__(kSampleThreads) void sample_kernel(const double *__restrict__ coords,
                                                                const double *__restrict__ cum,
                                                                const double *__restrict__ abw, long long n, int dim,
                                                                double h0, double h1, const double *__restrict__ u,
                                                                const double *__restrict__ z, long long ns,
                                                                double *__restrict__ out,
                                                                long long *__restrict__ index) {
  // each product and the sum rounded on its own: this file is compiled with
  // -ffp-contract=off (hipcc would otherwise fuse the product and the sum
  // into an FMA, which the reference's host arithmetic does not)
  const long long k = (long long)blockIdx.x * kSampleThreads + threadIdx.x;
  if (k >= ns) return;
  const long long i = upper_bound_index(cum, n, u[k]);
  if (index) index[k] = i;
  for (int d = 0; d < dim; ++d) {
    const double h = d == 0 ? h0 : h1;
    double q;
    if constexpr (KF32) {
      const float ah = __fmul_rn((float)abw[i], (float)h);
      q = (double)__fmul_rn(ah, (float)z[k * dim + d]);
    } else {
      q = __dmul_rn(__dmul_rn(abw[i], h), z[k * dim + d]);
    }
    out[k * dim + d] = __dadd_rn(coords[i * dim + d], q);
  }
}

void copy_in(bk_sampler *s, DevBuf &b, const double *host, size_t count, const char *what) {
  b.reserve(count * sizeof(double));
  hip_check(hipMemcpyAsync(b.ptr, host, count * sizeof(double), hipMemcpyHostToDevice, s->stream), what);
}

}  // namespace

extern "C" {

int bk_sampler_create(int device, int dim, size_t n, const double *coords, const double *cum_weights,
                      const double *abw, bk_sampler **out) {
  return guarded([&] {
    if (!out) fail(BK_ERR_INVALID_ARG, "null output pointer");
    *out = nullptr;
    if (dim < 1 || dim > 2) fail(BK_ERR_UNSUPPORTED, "sampling supports dimensions 1 and 2");
    if (n == 0 || !coords || !cum_weights || !abw) fail(BK_ERR_INVALID_ARG, "empty or null point arrays");
    for (size_t i = 1; i < n; ++i)
      if (!(cum_weights[i] >= cum_weights[i - 1])) fail(BK_ERR_INVALID_ARG, "cumulative weights must not decrease");
    int count = 0;
    hip_check(hipGetDeviceCount(&count), "hipGetDeviceCount");
    if (device < 0 || device >= count) fail(BK_ERR_INVALID_ARG, "device " + std::to_string(device) + " not present");
    DeviceGuard g(device);
    bk_sampler *s = new bk_sampler();
    s->device = device;
    s->dim = dim;
    s->n = n;
    try {
      hip_check(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking), "hipStreamCreateWithFlags");
      copy_in(s, s->coords, coords, n * dim, "upload coordinates");
      copy_in(s, s->cum, cum_weights, n, "upload cumulative weights");
      copy_in(s, s->abw, abw, n, "upload abw");
      hip_check(hipStreamSynchronize(s->stream), "hipStreamSynchronize");
    } catch (...) {
      for (DevBuf *b : {&s->coords, &s->cum, &s->abw}) b->release();
      if (s->stream) (void)hipStreamDestroy(s->stream);
      delete s;
      throw;
    }
    *out = s;
  });
}

int bk_sampler_destroy(bk_sampler *s) {
  return guarded([&] {
    if (!s) return;
    {
      DeviceGuard g(s->device);
      (void)hipStreamSynchronize(s->stream);
      for (DevBuf *b : {&s->coords, &s->cum, &s->abw, &s->u, &s->z, &s->out, &s->index}) b->release();
      (void)hipStreamDestroy(s->stream);
    }
    delete s;
  });
}

int bk_sampler_draw(bk_sampler *s, int kernel_float, const double *h, size_t n_samples, const double *u,
                    const double *z, double *out, long long *index_out) {
  return guarded([&] {
    if (!s) fail(BK_ERR_INVALID_ARG, "null sampler");
    if (n_samples == 0) return;
    if (!h || !u || !z || !out) fail(BK_ERR_INVALID_ARG, "null draw or output array");
    if (kernel_float != 0 && kernel_float != 1) fail(BK_ERR_INVALID_ARG, "kernel_float must be 0 or 1");
    if (n_samples > (size_t(1) << 40)) fail(BK_ERR_INVALID_ARG, "too many samples in one call");
    DeviceGuard g(s->device);
    const size_t nd = n_samples * (size_t)s->dim;
    copy_in(s, s->u, u, n_samples, "upload uniform draws");
    copy_in(s, s->z, z, nd, "upload kernel draws");
    s->out.reserve(nd * sizeof(double));
    long long *idx = nullptr;
    if (index_out) {
      s->index.reserve(n_samples * sizeof(long long));
      idx = s->index.as<long long>();
    }
    const double h0 = h[0], h1 = s->dim > 1 ? h[1] : h[0];
    const unsigned blocks = (unsigned)((n_samples + kSampleThreads - 1) / kSampleThreads);
    if (kernel_float)
      sample_kernel<true><<<blocks, kSampleThreads, 0, s->stream>>>(
          s->coords.as<double>(), s->cum.as<double>(), s->abw.as<double>(), (long long)s->n, s->dim, h0, h1,
          s->u.as<double>(), s->z.as<double>(), (long long)n_samples, s->out.as<double>(), idx);
    else
      sample_kernel<false><<<blocks, kSampleThreads, 0, s->stream>>>(
          s->coords.as<double>(), s->cum.as<double>(), s->abw.as<double>(), (long long)s->n, s->dim, h0, h1,
          s->u.as<double>(), s->z.as<double>(), (long long)n_samples, s->out.as<double>(), idx);
    hip_check(hipGetLastError(), "sample kernel launch");
    hip_check(hipMemcpyAsync(out, s->out.ptr, nd * sizeof(double), hipMemcpyDeviceToHost, s->stream),
              "read samples");
    if (index_out)
      hip_check(hipMemcpyAsync(index_out, idx, n_samples * sizeof(long long), hipMemcpyDeviceToHost, s->stream),
                "read indices");
    hip_check(hipStreamSynchronize(s->stream), "hipStreamSynchronize");
  });
}

}  // extern "C"
