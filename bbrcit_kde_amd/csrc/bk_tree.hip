// Device steps of the kd-tree build (the reference's include/Kdtree.h:337-471):
//   * bk_lexico_runs: the lexicographic sort of a point set, the runs of
//     exactly equal points in it (the duplicates that construction merges)
//     and the root box, in one device pass over the uploaded coordinates;
//   * bk_run_merge: the merge of the attributes of runs of duplicates, one
//     thread per run, left to right in the run's order, with the operations
//     the host merge uses (a sum, or std::min / std::max), so the merged
//     values are bit-identical to the host loop's.
// The partition (std::nth_element at the lower median, whose permutation is
// part of the API: the tree order of the points) stays on the host.
#include "../../include/bbrcit_hip.h"
#include "bk_common.h"
#include "bk_internal.h"

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <cfloat>
#include <string>
#include <vector>

namespace {

using bk::DevBuf;
using bk::DeviceGuard;
using bk::fail;
using bk::guarded;
using bk::hip_check;

constexpr int kThreads = 256;

// std::min / std::max as the host merge computes them
__device__ __forceinline__ double host_min(double a, double b) { return (b < a) ? b : a; }
__device__ __forceinline__ double host_max(double a, double b) { return (a < b) ? b : a; }

// flag[k] = 1 when the k-th point in sorted order starts a run of equal points
__global__ void run_flags_kernel(const double *__restrict__ coords, int dim, const long long *__restrict__ order,
                                 long long n, int *__restrict__ flag) {
  const long long k = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (k >= n) return;
  int f = 1;
  if (k > 0) {
    const double *a = coords + order[k] * dim;
    const double *b = coords + order[k - 1] * dim;
    f = 0;
    for (int d = 0; d < dim; ++d)
      if (!(a[d] == b[d])) f = 1;
  }
  flag[k] = f;
}

// start[u] = sorted position of the first point of run u (scan = inclusive sum of the flags)
__global__ void run_starts_kernel(const int *__restrict__ flag, const int *__restrict__ scan, long long n,
                                  long long *__restrict__ start) {
  const long long k = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (k < n && flag[k]) start[scan[k] - 1] = k;
}

// One thread per run: the run's first values merged with each next member's,
// in order; the members of run u are rows off[u] .. off[u + 1] - 1.
__global__ void run_merge_kernel(const long long *__restrict__ off, long long n_runs, int n_sum,
                                 const double *__restrict__ sums, int n_min, const double *__restrict__ mins,
                                 int n_max, const double *__restrict__ maxs, double *__restrict__ sums_out,
                                 double *__restrict__ mins_out, double *__restrict__ maxs_out) {
  const long long u = (long long)blockIdx.x * kThreads + threadIdx.x;
  if (u >= n_runs) return;
  const long long b = off[u], e = off[u + 1];
  for (int j = 0; j < n_sum; ++j) {
    double acc = sums[b * n_sum + j];
    for (long long k = b + 1; k < e; ++k) acc = acc + sums[k * n_sum + j];
    sums_out[u * n_sum + j] = acc;
  }
  for (int j = 0; j < n_min; ++j) {
    double acc = mins[b * n_min + j];
    for (long long k = b + 1; k < e; ++k) acc = host_min(acc, mins[k * n_min + j]);
    mins_out[u * n_min + j] = acc;
  }
  for (int j = 0; j < n_max; ++j) {
    double acc = maxs[b * n_max + j];
    for (long long k = b + 1; k < e; ++k) acc = host_max(acc, maxs[k * n_max + j]);
    maxs_out[u * n_max + j] = acc;
  }
}

// Per workgroup and coordinate: the minimum and maximum of its points,
// starting from DBL_MAX and DBL_MIN (the smallest positive double) as the
// host box does; part[(block * dim + d) * 2 + {0, 1}].
__global__ void box_partial_kernel(const double *__restrict__ coords, int dim, long long n,
                                   double *__restrict__ part) {
  __shared__ double slo[kThreads], shi[kThreads];
  for (int d = 0; d < dim; ++d) {
    double lo = DBL_MAX, hi = DBL_MIN;
    for (long long k = (long long)blockIdx.x * kThreads + threadIdx.x; k < n; k += (long long)gridDim.x * kThreads) {
      const double v = coords[k * dim + d];
      lo = host_min(v, lo);
      hi = host_max(v, hi);
    }
    slo[threadIdx.x] = lo;
    shi[threadIdx.x] = hi;
    __syncthreads();
    for (int o = kThreads / 2; o > 0; o >>= 1) {
      if ((int)threadIdx.x < o) {
        slo[threadIdx.x] = host_min(slo[threadIdx.x + o], slo[threadIdx.x]);
        shi[threadIdx.x] = host_max(shi[threadIdx.x + o], shi[threadIdx.x]);
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      part[((long long)blockIdx.x * dim + d) * 2] = slo[0];
      part[((long long)blockIdx.x * dim + d) * 2 + 1] = shi[0];
    }
    __syncthreads();
  }
}

struct Stream {
  hipStream_t s = nullptr;
  Stream() { hip_check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreateWithFlags"); }
  ~Stream() {
    if (s) (void)hipStreamDestroy(s);
  }
};

void upload(DevBuf &b, const double *host, size_t count, hipStream_t s, const char *what) {
  b.reserve(std::max<size_t>(count, 1) * sizeof(double));
  if (count) hip_check(hipMemcpyAsync(b.ptr, host, count * sizeof(double), hipMemcpyHostToDevice, s), what);
}

void check_groups(int n_sum, const double *sums, int n_min, const double *mins, int n_max, const double *maxs) {
  if (n_sum < 0 || n_min < 0 || n_max < 0 || n_sum > 8 || n_min > 8 || n_max > 8)
    fail(BK_ERR_INVALID_ARG, "between 0 and 8 values per attribute group");
  if ((n_sum && !sums) || (n_min && !mins) || (n_max && !maxs)) fail(BK_ERR_INVALID_ARG, "null attribute array");
}

void require_device(int device) {
  int count = 0;
  hip_check(hipGetDeviceCount(&count), "hipGetDeviceCount");
  if (device < 0 || device >= count) fail(BK_ERR_INVALID_ARG, "no HIP device " + std::to_string(device));
}

}  // namespace

extern "C" {

int bk_lexico_runs(int device, int dim, size_t n, const double *coords, long long *order_out, size_t *n_runs_out,
                   long long *run_start_out, double *box_out) {
  return guarded([&] {
    if (dim < 1 || dim > 16) fail(BK_ERR_INVALID_ARG, "dimension must be between 1 and 16");
    if (!n_runs_out) fail(BK_ERR_INVALID_ARG, "null output pointer");
    *n_runs_out = 0;
    if (n == 0) return;
    if (!coords || !order_out || !run_start_out || !box_out) fail(BK_ERR_INVALID_ARG, "null coordinate or output array");
    if (n > 0x7fffffffULL) fail(BK_ERR_INVALID_ARG, "at most 2^31 - 1 points");
    require_device(device);
    DeviceGuard guard(device);
    Stream st;
    const hipStream_t s = st.s;
    struct Bufs {
      DevBuf c, ka, kb, va, vb, tmp, flag, scan, start, part;
      ~Bufs() {
        for (DevBuf *b : {&c, &ka, &kb, &va, &vb, &tmp, &flag, &scan, &start, &part}) b->release();
      }
    } w;
    const long long nn = (long long)n;
    upload(w.c, coords, n * (size_t)dim, s, "upload coordinates");
    bk::lexico_sort(w.c.as<double>(), dim, nn, w.ka, w.kb, w.va, w.vb, w.tmp, s);
    const long long *order = w.va.as<long long>();
    hip_check(hipMemcpyAsync(order_out, order, n * sizeof(long long), hipMemcpyDeviceToHost, s), "read order");
    const unsigned grid = (unsigned)((n + kThreads - 1) / kThreads);
    w.flag.reserve(n * sizeof(int));
    w.scan.reserve(n * sizeof(int));
    run_flags_kernel<<<grid, kThreads, 0, s>>>(w.c.as<double>(), dim, order, nn, w.flag.as<int>());
    hip_check(hipGetLastError(), "run flag launch");
    size_t tmp_bytes = 0;
    hip_check(hipcub::DeviceScan::InclusiveSum(nullptr, tmp_bytes, w.flag.as<int>(), w.scan.as<int>(), (int)n, s),
              "scan size query");
    w.tmp.reserve(tmp_bytes + 256);
    hip_check(hipcub::DeviceScan::InclusiveSum(w.tmp.ptr, tmp_bytes, w.flag.as<int>(), w.scan.as<int>(), (int)n, s),
              "scan");
    w.start.reserve(n * sizeof(long long));
    run_starts_kernel<<<grid, kThreads, 0, s>>>(w.flag.as<int>(), w.scan.as<int>(), nn, w.start.as<long long>());
    hip_check(hipGetLastError(), "run start launch");
    const unsigned box_blocks = (unsigned)std::min<long long>(1024, (nn + kThreads - 1) / kThreads);
    w.part.reserve((size_t)box_blocks * dim * 2 * sizeof(double));
    box_partial_kernel<<<box_blocks, kThreads, 0, s>>>(w.c.as<double>(), dim, nn, w.part.as<double>());
    hip_check(hipGetLastError(), "box launch");
    int n_runs = 0;
    hip_check(hipMemcpyAsync(&n_runs, w.scan.as<int>() + (n - 1), sizeof(int), hipMemcpyDeviceToHost, s),
              "read run count");
    std::vector<double> part((size_t)box_blocks * dim * 2);
    hip_check(hipMemcpyAsync(part.data(), w.part.ptr, part.size() * sizeof(double), hipMemcpyDeviceToHost, s),
              "read box");
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    if (n_runs < 1 || (size_t)n_runs > n) fail(BK_ERR_HIP, "inconsistent run count");
    hip_check(hipMemcpyAsync(run_start_out, w.start.ptr, (size_t)n_runs * sizeof(long long), hipMemcpyDeviceToHost,
                             s),
              "read run starts");
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    for (int d = 0; d < dim; ++d) {
      double lo = DBL_MAX, hi = DBL_MIN;
      for (unsigned b = 0; b < box_blocks; ++b) {
        const double l = part[((size_t)b * dim + d) * 2], h = part[((size_t)b * dim + d) * 2 + 1];
        lo = (l < lo) ? l : lo;
        hi = (hi < h) ? h : hi;
      }
      box_out[2 * d] = lo;
      box_out[2 * d + 1] = hi;
    }
    *n_runs_out = (size_t)n_runs;
  });
}

int bk_run_merge(int device, size_t n_runs, const long long *run_offsets, int n_sum, const double *sums, int n_min,
                 const double *mins, int n_max, const double *maxs, double *sums_out, double *mins_out,
                 double *maxs_out) {
  return guarded([&] {
    if (n_runs == 0) return;
    if (!run_offsets) fail(BK_ERR_INVALID_ARG, "null run offsets");
    check_groups(n_sum, sums, n_min, mins, n_max, maxs);
    if ((n_sum && !sums_out) || (n_min && !mins_out) || (n_max && !maxs_out))
      fail(BK_ERR_INVALID_ARG, "null merged attribute array");
    if (run_offsets[0] != 0) fail(BK_ERR_INVALID_ARG, "run offsets must start at 0");
    for (size_t u = 0; u < n_runs; ++u)
      if (run_offsets[u + 1] <= run_offsets[u]) fail(BK_ERR_INVALID_ARG, "every run needs at least one member");
    const size_t rows = (size_t)run_offsets[n_runs];
    require_device(device);
    DeviceGuard guard(device);
    Stream st;
    const hipStream_t s = st.s;
    struct Bufs {
      DevBuf off, sm, mn, mx, so, mno, mxo;
      ~Bufs() {
        for (DevBuf *b : {&off, &sm, &mn, &mx, &so, &mno, &mxo}) b->release();
      }
    } w;
    w.off.reserve((n_runs + 1) * sizeof(long long));
    hip_check(hipMemcpyAsync(w.off.ptr, run_offsets, (n_runs + 1) * sizeof(long long), hipMemcpyHostToDevice, s),
              "upload run offsets");
    upload(w.sm, sums, rows * (size_t)n_sum, s, "upload summed attributes");
    upload(w.mn, mins, rows * (size_t)n_min, s, "upload minimised attributes");
    upload(w.mx, maxs, rows * (size_t)n_max, s, "upload maximised attributes");
    w.so.reserve(std::max<size_t>(n_runs * n_sum, 1) * sizeof(double));
    w.mno.reserve(std::max<size_t>(n_runs * n_min, 1) * sizeof(double));
    w.mxo.reserve(std::max<size_t>(n_runs * n_max, 1) * sizeof(double));
    run_merge_kernel<<<(unsigned)((n_runs + kThreads - 1) / kThreads), kThreads, 0, s>>>(
        w.off.as<long long>(), (long long)n_runs, n_sum, w.sm.as<double>(), n_min, w.mn.as<double>(), n_max,
        w.mx.as<double>(), w.so.as<double>(), w.mno.as<double>(), w.mxo.as<double>());
    hip_check(hipGetLastError(), "run merge launch");
    if (n_sum)
      hip_check(hipMemcpyAsync(sums_out, w.so.ptr, n_runs * n_sum * sizeof(double), hipMemcpyDeviceToHost, s),
                "read merged sums");
    if (n_min)
      hip_check(hipMemcpyAsync(mins_out, w.mno.ptr, n_runs * n_min * sizeof(double), hipMemcpyDeviceToHost, s),
                "read merged minima");
    if (n_max)
      hip_check(hipMemcpyAsync(maxs_out, w.mxo.ptr, n_runs * n_max * sizeof(double), hipMemcpyDeviceToHost, s),
                "read merged maxima");
    hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
  });
}

}  // extern "C"
