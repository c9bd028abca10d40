// Spatial ordering of query lists: a device radix sort of Morton keys.
//
// A workgroup of the FP32 Gaussian kernel skips every reference tile whose
// bounding box lies beyond the kernel's reach from the bounding box of the
// workgroup's queries.  That only pays when a workgroup's queries are close
// together, so callers that evaluate scattered query subsets (the
// surrogate's sampled check and its fallback) sort them along a Z curve
// first and scatter the results back through the same list.
#include "bk_common.h"
#include "bk_internal.h"

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <string>
#include <utility>

namespace bk {

namespace {

constexpr int kMortonBits = 10;  // cells per axis: 2^kMortonBits

__device__ __forceinline__ unsigned int spread_bits(unsigned int v) {
  v &= 0x0000FFFFu;
  v = (v | (v << 8)) & 0x00FF00FFu;
  v = (v | (v << 4)) & 0x0F0F0F0Fu;
  v = (v | (v << 2)) & 0x33333333u;
  v = (v | (v << 1)) & 0x55555555u;
  return v;
}

// Keys of FP32 or FP64 coordinates (the cell of an FP64 point is found from
// its FP32 rounding: the order only has to group nearby points).
template <typename T>
__global__ void morton_keys_kernel(const T *__restrict__ qx, const T *__restrict__ qy,
                                   const long long *__restrict__ list, long long n, float lo, float scale,
                                   int index_bits, unsigned long long *__restrict__ keys) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const long long i = list ? list[k] : k;
  constexpr float top = (float)((1 << kMortonBits) - 1);
  // fminf/fmaxf return the number when the other operand is NaN
  const float fx = __builtin_fminf(__builtin_fmaxf(((float)qx[i] - lo) * scale, 0.0f), top);
  const float fy = __builtin_fminf(__builtin_fmaxf(((float)qy[i] - lo) * scale, 0.0f), top);
  const unsigned int code = spread_bits((unsigned int)fx) | (spread_bits((unsigned int)fy) << 1);
  keys[k] = ((unsigned long long)code << index_bits) | (unsigned long long)i;
}

__global__ void keys_to_list_kernel(const unsigned long long *__restrict__ keys, long long n,
                                    unsigned long long mask, long long *__restrict__ list) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) list[k] = (long long)(keys[k] & mask);
}

template <typename T>
__global__ void gather_xy_kernel(const T *__restrict__ qx, const T *__restrict__ qy,
                                 const long long *__restrict__ list, long long n, T *__restrict__ gx,
                                 T *__restrict__ gy) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const long long i = list[k];
  gx[k] = qx[i];
  gy[k] = qy[i];
}

__global__ void scatter_f64_kernel(const long long *__restrict__ list, long long n, const double *__restrict__ vals,
                                   double *__restrict__ out) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) out[list[k]] = vals[k];
}

// Order-preserving 64-bit key of a double: -0 and +0 share the key of +0
// (ExactLexicoLess compares them equal); negative values take their
// complemented bits, the others their bits with the sign bit set, so
// unsigned order is numeric order.
__device__ __forceinline__ unsigned long long lexico_key(double v) {
  if (v == 0.0) v = 0.0;
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  return (b >> 63) ? ~b : (b | 0x8000000000000000ULL);
}

// keys[k] = key of coordinate d of point order[k]
__global__ void lexico_keys_kernel(const double *__restrict__ coords, int dim, int d,
                                   const long long *__restrict__ order, long long n,
                                   unsigned long long *__restrict__ keys) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  keys[k] = lexico_key(coords[order[k] * dim + d]);
}

// keys[i] = ~(tiles listed for pair i = b + sp * nblocks within split sp's
// tile range), vals[i] = i
__global__ void listed_count_kernel(const unsigned long long *__restrict__ mask, long long nwords, int nblocks,
                                    int splits, long long split_tiles, unsigned int *__restrict__ keys,
                                    int *__restrict__ vals) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)nblocks * splits) return;
  const long long b = i % nblocks, sp = i / nblocks;
  const unsigned long long *m = mask + b * nwords;
  const long long t0 = sp * split_tiles;
  long long t1 = t0 + split_tiles;
  if (t1 > nwords * 64) t1 = nwords * 64;
  unsigned int c = 0;
  for (long long t = t0; t < t1;) {
    const long long w = t >> 6;
    const long long stop = (w + 1) * 64 < t1 ? (w + 1) * 64 : t1;
    const int len = (int)(stop - t);
    unsigned long long bits = m[w] >> (t & 63);
    if (len < 64) bits &= (1ULL << len) - 1ULL;
    c += (unsigned int)__builtin_popcountll(bits);
    t = stop;
  }
  keys[i] = ~c;
  vals[i] = (int)i;
}

__global__ void iota_kernel(long long *__restrict__ v, long long n) {
  const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) v[k] = k;
}

}  // namespace

hipError_t launch_gather_xy(const float *qx, const float *qy, const long long *list, long long n, float *gx,
                            float *gy, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  gather_xy_kernel<float><<<(unsigned)((n + 255) / 256), 256, 0, s>>>(qx, qy, list, n, gx, gy);
  return hipGetLastError();
}

hipError_t launch_gather_xy(const double *qx, const double *qy, const long long *list, long long n, double *gx,
                            double *gy, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  gather_xy_kernel<double><<<(unsigned)((n + 255) / 256), 256, 0, s>>>(qx, qy, list, n, gx, gy);
  return hipGetLastError();
}

hipError_t launch_scatter_f64(const long long *list, long long n, const double *vals, double *out, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  scatter_f64_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(list, n, vals, out);
  return hipGetLastError();
}

void order_by_keys(unsigned int *keys, int *vals, long long n, int *order_out, DevBuf &keys_b, DevBuf &tmp,
                   hipStream_t s) {
  if (n <= 0) return;
  if (n > 0x7fffffffLL) fail(BK_ERR_INVALID_ARG, "order_by_keys: too many items");
  keys_b.reserve((size_t)n * sizeof(unsigned int));
  size_t tmp_bytes = 0;
  hip_check(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, keys, keys_b.as<unsigned int>(), vals, order_out,
                                               (int)n, 0, 32, s),
            "radix sort size query");
  tmp.reserve(tmp_bytes + 256);
  hip_check(hipcub::DeviceRadixSort::SortPairs(tmp.ptr, tmp_bytes, keys, keys_b.as<unsigned int>(), vals, order_out,
                                               (int)n, 0, 32, s),
            "radix sort");
}

void listed_order(const unsigned long long *mask, long long nwords, int nblocks, int splits, long long split_tiles,
                  int *order_out, DevBuf &keys_a, DevBuf &keys_b, DevBuf &vals, DevBuf &tmp, hipStream_t s) {
  const long long n = (long long)nblocks * splits;
  if (n <= 0) return;
  if (n > 0x7fffffffLL) fail(BK_ERR_INVALID_ARG, "listed_order: too many (block, split) pairs");
  keys_a.reserve((size_t)n * sizeof(unsigned int));
  keys_b.reserve((size_t)n * sizeof(unsigned int));
  vals.reserve((size_t)n * sizeof(int));
  size_t tmp_bytes = 0;
  hip_check(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, keys_a.as<unsigned int>(), keys_b.as<unsigned int>(),
                                               vals.as<int>(), order_out, (int)n, 0, 32, s),
            "radix sort size query");
  tmp.reserve(tmp_bytes + 256);
  listed_count_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(mask, nwords, nblocks, splits, split_tiles,
                                                                   keys_a.as<unsigned int>(), vals.as<int>());
  hip_check(hipGetLastError(), "listed count launch");
  hip_check(hipcub::DeviceRadixSort::SortPairs(tmp.ptr, tmp_bytes, keys_a.as<unsigned int>(),
                                               keys_b.as<unsigned int>(), vals.as<int>(), order_out, (int)n, 0, 32, s),
            "radix sort");
}

namespace {
template <typename T>
void spatial_sort_impl(const T *qx, const T *qy, const long long *list, long long n, long long index_limit,
                       float lo, float hi, long long *list_out, DevBuf &keys_a, DevBuf &keys_b, DevBuf &tmp,
                       hipStream_t s) {
  if (n <= 0) return;
  if (!list) index_limit = n;
  if (index_limit < n || index_limit > (1LL << 32))
    fail(BK_ERR_INVALID_ARG, "spatial_sort: indices must lie in [0, 2^32) and cover the list");
  if (!(hi > lo)) fail(BK_ERR_INVALID_ARG, "spatial_sort: empty box");
  // the key holds the index in its low bits: enough bits for every index
  int index_bits = 1;
  while ((1LL << index_bits) < index_limit) ++index_bits;
  const int end_bit = index_bits + 2 * kMortonBits;
  keys_a.reserve((size_t)n * sizeof(unsigned long long));
  keys_b.reserve((size_t)n * sizeof(unsigned long long));
  size_t tmp_bytes = 0;
  hip_check(hipcub::DeviceRadixSort::SortKeys(nullptr, tmp_bytes, keys_a.as<unsigned long long>(),
                                              keys_b.as<unsigned long long>(), (int)n, 0, end_bit, s),
            "radix sort size query");
  tmp.reserve(tmp_bytes + 256);
  const unsigned grid = (unsigned)((n + 255) / 256);
  const float scale = (float)((double)(1 << kMortonBits) / ((double)hi - (double)lo));
  morton_keys_kernel<T>
      <<<grid, 256, 0, s>>>(qx, qy, list, n, lo, scale, index_bits, keys_a.as<unsigned long long>());
  hip_check(hipGetLastError(), "morton key launch");
  hip_check(hipcub::DeviceRadixSort::SortKeys(tmp.ptr, tmp_bytes, keys_a.as<unsigned long long>(),
                                              keys_b.as<unsigned long long>(), (int)n, 0, end_bit, s),
            "radix sort");
  keys_to_list_kernel<<<grid, 256, 0, s>>>(keys_b.as<unsigned long long>(), n, (1ULL << index_bits) - 1ULL,
                                           list_out);
  hip_check(hipGetLastError(), "key unpack launch");
}
}  // namespace

void spatial_sort(const float *qx, const float *qy, const long long *list, long long n, long long index_limit,
                  float lo, float hi, long long *list_out, DevBuf &keys_a, DevBuf &keys_b, DevBuf &tmp,
                  hipStream_t s) {
  spatial_sort_impl<float>(qx, qy, list, n, index_limit, lo, hi, list_out, keys_a, keys_b, tmp, s);
}

void spatial_sort(const double *qx, const double *qy, const long long *list, long long n, long long index_limit,
                  float lo, float hi, long long *list_out, DevBuf &keys_a, DevBuf &keys_b, DevBuf &tmp,
                  hipStream_t s) {
  spatial_sort_impl<double>(qx, qy, list, n, index_limit, lo, hi, list_out, keys_a, keys_b, tmp, s);
}

}  // namespace bk

// Lexicographic order of n points by a stable device radix sort, least
// significant coordinate first (see include/bbrcit_hip.h).
namespace bk {
void lexico_sort(const double *coords, int dim, long long n, DevBuf &ka, DevBuf &kb, DevBuf &va, DevBuf &vb,
                 DevBuf &tmp, hipStream_t s) {
  ka.reserve((size_t)n * sizeof(unsigned long long));
  kb.reserve((size_t)n * sizeof(unsigned long long));
  va.reserve((size_t)n * sizeof(long long));
  vb.reserve((size_t)n * sizeof(long long));
  size_t tmp_bytes = 0;
  hip_check(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, ka.as<unsigned long long>(),
                                               kb.as<unsigned long long>(), va.as<long long>(), vb.as<long long>(),
                                               (int)n, 0, 64, s),
            "radix sort size query");
  tmp.reserve(tmp_bytes + 256);
  const unsigned grid = (unsigned)((n + 255) / 256);
  iota_kernel<<<grid, 256, 0, s>>>(va.as<long long>(), n);
  hip_check(hipGetLastError(), "iota launch");
  for (int d = dim - 1; d >= 0; --d) {
    lexico_keys_kernel<<<grid, 256, 0, s>>>(coords, dim, d, va.as<long long>(), n, ka.as<unsigned long long>());
    hip_check(hipGetLastError(), "key launch");
    hip_check(hipcub::DeviceRadixSort::SortPairs(tmp.ptr, tmp_bytes, ka.as<unsigned long long>(),
                                                 kb.as<unsigned long long>(), va.as<long long>(),
                                                 vb.as<long long>(), (int)n, 0, 64, s),
              "radix sort");
    std::swap(va, vb);
  }
}
}  // namespace bk

extern "C" int bk_lexico_order(int device, int dim, size_t n, const double *coords, long long *order_out) {
  using namespace bk;
  return guarded([&] {
    if (dim < 1 || dim > 16) fail(BK_ERR_INVALID_ARG, "dimension must be between 1 and 16");
    if (n == 0) return;
    if (!coords || !order_out) fail(BK_ERR_INVALID_ARG, "null coordinate or order array");
    if (n > 0x7fffffffULL) fail(BK_ERR_INVALID_ARG, "at most 2^31 - 1 points");
    int count = 0;
    hip_check(hipGetDeviceCount(&count), "hipGetDeviceCount");
    if (device < 0 || device >= count) fail(BK_ERR_INVALID_ARG, "no HIP device " + std::to_string(device));
    DeviceGuard guard(device);
    struct Work {
      hipStream_t s = nullptr;
      DevBuf c, ka, kb, va, vb, tmp;
      ~Work() {
        for (DevBuf *b : {&c, &ka, &kb, &va, &vb, &tmp}) b->release();
        if (s) (void)hipStreamDestroy(s);
      }
    } w;
    hip_check(hipStreamCreate(&w.s), "hipStreamCreate");
    w.c.reserve(n * (size_t)dim * sizeof(double));
    hip_check(hipMemcpyAsync(w.c.ptr, coords, n * (size_t)dim * sizeof(double), hipMemcpyHostToDevice, w.s),
              "upload coordinates");
    lexico_sort(w.c.as<double>(), dim, (long long)n, w.ka, w.kb, w.va, w.vb, w.tmp, w.s);
    hip_check(hipMemcpyAsync(order_out, w.va.ptr, n * sizeof(long long), hipMemcpyDeviceToHost, w.s),
              "download order");
    hip_check(hipStreamSynchronize(w.s), "hipStreamSynchronize");
  });
}
