// Host-side helpers shared by the C-ABI translation units: error codes
// carried as C++ exceptions up to the ABI boundary, a grow-only device
// buffer, and a scoped device switch.
#pragma once

#include "../../include/bbrcit_hip.h"

#include <hip/hip_runtime.h>

#include <new>
#include <stdexcept>
#include <string>

namespace bk {

struct BkError {
  int code;
  std::string msg;
};

[[noreturn]] inline void fail(int code, const std::string &msg) { throw BkError{code, msg}; }

inline void hip_check(hipError_t e, const char *what) {
  if (e != hipSuccess) fail(BK_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

// Message of the last failed call on this thread (returned by bk_last_error).
std::string &last_error();

// Run `f`, turning exceptions into a status code and the thread's message.
template <class F>
int guarded(F &&f) {
  try {
    f();
    last_error().clear();
    return BK_OK;
  } catch (const BkError &e) {
    last_error() = e.msg;
    return e.code;
  } catch (const std::bad_alloc &) {
    last_error() = "host allocation failed";
    return BK_ERR_ALLOC;
  } catch (const std::exception &e) {
    last_error() = e.what();
    return BK_ERR_STATE;
  } catch (...) {
    last_error() = "unknown failure";
    return BK_ERR_STATE;
  }
}

// Device buffer that only grows.
struct DevBuf {
  void *ptr = nullptr;
  size_t bytes = 0;
  void reserve(size_t need) {
    if (need <= bytes) return;
    if (ptr) {
      hip_check(hipFree(ptr), "hipFree");
      ptr = nullptr;
      bytes = 0;
    }
    if (hipMalloc(&ptr, need) != hipSuccess) {
      ptr = nullptr;
      fail(BK_ERR_ALLOC, "hipMalloc of " + std::to_string(need) + " bytes failed");
    }
    bytes = need;
  }
  void release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
  }
  template <class T>
  T *as() const {
    return static_cast<T *>(ptr);
  }
};

// Makes `dev` current for the scope and restores the previous device.
struct DeviceGuard {
  int prev = -1;
  explicit DeviceGuard(int dev) {
    hip_check(hipGetDevice(&prev), "hipGetDevice");
    if (prev != dev) hip_check(hipSetDevice(dev), "hipSetDevice");
  }
  ~DeviceGuard() {
    int cur = -1;
    if (hipGetDevice(&cur) == hipSuccess && cur != prev && prev >= 0) (void)hipSetDevice(prev);
  }
  DeviceGuard(const DeviceGuard &) = delete;
  DeviceGuard &operator=(const DeviceGuard &) = delete;
};

// Device and current stream of a context (defined with the context).
int ctx_device(const bk_ctx *ctx);
hipStream_t ctx_stream(const bk_ctx *ctx);

// Normalised direct sums over all references of an FP32 2-D context for nq
// device-resident FP32 queries, into device memory `out`, on stream s.
void eval_queries_dev(bk_ctx *ctx, const bk_kernel_desc *desc, const float *qx, const float *qy, size_t nq,
                      double *out, hipStream_t s);

}  // namespace bk
