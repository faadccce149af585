// tt_hip.h -- HIP helpers shared by the device translation units: error
// checking that turns HIP failures into C-ABI error codes, an owning device
// buffer, and launch-size helpers.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "tt_internal.h"

namespace tt {

struct HipError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

inline void hip_check(hipError_t e, const char *what, const char *file, int line) {
  if (e != hipSuccess) {
    char buf[512];
    std::snprintf(buf, sizeof(buf), "%s failed at %s:%d: %s", what, file, line, hipGetErrorString(e));
    throw HipError(buf);
  }
}
#define TT_HIP(x) ::tt::hip_check((x), #x, __FILE__, __LINE__)
#define TT_LAUNCH_CHECK() ::tt::hip_check(hipGetLastError(), "kernel launch", __FILE__, __LINE__)

// Owning device allocation.
template <typename T>
struct DBuf {
  T *p = nullptr;
  size_t n = 0;
  DBuf() = default;
  explicit DBuf(size_t count) { alloc(count); }
  DBuf(const DBuf &) = delete;
  DBuf &operator=(const DBuf &) = delete;
  DBuf(DBuf &&o) noexcept : p(o.p), n(o.n) { o.p = nullptr, o.n = 0; }
  DBuf &operator=(DBuf &&o) noexcept {
    if (this != &o) {
      release();
      p = o.p, n = o.n;
      o.p = nullptr, o.n = 0;
    }
    return *this;
  }
  ~DBuf() { release(); }
  void alloc(size_t count) {
    release();
    n = count;
    if (count) TT_HIP(hipMalloc(&p, sizeof(T) * count));
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr, n = 0;
  }
  void upload(const T *h, size_t count, hipStream_t s = 0) {
    if (count > n) alloc(count);
    if (count) TT_HIP(hipMemcpyAsync(p, h, sizeof(T) * count, hipMemcpyHostToDevice, s));
  }
  void upload(const std::vector<T> &v, hipStream_t s = 0) { upload(v.data(), v.size(), s); }
  void download(T *h, size_t count, hipStream_t s = 0) const {
    if (count) TT_HIP(hipMemcpyAsync(h, p, sizeof(T) * count, hipMemcpyDeviceToHost, s));
  }
  void zero(hipStream_t s = 0) {
    if (n) TT_HIP(hipMemsetAsync(p, 0, sizeof(T) * n, s));
  }
  size_t bytes() const { return sizeof(T) * n; }
};

inline unsigned ceil_div(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

// Runs f and converts exceptions into a negative C-ABI status.
template <typename F>
int guarded(const char *fn, F &&f) {
  try {
    return f();
  } catch (std::exception &e) {
    tt_set_error("%s: %s", fn, e.what());
    return -1;
  }
}

}  // namespace tt
