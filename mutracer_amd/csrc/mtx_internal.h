// Host-side helpers shared by the library's translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/mtx.h"
#include "mtx_accel.h"

namespace mtx {

// Error carried to the C ABI boundary: a status code and a message.
struct Error : std::runtime_error {
  int32_t code;
  Error(int32_t c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define MTX_HIP(expr)                                                                   \
  do {                                                                                  \
    hipError_t _e = (expr);                                                             \
    if (_e != hipSuccess) {                                                             \
      (void)hipGetLastError(); /* do not leave the failure as the thread's last error */ \
      throw ::mtx::Error(MTX_CONTEXT_FAILED, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    }                                                                                   \
  } while (0)

#define MTX_REQUIRE(cond, code, msg)                  \
  do {                                                \
    if (!(cond)) throw ::mtx::Error((code), (msg));   \
  } while (0)

// Owning device allocation.
template <typename T>
struct DevBuf {
  T *ptr = nullptr;
  size_t count = 0;
  DevBuf() = default;
  explicit DevBuf(size_t n) { alloc(n); }
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  DevBuf(DevBuf &&o) noexcept : ptr(o.ptr), count(o.count) { o.ptr = nullptr; o.count = 0; }
  DevBuf &operator=(DevBuf &&o) noexcept {
    if (this != &o) { release(); ptr = o.ptr; count = o.count; o.ptr = nullptr; o.count = 0; }
    return *this;
  }
  ~DevBuf() { release(); }
  void alloc(size_t n) {
    release();
    if (n) MTX_HIP(hipMalloc((void **)&ptr, n * sizeof(T)));
    count = n;
  }
  void release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    count = 0;
  }
  size_t bytes() const { return count * sizeof(T); }
  void upload(const T *h, size_t n, hipStream_t s = 0) {
    if (count < n) alloc(n);
    if (n) MTX_HIP(hipMemcpyAsync(ptr, h, n * sizeof(T), hipMemcpyHostToDevice, s));
  }
  // Allocates at least one element and uploads the n given (n may be 0).
  void upload_min1(const T *h, size_t n, hipStream_t s = 0) {
    if (count < (n ? n : 1)) alloc(n ? n : 1);
    upload(h, n, s);
  }
  void download(T *h, size_t n, hipStream_t s = 0) const {
    if (n) MTX_HIP(hipMemcpyAsync(h, ptr, n * sizeof(T), hipMemcpyDeviceToHost, s));
  }
  std::vector<T> to_host(size_t n, hipStream_t s = 0) const {
    std::vector<T> v(n);
    download(v.data(), n, s);
    MTX_HIP(hipStreamSynchronize(s));
    return v;
  }
};

inline void set_error(char *err, size_t cap, const std::string &msg) {
  if (!err || !cap) return;
  size_t n = msg.size() < cap - 1 ? msg.size() : cap - 1;
  // Truncate on a UTF-8 boundary.
  while (n > 0 && n < msg.size() && ((unsigned char)msg[n] & 0xC0) == 0x80) n--;
  std::memcpy(err, msg.data(), n);
  err[n] = 0;
}

}  // namespace mtx
