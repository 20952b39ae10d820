// Shared definitions for librespack_gpu (HIP, gfx950).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace rp {

struct Error : std::runtime_error {
  using std::runtime_error::runtime_error;
};

#define RP_HIP(x)                                                                      \
  do {                                                                                 \
    hipError_t e_ = (x);                                                               \
    if (e_ != hipSuccess)                                                              \
      throw ::rp::Error(std::string(#x) + " failed: " + hipGetErrorString(e_) + " (" + \
                        __FILE__ + ":" + std::to_string(__LINE__) + ")");              \
  } while (0)

#define RP_CHECK(cond, msg)                  \
  do {                                       \
    if (!(cond)) throw ::rp::Error(msg);     \
  } while (0)

// Owning device buffer (hipMalloc'd; freed on destruction or re-allocation).
template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n) {
    o.p = nullptr;
    o.n = 0;
  }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      release();
      p = o.p;
      n = o.n;
      o.p = nullptr;
      o.n = 0;
    }
    return *this;
  }
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  // Grows (never shrinks) the allocation to hold at least m elements.
  void reserve(size_t m) {
    if (m <= n && p) return;
    release();
    if (m == 0) return;
    RP_HIP(hipMalloc(&p, m * sizeof(T)));
    n = m;
  }
  void upload(const T* h, size_t m, hipStream_t s) {
    reserve(m);
    if (m) RP_HIP(hipMemcpyAsync(p, h, m * sizeof(T), hipMemcpyHostToDevice, s));
  }
  void upload(const std::vector<T>& v, hipStream_t s) { upload(v.data(), v.size(), s); }
  void swap(DevBuf& o) {
    std::swap(p, o.p);
    std::swap(n, o.n);
  }
};

// ------------------------------------------------------------ device complex
struct zd {
  double x, y;
};
__host__ __device__ inline zd mkz(double a, double b) { return zd{a, b}; }
__host__ __device__ inline zd operator+(zd a, zd b) { return {a.x + b.x, a.y + b.y}; }
__host__ __device__ inline zd operator-(zd a, zd b) { return {a.x - b.x, a.y - b.y}; }
__host__ __device__ inline zd operator-(zd a) { return {-a.x, -a.y}; }
__host__ __device__ inline zd operator*(zd a, zd b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
__host__ __device__ inline zd operator*(zd a, double s) { return {a.x * s, a.y * s}; }
__host__ __device__ inline zd operator*(double s, zd a) { return {a.x * s, a.y * s}; }
__host__ __device__ inline zd operator/(zd a, double s) { return {a.x / s, a.y / s}; }
__host__ __device__ inline zd conjz(zd a) { return {a.x, -a.y}; }
// Complex division with Smith's scaling (robust for the wide dynamic range of
// the tetrahedron closed forms).
__host__ __device__ inline zd operator/(zd a, zd b) {
  if (fabs(b.x) >= fabs(b.y)) {
    double r = b.y / b.x, den = b.x + b.y * r;
    return {(a.x + a.y * r) / den, (a.y - a.x * r) / den};
  } else {
    double r = b.x / b.y, den = b.x * r + b.y;
    return {(a.x * r + a.y) / den, (a.y * r - a.x) / den};
  }
}
__host__ __device__ inline zd operator/(double s, zd b) { return zd{s, 0.0} / b; }
__host__ __device__ inline double absz(zd a) { return hypot(a.x, a.y); }
__device__ __forceinline__ zd logz(zd a) { return {log(hypot(a.x, a.y)), atan2(a.y, a.x)}; }

}  // namespace rp
