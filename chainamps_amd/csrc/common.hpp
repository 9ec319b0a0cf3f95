// Common utilities: error reporting, HIP checks, device buffers.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace cam {

struct Error : std::runtime_error {
  explicit Error(const std::string& s) : std::runtime_error(s) {}
};

void set_last_error(const std::string& s);
const char* last_error();

#define CAM_THROW(msg) throw ::cam::Error(std::string(msg))
#define CAM_REQUIRE(cond, msg)                                   \
  do {                                                           \
    if (!(cond)) throw ::cam::Error(std::string(msg));           \
  } while (0)
#define HIP_CHECK(expr)                                                              \
  do {                                                                               \
    hipError_t e_ = (expr);                                                          \
    if (e_ != hipSuccess)                                                            \
      throw ::cam::Error(std::string("HIP error ") + hipGetErrorString(e_) + " at " + \
                         __FILE__ + ":" + std::to_string(__LINE__) + " (" #expr ")"); \
  } while (0)

// Owning device allocation (hipMalloc'd), move-only.
template <typename T>
class DevBuf {
 public:
  DevBuf() = default;
  explicit DevBuf(size_t n) { alloc(n); }
  ~DevBuf() { reset(); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; o.n_ = 0; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) { reset(); p_ = o.p_; n_ = o.n_; o.p_ = nullptr; o.n_ = 0; }
    return *this;
  }
  void alloc(size_t n) {
    reset();
    if (n) HIP_CHECK(hipMalloc(reinterpret_cast<void**>(&p_), n * sizeof(T)));
    n_ = n;
  }
  // grow-only (contents not preserved)
  void ensure(size_t n) { if (n > n_) alloc(n); }
  void reset() {
    if (p_) (void)hipFree(p_);
    p_ = nullptr;
    n_ = 0;
  }
  T* get() const { return p_; }
  size_t size() const { return n_; }
  void upload(const T* h, size_t n, hipStream_t s) {
    CAM_REQUIRE(n <= n_, "upload overflow");
    if (n) HIP_CHECK(hipMemcpyAsync(p_, h, n * sizeof(T), hipMemcpyHostToDevice, s));
  }
  void upload(const std::vector<T>& v, hipStream_t s) { ensure(v.size()); upload(v.data(), v.size(), s); }

 private:
  T* p_ = nullptr;
  size_t n_ = 0;
};

// Non-owning view into a workspace DevBuf (same get() interface).
template <typename T>
struct DevView {
  T* p = nullptr;
  T* get() const { return p; }
};

struct cplx {
  double re, im;
};

}  // namespace cam
