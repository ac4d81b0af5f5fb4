// Shared runtime pieces of libwhir_amd: error reporting across the C ABI,
// per-thread HIP streams, stream-ordered scratch allocations and launch helpers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

#include <stdexcept>
#include <string>

namespace whir {

// Errors inside the library are C++ exceptions; every extern "C" entry point
// catches them and turns them into a nonzero status plus whir_last_error().
struct Error : std::runtime_error {
  explicit Error(const std::string& m) : std::runtime_error(m) {}
};

void set_last_error(const std::string& msg);
const char* last_error();

#define WHIR_HIP(expr)                                                              \
  do {                                                                              \
    hipError_t _e = (expr);                                                         \
    if (_e != hipSuccess)                                                           \
      throw ::whir::Error(std::string(#expr) + ": " + hipGetErrorString(_e));      \
  } while (0)

#define WHIR_REQUIRE(cond, msg)                                                     \
  do {                                                                              \
    if (!(cond)) throw ::whir::Error(msg);                                          \
  } while (0)

// The stream the calling host thread launches on (created at default priority
// on first use; one per thread so concurrent callers overlap on the device).
hipStream_t thread_stream();

// Stream-ordered device scratch that frees itself on scope exit.
struct DevBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
  hipStream_t stream = nullptr;
  DevBuf() = default;
  DevBuf(size_t n, hipStream_t s) : bytes(n), stream(s) {
    if (n) WHIR_HIP(hipMallocAsync(&ptr, n, s));
  }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : ptr(o.ptr), bytes(o.bytes), stream(o.stream) { o.ptr = nullptr; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    release();
    ptr = o.ptr; bytes = o.bytes; stream = o.stream; o.ptr = nullptr;
    return *this;
  }
  ~DevBuf() { release(); }
  void release() {
    if (ptr) (void)hipFreeAsync(ptr, stream);
    ptr = nullptr;
  }
  template <class T> T* as() const { return static_cast<T*>(ptr); }
};

inline unsigned blocks_for(size_t n, unsigned threads) {
  size_t b = (n + threads - 1) / threads;
  return (unsigned)(b ? b : 1);
}

// Zero a device range with a kernel (no memset nodes: the launch may be captured).
void fill_zero(void* dst, size_t bytes, hipStream_t s);

}  // namespace whir
