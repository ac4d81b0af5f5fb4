// Runtime plumbing: last-error storage, per-thread streams, zero fill.
#include "common.h"

#include <mutex>
#include <vector>

namespace whir {

static thread_local std::string g_last_error;

void set_last_error(const std::string& msg) { g_last_error = msg; }
const char* last_error() { return g_last_error.c_str(); }

namespace {
struct ThreadStream {
  hipStream_t s = nullptr;
  int device = -1;
  ~ThreadStream() {
    if (s) (void)hipStreamDestroy(s);
  }
};
thread_local ThreadStream t_stream;
}  // namespace

hipStream_t thread_stream() {
  int dev = 0;
  WHIR_HIP(hipGetDevice(&dev));
  if (!t_stream.s || t_stream.device != dev) {
    if (t_stream.s) (void)hipStreamDestroy(t_stream.s);
    WHIR_HIP(hipStreamCreateWithFlags(&t_stream.s, hipStreamNonBlocking));
    t_stream.device = dev;
  }
  return t_stream.s;
}

__global__ void k_fill_zero(uint4* __restrict__ p, size_t n16, uint8_t* __restrict__ tail, size_t ntail) {
  size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t k = i; k < n16; k += stride) p[k] = make_uint4(0, 0, 0, 0);
  if (i < ntail) tail[i] = 0;
}

void fill_zero(void* dst, size_t bytes, hipStream_t s) {
  if (!bytes) return;
  WHIR_REQUIRE(((uintptr_t)dst & 15) == 0, "fill_zero: destination must be 16-byte aligned");
  size_t n16 = bytes / 16, ntail = bytes % 16;
  unsigned threads = 256;
  unsigned blocks = (unsigned)std::min<size_t>(blocks_for(n16 ? n16 : 1, threads), 8192);
  k_fill_zero<<<blocks, threads, 0, s>>>((uint4*)dst, n16, (uint8_t*)dst + n16 * 16, ntail);
  WHIR_HIP(hipGetLastError());
}

}  // namespace whir
