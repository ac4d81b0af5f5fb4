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
// Frees this thread's stream (with its buffers and mailbox) and its pinned staging
// and readback areas now; a worker thread calls it before it ends, so that no HIP
// call runs in thread-exit destructors (where profiling tools intercepting HIP have
// already torn down their per-thread state).  The thread may use HIP again after.
void release_thread_resources();

// The batched-commit plans of the calling thread (prover.hip): released by
// release_thread_resources before the thread's stream.
void release_commit_plans();

// Enqueue a host->device copy on stream s; `src` may be reused as soon as this
// returns (small copies go through a pinned per-thread staging area).
void upload_async(void* dst, const void* src, size_t bytes, hipStream_t s);
// Copy device->host through a pinned per-thread buffer and wait for it.
void download_sync(void* dst, const void* src, size_t bytes, hipStream_t s);

// Pinned host words (RESULT_MAILBOX_WORDS per stream) that kernels write small
// results into directly, saving a device-to-host copy per result; read them after
// mailbox_wait (or a stream synchronization).  The words are fine-grained (coherent)
// host memory, so device stores to them are not held in the GPU's L2.
constexpr size_t RESULT_MAILBOX_WORDS = 65536;
uint64_t* result_mailbox(hipStream_t s);
// Waits until the results the host is about to read from the mailbox have landed: a
// new sequence number is stored to a coherent flag word after the mailbox with a
// system-scope release, and the host spins on the flag (about 6 us from kernel end to
// host, against about 12 us for hipStreamSynchronize on MI355X).  The store comes from
// the kernel that wrote the awaited result when the wait was armed (mailbox_arm) and
// that kernel took the request (mailbox_take): everything queued before that launch
// has then finished, and work queued after it (none of which the host reads) may still
// run.  Otherwise a one-thread kernel queued behind the stream's work stores it, and
// everything queued so far has finished.  A stream error ends the wait with that
// error.  Not for use while s is being captured.
void mailbox_wait(hipStream_t s);

// A request that a kernel storing a result into the mailbox also raise the flag.
struct MailboxSig {
  uint64_t* flag = nullptr;
  uint64_t seq = 0;
};
// Arms the calling thread's next mailbox_wait on s: the launch that stores the awaited
// result may take the request and raise the flag itself, which saves the separate
// signal kernel.  Call it right before the call whose LAST launch stores that result,
// so that nothing the host reads is still being written when the flag rises.
void mailbox_arm(hipStream_t s);
// The armed request when `out` (the destination of the launch about to be made) lies
// in the armed stream's mailbox, which disarms it; otherwise an empty request.  Only a
// launch after which nothing the host will read is still to be written may take it.
MailboxSig mailbox_take(const void* out);
#if defined(__HIP__)
// Device side, by the one thread that stored the result (after its stores).
__device__ inline void mailbox_raise(const MailboxSig& sig) {
  if (sig.flag) {
    __threadfence_system();
    __hip_atomic_store(sig.flag, sig.seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}
#endif

// Device buffers recycled per stream: a buffer released on stream s goes to s's
// free list and is handed out again only to allocations on s, so stream order
// alone keeps its last readers ahead of its next writers.  Sizes repeat from
// one proof to the next, so after the first proof nothing is allocated.
// Allocations come from the library's own stream-ordered pool (one per device); an
// allocation that finds the device out of memory first returns every stream's cached
// buffers and tries once more.
void* pool_alloc(size_t bytes, hipStream_t s);
void pool_free(void* p, size_t bytes, hipStream_t s);
// Drop every cached buffer of stream s (e.g. before the stream is destroyed), or of
// every stream; the freed memory goes back to the device.
void pool_trim(hipStream_t s);
void pool_trim_all();
// Bytes held in the free lists of stream s (all streams when s is null).
size_t pool_cached_bytes(hipStream_t s);

// Stream-ordered device scratch that returns to the pool on scope exit.
struct DevBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
  hipStream_t stream = nullptr;
  DevBuf() = default;
  DevBuf(size_t n, hipStream_t s) : bytes(n), stream(s) {
    if (n) ptr = pool_alloc(n, s);
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
    if (ptr) pool_free(ptr, bytes, stream);
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

// Set while the calling thread captures its stream into a HIP graph: code that would
// allocate, synchronize, tune or time with events must not run then (the profiler's
// scopes record nothing).
bool capture_active();
void set_capture_active(bool on);

// Kernel timing with HIP events on the launching stream, off unless enabled
// (whir_profile_enable).  A scope brackets one launch of a named kernel family
// and credits it with its algorithmic HBM bytes.
bool profile_on();
void profile_set(bool on);
void profile_reset();
std::string profile_json();  // {"name": {"launches", "total_ms", "bytes"}}
// Interned "base<a>" / "base<a, b>": the profiler names a launch after its kernel
// template as rocprofv3 prints it, so the two reports line up.
const char* kernel_name(const char* base, int a, int b = -1);
struct ProfScope {
  ProfScope(const char* name, hipStream_t s, double bytes);
  ~ProfScope();
  const char* name;
  hipStream_t s;
  double bytes;
  hipEvent_t start = nullptr;
};

}  // namespace whir
