// Runtime plumbing: last-error storage, per-thread streams, zero fill.
#include "common.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

namespace whir {

static thread_local std::string g_last_error;

void set_last_error(const std::string& msg) { g_last_error = msg; }
const char* last_error() { return g_last_error.c_str(); }

namespace {
struct ThreadStream {
  hipStream_t s = nullptr;
  int device = -1;
  void release();
  ~ThreadStream() { release(); }
};
thread_local ThreadStream t_stream;
}  // namespace

namespace {
std::mutex g_pool_mu;
std::vector<hipMemPool_t> g_pools;  // per device: the library's own stream-ordered pool

// The library allocates from a pool of its own, so that keeping freed memory cached
// (release threshold UINT64_MAX: with the default 0 every synchronization hands it
// back and the next proof maps its half-GiB codewords again) leaves the device's
// default pool, which the host application may use, as it is.
hipMemPool_t device_pool(int dev) {
  std::lock_guard<std::mutex> g(g_pool_mu);
  if ((size_t)dev >= g_pools.size()) g_pools.resize(dev + 1, nullptr);
  if (!g_pools[dev]) {
    hipMemPoolProps props = {};
    props.allocType = hipMemAllocationTypePinned;
    props.handleTypes = hipMemHandleTypeNone;
    props.location.type = hipMemLocationTypeDevice;
    props.location.id = dev;
    hipMemPool_t pool = nullptr;
    WHIR_HIP(hipMemPoolCreate(&pool, &props));
    uint64_t thr = UINT64_MAX;
    WHIR_HIP(hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &thr));
    g_pools[dev] = pool;
  }
  return g_pools[dev];
}

// Per-thread pinned staging area for small host->device uploads, so they can be
// enqueued without waiting for the stream.  Regions are reused only after the
// whole area has wrapped and the copies enqueued before the wrap have finished.
struct Staging {
  uint8_t* host = nullptr;
  size_t cap = 0, off = 0;
  hipEvent_t wrap = nullptr;
  int device = -1;
  void release() {
    if (host) (void)hipHostFree(host);
    if (wrap) (void)hipEventDestroy(wrap);
    host = nullptr;
    wrap = nullptr;
    cap = off = 0;
    device = -1;
  }
  ~Staging() { release(); }
};
thread_local Staging t_staging;
}  // namespace

namespace {
std::mutex g_cache_mu;
// (stream, size rounded to 256 B) -> free buffers
std::map<std::pair<hipStream_t, size_t>, std::vector<void*>> g_cache;
std::map<hipStream_t, size_t> g_cached_bytes;  // bytes held in each stream's free lists
size_t g_cached_total = 0;
size_t round_bytes(size_t b) { return (b + 255) & ~(size_t)255; }

hipError_t alloc_from_pool(void** p, size_t rb, hipStream_t s) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  return hipMallocFromPoolAsync(p, rb, device_pool(dev), s);
}
}  // namespace

void* pool_alloc(size_t bytes, hipStream_t s) {
  size_t rb = round_bytes(bytes);
  {
    std::lock_guard<std::mutex> g(g_cache_mu);
    auto it = g_cache.find({s, rb});
    if (it != g_cache.end() && !it->second.empty()) {
      void* p = it->second.back();
      it->second.pop_back();
      g_cached_bytes[s] -= rb;
      g_cached_total -= rb;
      return p;
    }
  }
  void* p = nullptr;
  hipError_t e = alloc_from_pool(&p, rb, s);
  if (e == hipErrorOutOfMemory) {
    // Buffers cached by this and other streams hold device memory nobody uses: hand
    // them back (each on its own stream, after that stream's earlier work) and try once more.
    (void)hipGetLastError();
    pool_trim_all();
    e = alloc_from_pool(&p, rb, s);
  }
  if (e != hipSuccess)
    throw Error(std::string("device allocation of ") + std::to_string(rb) + " bytes: " + hipGetErrorString(e));
  return p;
}

void pool_free(void* p, size_t bytes, hipStream_t s) {
  const size_t rb = round_bytes(bytes);
  std::lock_guard<std::mutex> g(g_cache_mu);
  g_cache[{s, rb}].push_back(p);
  g_cached_bytes[s] += rb;
  g_cached_total += rb;
}

size_t pool_cached_bytes(hipStream_t s) {
  std::lock_guard<std::mutex> g(g_cache_mu);
  if (!s) return g_cached_total;
  auto it = g_cached_bytes.find(s);
  return it == g_cached_bytes.end() ? 0 : it->second;
}

// Takes every cached buffer of stream s (all streams when s is null) out of the lists.
static std::vector<std::pair<hipStream_t, void*>> take_cached(hipStream_t s) {
  std::vector<std::pair<hipStream_t, void*>> drop;
  std::lock_guard<std::mutex> g(g_cache_mu);
  for (auto it = g_cache.begin(); it != g_cache.end();) {
    if (!s || it->first.first == s) {
      for (void* p : it->second) drop.emplace_back(it->first.first, p);
      g_cached_bytes[it->first.first] -= it->first.second * it->second.size();
      g_cached_total -= it->first.second * it->second.size();
      it = g_cache.erase(it);
    } else {
      ++it;
    }
  }
  return drop;
}

static void free_taken(const std::vector<std::pair<hipStream_t, void*>>& drop) {
  std::vector<hipStream_t> streams;
  for (const auto& d : drop) {
    (void)hipFreeAsync(d.second, d.first);
    if (std::find(streams.begin(), streams.end(), d.first) == streams.end()) streams.push_back(d.first);
  }
  for (hipStream_t st : streams) (void)hipStreamSynchronize(st);
  if (!drop.empty()) {
    int dev = 0;
    // Give the freed memory back to the device too (the pool would otherwise keep it
    // reserved for its next allocations).
    if (hipGetDevice(&dev) == hipSuccess) (void)hipMemPoolTrimTo(device_pool(dev), 0);
  }
}

void pool_trim(hipStream_t s) {
  if (s) free_taken(take_cached(s));
}

void pool_trim_all() { free_taken(take_cached(nullptr)); }

static void release_stream_slots(hipStream_t s);

void ThreadStream::release() {
  if (s) {
    pool_trim(s);
    (void)hipStreamSynchronize(s);
    release_stream_slots(s);
    (void)hipStreamDestroy(s);
    s = nullptr;
    device = -1;
  }
}

hipStream_t thread_stream() {
  int dev = 0;
  WHIR_HIP(hipGetDevice(&dev));
  if (!t_stream.s || t_stream.device != dev) {
    if (t_stream.s) (void)hipStreamDestroy(t_stream.s);
    (void)device_pool(dev);
    WHIR_HIP(hipStreamCreateWithFlags(&t_stream.s, hipStreamNonBlocking));
    t_stream.device = dev;
  }
  return t_stream.s;
}

void upload_async(void* dst, const void* src, size_t bytes, hipStream_t s) {
  if (!bytes) return;
  Staging& st = t_staging;
  int dev = 0;
  WHIR_HIP(hipGetDevice(&dev));
  if (!st.host || st.device != dev) {
    if (st.host) (void)hipHostFree(st.host);
    if (st.wrap) (void)hipEventDestroy(st.wrap);
    st.cap = (size_t)8 << 20;
    WHIR_HIP(hipHostMalloc((void**)&st.host, st.cap, hipHostMallocDefault));
    WHIR_HIP(hipEventCreateWithFlags(&st.wrap, hipEventDisableTiming));
    st.off = 0;
    st.device = dev;
  }
  if (bytes > st.cap / 4) {  // large: plain copy, then wait so `src` may be released
    WHIR_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
    WHIR_HIP(hipStreamSynchronize(s));
    return;
  }
  if (st.off + bytes > st.cap) {
    WHIR_HIP(hipEventRecord(st.wrap, s));
    WHIR_HIP(hipEventSynchronize(st.wrap));
    st.off = 0;
  }
  memcpy(st.host + st.off, src, bytes);
  WHIR_HIP(hipMemcpyAsync(dst, st.host + st.off, bytes, hipMemcpyHostToDevice, s));
  st.off += (bytes + 255) & ~(size_t)255;
}

namespace {
struct Readback {
  uint8_t* host = nullptr;
  size_t cap = 0;
  int device = -1;
  void release() {
    if (host) (void)hipHostFree(host);
    host = nullptr;
    cap = 0;
    device = -1;
  }
  ~Readback() { release(); }
};
thread_local Readback t_readback;
}  // namespace

void release_thread_resources() {
  release_commit_plans();
  t_stream.release();
  t_staging.release();
  t_readback.release();
}

void download_sync(void* dst, const void* src, size_t bytes, hipStream_t s) {
  if (!bytes) {
    WHIR_HIP(hipStreamSynchronize(s));
    return;
  }
  Readback& rb = t_readback;
  int dev = 0;
  WHIR_HIP(hipGetDevice(&dev));
  if (!rb.host || rb.device != dev || rb.cap < bytes) {
    if (rb.host) (void)hipHostFree(rb.host);
    rb.cap = std::max<size_t>(bytes, (size_t)1 << 20);
    WHIR_HIP(hipHostMalloc((void**)&rb.host, rb.cap, hipHostMallocDefault));
    rb.device = dev;
  }
  WHIR_HIP(hipMemcpyAsync(rb.host, src, bytes, hipMemcpyDeviceToHost, s));
  WHIR_HIP(hipStreamSynchronize(s));
  memcpy(dst, rb.host, bytes);
}

__global__ void k_fill_zero(uint4* __restrict__ p, size_t n16, uint8_t* __restrict__ tail, size_t ntail) {
  size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t k = i; k < n16; k += stride) p[k] = make_uint4(0, 0, 0, 0);
  if (i < ntail) tail[i] = 0;
}

// ------------------------------------------------------------------ profiler
namespace {
struct ProfRecord {
  hipEvent_t start, stop;
  double bytes;
};
std::mutex g_prof_mu;
bool g_prof_on = false;
std::map<std::string, std::vector<ProfRecord>> g_prof;
std::vector<hipEvent_t> g_event_pool;  // recycled events (creating one per launch costs more than the launch)

hipEvent_t take_event() {
  {
    std::lock_guard<std::mutex> g(g_prof_mu);
    if (!g_event_pool.empty()) {
      hipEvent_t e = g_event_pool.back();
      g_event_pool.pop_back();
      return e;
    }
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}
}  // namespace

bool profile_on() { return g_prof_on; }

static thread_local bool t_capture = false;
bool capture_active() { return t_capture; }
void set_capture_active(bool on) { t_capture = on; }

const char* kernel_name(const char* base, int a, int b) {
  static std::mutex mu;
  static std::map<std::tuple<std::string, int, int>, std::string> names;
  std::lock_guard<std::mutex> g(mu);
  auto& s = names[std::make_tuple(std::string(base), a, b)];
  if (s.empty()) {
    s = std::string(base) + "<" + std::to_string(a);
    if (b >= 0) s += ", " + std::to_string(b);
    s += ">";
  }
  return s.c_str();
}

ProfScope::ProfScope(const char* n, hipStream_t st, double b) : name(n), s(st), bytes(b) {
  if (!g_prof_on || t_capture) return;
  start = take_event();
  if (start) (void)hipEventRecord(start, s);
}

ProfScope::~ProfScope() {
  if (!start) return;
  hipEvent_t stop = take_event();
  if (!stop) return;
  (void)hipEventRecord(stop, s);
  std::lock_guard<std::mutex> g(g_prof_mu);
  g_prof[name].push_back({start, stop, bytes});
}

void profile_set(bool on) {
  g_prof_on = on;
  if (on) {  // pre-create a pool so the timed region records into existing events
    std::vector<hipEvent_t> fresh;
    for (int i = 0; i < 2048; i++) {
      hipEvent_t e = nullptr;
      if (hipEventCreate(&e) == hipSuccess) fresh.push_back(e);
    }
    std::lock_guard<std::mutex> g(g_prof_mu);
    g_event_pool.insert(g_event_pool.end(), fresh.begin(), fresh.end());
  }
}

void profile_reset() {
  std::lock_guard<std::mutex> g(g_prof_mu);
  for (auto& kv : g_prof)
    for (auto& r : kv.second) {
      (void)hipEventSynchronize(r.stop);
      g_event_pool.push_back(r.start);
      g_event_pool.push_back(r.stop);
    }
  g_prof.clear();
}

std::string profile_json() {
  std::lock_guard<std::mutex> g(g_prof_mu);
  std::string out = "{";
  bool first = true;
  for (auto& kv : g_prof) {
    double ms = 0, bytes = 0;
    for (auto& r : kv.second) {
      float e = 0;
      (void)hipEventSynchronize(r.stop);
      if (hipEventElapsedTime(&e, r.start, r.stop) == hipSuccess) ms += e;
      bytes += r.bytes;
    }
    char buf[256];
    snprintf(buf, sizeof buf, "%s\"%s\": {\"launches\": %zu, \"total_ms\": %.6f, \"bytes\": %.0f}", first ? "" : ", ",
             kv.first.c_str(), kv.second.size(), ms, bytes);
    out += buf;
    first = false;
  }
  return out + "}";
}

namespace {
// The flag word of mailbox_wait sits in its own 64-byte line after the mailbox.
constexpr size_t MAILBOX_FLAG_WORD = RESULT_MAILBOX_WORDS + 8;
struct StreamSlots {
  uint64_t* mailbox = nullptr;
  uint64_t seq = 0;  // last sequence number mailbox_wait queued on the stream
};
std::mutex g_slots_mu;
std::map<hipStream_t, StreamSlots> g_slots;

StreamSlots& stream_slots(hipStream_t s) {
  std::lock_guard<std::mutex> g(g_slots_mu);
  StreamSlots& sl = g_slots[s];
  if (!sl.mailbox) {
    WHIR_HIP(hipHostMalloc((void**)&sl.mailbox, (MAILBOX_FLAG_WORD + 8) * 8, hipHostMallocCoherent));
    sl.mailbox[MAILBOX_FLAG_WORD] = 0;
  }
  return sl;  // std::map references stay valid until the entry is erased
}

__global__ void k_mailbox_signal(uint64_t* flag, uint64_t seq) {
  if (threadIdx.x == 0) {
    __threadfence_system();
    __hip_atomic_store(flag, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
  }
}
}  // namespace

uint64_t* result_mailbox(hipStream_t s) { return stream_slots(s).mailbox; }

namespace {
// The calling thread's armed wait: stream, its mailbox, the sequence number, and
// whether a launch has taken the request.
struct ArmedWait {
  hipStream_t s = nullptr;
  uint64_t* mbox = nullptr;
  uint64_t seq = 0;
  bool armed = false;
  bool taken = false;
};
thread_local ArmedWait t_armed;
}  // namespace

// WHIR_MAILBOX_SYNC=1: mailbox_wait synchronizes the stream and nothing is armed (the
// plain wait, kept for comparison).
static bool mailbox_plain_sync() {
  static const bool on = [] {
    const char* e = getenv("WHIR_MAILBOX_SYNC");
    return e && atoi(e) != 0;
  }();
  return on;
}

void mailbox_arm(hipStream_t s) {
  if (mailbox_plain_sync()) return;
  StreamSlots& sl = stream_slots(s);
  t_armed = ArmedWait{s, sl.mailbox, ++sl.seq, true, false};
}

MailboxSig mailbox_take(const void* out) {
  ArmedWait& a = t_armed;
  const uint64_t* p = (const uint64_t*)out;
  if (!a.armed || p < a.mbox || p >= a.mbox + RESULT_MAILBOX_WORDS) return {};
  a.armed = false;
  a.taken = true;
  return MailboxSig{a.mbox + MAILBOX_FLAG_WORD, a.seq};
}

void mailbox_wait(hipStream_t s) {
  WHIR_REQUIRE(!capture_active(), "mailbox_wait during a stream capture");
  if (mailbox_plain_sync()) {
    t_armed = ArmedWait{};
    WHIR_HIP(hipStreamSynchronize(s));
    return;
  }
  StreamSlots& sl = stream_slots(s);
  uint64_t* flag = sl.mailbox + MAILBOX_FLAG_WORD;
  uint64_t want;
  bool signal = true;
  if (t_armed.s == s && (t_armed.armed || t_armed.taken)) {
    want = t_armed.seq;
    signal = !t_armed.taken;
  } else {
    want = ++sl.seq;
  }
  t_armed = ArmedWait{};
  if (signal) {
    k_mailbox_signal<<<1, 64, 0, s>>>(flag, want);
    WHIR_HIP(hipGetLastError());
  }
  for (uint32_t n = 1; __atomic_load_n(flag, __ATOMIC_ACQUIRE) < want; n++) {
    if ((n & 4095) == 0) {
      // Every few thousand polls ask the runtime: an error ends the wait with that
      // error; an idle stream whose flag has not shown yet falls back to a plain
      // synchronization.
      const hipError_t e = hipStreamQuery(s);
      if (e == hipSuccess) {
        WHIR_HIP(hipStreamSynchronize(s));
        break;
      }
      if (e != hipErrorNotReady) WHIR_HIP(e);
    }
    __builtin_ia32_pause();
  }
}

static void release_stream_slots(hipStream_t s) {
  std::lock_guard<std::mutex> g(g_slots_mu);
  auto it = g_slots.find(s);
  if (it == g_slots.end()) return;
  if (it->second.mailbox) (void)hipHostFree(it->second.mailbox);
  g_slots.erase(it);
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
