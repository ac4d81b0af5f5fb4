// extern "C" boundary of libwhir_amd (declarations in include/whir_amd.h).
#include <math.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "../../include/whir_amd.h"
#include "algebra.h"
#include "common.h"
#include "gl.h"
#include "hash.h"
#include "keccak_host.h"
#include "ntt.h"

using namespace whir;

#define API_BEGIN try {
#define API_END                                \
  }                                            \
  catch (const std::exception& e) {            \
    set_last_error(e.what());                  \
    return 1;                                  \
  }                                            \
  catch (...) {                                \
    set_last_error("unknown error");           \
    return 1;                                  \
  }                                            \
  set_last_error("");                          \
  return 0;

namespace {

void check_degree(int D) { WHIR_REQUIRE(D >= 1 && D <= 3, "degree must be 1, 2 or 3"); }

struct HostIO {
  hipStream_t s;
  explicit HostIO() : s(thread_stream()) {}
  DevBuf up(const void* src, size_t bytes) {
    DevBuf b(bytes, s);
    if (bytes) WHIR_HIP(hipMemcpyAsync(b.ptr, src, bytes, hipMemcpyHostToDevice, s));
    return b;
  }
  void down(void* dst, const void* src, size_t bytes) {
    if (bytes) WHIR_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
  }
  void sync() { WHIR_HIP(hipStreamSynchronize(s)); }
};

uint64_t next_order_u64(uint64_t size) {
  uint64_t best = 0;
  for (int a = 0; a <= 32; a++) {
    uint64_t c1 = 1ull << a, c3 = 3ull << a;
    if (c1 >= size && (!best || c1 < best)) best = c1;
    if (c3 >= size && (!best || c3 < best)) best = c3;
  }
  WHIR_REQUIRE(best, "next_order: size exceeds the largest supported order 3*2^32");
  return best;
}

// coset size of cooley_tukey.rs:436-439
uint64_t coset_size(uint64_t masked_len, uint64_t N) {
  uint64_t c = next_order_u64(masked_len);
  while (N % c) c = next_order_u64(c + 1);
  return c;
}

int log2_exact(uint64_t n) {
  int k = 0;
  while ((1ull << k) < n) k++;
  return (1ull << k) == n ? k : -1;
}

void rs_encode_common(int D, const uint64_t* d_msgs, size_t M, size_t L, const uint64_t* d_masks, size_t ml,
                      size_t N, uint64_t* d_out_cols, hipStream_t s) {
  WHIR_REQUIRE(N > 0 && gl::NTT_ORDER % N == 0, "interleaved_encode: codeword length must divide 3*2^32");
  WHIR_REQUIRE(L + ml <= N, "interleaved_encode: masked message longer than the codeword");
  uint64_t C = coset_size(L + ml, N);
  int logC = log2_exact(C);
  WHIR_REQUIRE(logC >= 0 && log2_exact(N) >= 0,
               "interleaved_encode: the GPU engine needs power-of-two codeword and coset sizes");
  RsArgs a{d_msgs, L, L, d_masks, ml, d_out_cols, N, N / C, logC, M};
  size_t sw = rs_scratch_words(D, M, N, C);
  DevBuf scratch(sw * 8, s);
  rs_encode_cols(D, a, scratch.as<uint64_t>(), s);
}

const char* kEngineNames[5] = {"copy", nullptr, "keccak", nullptr, "Blake3"};
// DER bodies of the OIDs of SHA-256 (2.16.840.1.101.3.4.2.1) and SHA3-256 (.8).
const uint8_t kOidSha256[9] = {0x60, 0x86, 0x48, 0x01, 0x65, 0x03, 0x04, 0x02, 0x01};
const uint8_t kOidSha3[9] = {0x60, 0x86, 0x48, 0x01, 0x65, 0x03, 0x04, 0x02, 0x08};

}  // namespace

extern "C" {

const char* whir_last_error(void) { return last_error(); }
int whir_abi_version(void) { return 1; }

int whir_device_count(int* count) {
  API_BEGIN
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
  *count = n;
  API_END
}

int whir_set_device(int device) {
  API_BEGIN
  WHIR_HIP(hipSetDevice(device));
  API_END
}

int whir_dev_alloc(size_t bytes, void** ptr) {
  API_BEGIN
  WHIR_HIP(hipMalloc(ptr, std::max<size_t>(bytes, 1)));
  API_END
}
int whir_dev_free(void* ptr) {
  API_BEGIN
  WHIR_HIP(hipFree(ptr));
  API_END
}
int whir_host_alloc(size_t bytes, void** ptr) {
  API_BEGIN
  WHIR_HIP(hipHostMalloc(ptr, std::max<size_t>(bytes, 1), hipHostMallocDefault));
  API_END
}
int whir_host_free(void* ptr) {
  API_BEGIN
  WHIR_HIP(hipHostFree(ptr));
  API_END
}
int whir_memcpy_h2d(void* dst, const void* src, size_t bytes) {
  API_BEGIN
  hipStream_t s = thread_stream();
  WHIR_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
  WHIR_HIP(hipStreamSynchronize(s));
  API_END
}
int whir_memcpy_d2h(void* dst, const void* src, size_t bytes) {
  API_BEGIN
  hipStream_t s = thread_stream();
  WHIR_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
  WHIR_HIP(hipStreamSynchronize(s));
  API_END
}
int whir_memcpy_d2d(void* dst, const void* src, size_t bytes) {
  API_BEGIN
  hipStream_t s = thread_stream();
  WHIR_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, s));
  API_END
}
int whir_synchronize(void) {
  API_BEGIN
  WHIR_HIP(hipStreamSynchronize(thread_stream()));
  API_END
}

int whir_profile_enable(int on) {
  API_BEGIN
  profile_set(on != 0);
  API_END
}

int whir_profile_reset(void) {
  API_BEGIN
  profile_reset();
  API_END
}

int whir_profile_report(char* buf, size_t cap, size_t* len) {
  API_BEGIN
  std::string s = profile_json();
  *len = s.size();
  if (buf && cap) {
    size_t n = std::min(cap - 1, s.size());
    memcpy(buf, s.data(), n);
    buf[n] = 0;
  }
  API_END
}

// ------------------------------------------------------------------ ReedSolomon
int whir_rs_next_order(uint64_t size, uint64_t* out) {
  API_BEGIN
  *out = next_order_u64(size);
  API_END
}

int whir_rs_generator(uint64_t N, uint64_t* out) {
  API_BEGIN
  WHIR_REQUIRE(N > 0 && gl::NTT_ORDER % N == 0, "generator: codeword length must divide 3*2^32");
  *out = gl::root_of_unity(N);
  API_END
}

int whir_rs_evaluation_points(uint64_t masked_len, uint64_t N, const uint64_t* idx, size_t count, uint64_t* out) {
  API_BEGIN
  WHIR_REQUIRE(masked_len <= N, "evaluation_points: masked length exceeds codeword length");
  WHIR_REQUIRE(N > 0 && gl::NTT_ORDER % N == 0, "evaluation_points: codeword length must divide 3*2^32");
  uint64_t g = gl::root_of_unity(N);
  uint64_t C = coset_size(masked_len, N), Q = N / C;
  for (size_t k = 0; k < count; k++) {
    WHIR_REQUIRE(idx[k] < N, "evaluation_points: index out of range");
    uint64_t row = idx[k] / C, col = idx[k] % C;
    out[k] = gl::pow(g, row + col * Q);
  }
  API_END
}

int whir_rs_interleaved_encode(int D, const uint64_t* const* messages, size_t M, size_t L, const uint64_t* masks,
                               size_t ml, size_t N, uint64_t* out) {
  API_BEGIN
  check_degree(D);
  if (M == 0) {
    WHIR_REQUIRE(ml == 0, "interleaved_encode: masks without messages");
    return 0;
  }
  HostIO io;
  const size_t ew = (size_t)D;
  DevBuf msgs(M * L * ew * 8, io.s);
  for (size_t j = 0; j < M; j++)
    if (L) WHIR_HIP(hipMemcpyAsync(msgs.as<uint64_t>() + j * L * ew, messages[j], L * ew * 8, hipMemcpyHostToDevice, io.s));
  DevBuf dmask = io.up(masks, M * ml * ew * 8);
  DevBuf cols(M * N * ew * 8, io.s), rows(M * N * ew * 8, io.s);
  rs_encode_common(D, msgs.as<uint64_t>(), M, L, ml ? dmask.as<uint64_t>() : nullptr, ml, N, cols.as<uint64_t>(), io.s);
  cols_to_rows(D, cols.as<uint64_t>(), rows.as<uint64_t>(), N, M, io.s);
  io.down(out, rows.ptr, M * N * ew * 8);
  io.sync();
  API_END
}

int whir_rs_interleaved_encode_dev(int D, const uint64_t* msgs, size_t M, size_t L, const uint64_t* masks, size_t ml,
                                   size_t N, uint64_t* out, int row_major) {
  API_BEGIN
  check_degree(D);
  if (M == 0) return 0;
  hipStream_t s = thread_stream();
  if (!row_major) {
    rs_encode_common(D, msgs, M, L, masks, ml, N, out, s);
  } else {
    DevBuf cols(M * N * D * 8, s);
    rs_encode_common(D, msgs, M, L, masks, ml, N, cols.as<uint64_t>(), s);
    cols_to_rows(D, cols.as<uint64_t>(), out, N, M, s);
  }
  API_END
}

int whir_ntt_autotune(int enable) {
  API_BEGIN
  ntt_autotune(enable != 0);
  API_END
}

int whir_ntt_set_schedule(int D, size_t log_n, size_t log_c, size_t messages, int log_c1, int variant1, int tile1,
                          int variant2, int tile2) {
  API_BEGIN
  WHIR_REQUIRE(log_n <= 40 && log_c <= 40, "ntt schedule: sizes out of range");
  ntt_set_schedule(D, (int)log_n, (int)log_c, messages, log_c1, variant1, tile1, variant2, tile2);
  API_END
}

int whir_ntt_clear_schedules(void) {
  API_BEGIN
  ntt_clear_schedules();
  API_END
}

int whir_ntt_schedules(char* buf, size_t cap, size_t* len) {
  API_BEGIN
  WHIR_REQUIRE(len, "ntt_schedules: null length");
  std::string js = ntt_schedules_json();
  *len = js.size();
  if (buf && cap) {
    size_t n = std::min(cap - 1, js.size());
    memcpy(buf, js.data(), n);
    buf[n] = 0;
  }
  API_END
}

// ------------------------------------------------------------------ hashes
int whir_hash_engine_id(int engine, uint8_t out[32]) {
  API_BEGIN
  WHIR_REQUIRE(engine >= 0 && engine <= 4, "unknown hash engine");
  std::vector<uint8_t> msg = {'w', 'h', 'i', 'r', ':', ':', 'h', 'a', 's', 'h'};
  if (engine == WHIR_HASH_SHA2) msg.insert(msg.end(), kOidSha256, kOidSha256 + 9);
  else if (engine == WHIR_HASH_SHA3) msg.insert(msg.end(), kOidSha3, kOidSha3 + 9);
  else msg.insert(msg.end(), kEngineNames[engine], kEngineNames[engine] + strlen(kEngineNames[engine]));
  auto h = sha3_256(msg.data(), msg.size());
  memcpy(out, h.data(), 32);
  API_END
}

int whir_hash_supports_size(int engine, size_t size, int* out) {
  API_BEGIN
  *out = engine_supports_size(engine, size) ? 1 : 0;
  API_END
}

int whir_hash_preferred_batch_size(int engine, size_t* out) {
  API_BEGIN
  *out = engine_preferred_batch(engine);
  API_END
}

int whir_hash_many(int engine, size_t size, const uint8_t* input, size_t count, uint8_t* output) {
  API_BEGIN
  WHIR_REQUIRE(engine_supports_size(engine, size), "hash_many: engine does not support this message size");
  if (!count) return 0;
  HostIO io;
  DevBuf in = io.up(input, size * count);
  DevBuf out(32 * count, io.s);
  hash_many_dev(engine, size, in.as<uint8_t>(), count, out.as<uint8_t>(), io.s);
  io.down(output, out.ptr, 32 * count);
  io.sync();
  API_END
}

int whir_hash_many_dev(int engine, size_t size, const uint8_t* input, size_t count, uint8_t* output) {
  API_BEGIN
  hash_many_dev(engine, size, input, count, output, thread_stream());
  API_END
}

// ------------------------------------------------------------------ Merkle
int whir_merkle_num_nodes(size_t n, size_t* out) {
  API_BEGIN
  *out = merkle_num_nodes(n);
  API_END
}

int whir_merkle_commit(int engine, const uint8_t* leaves, size_t n, uint8_t* nodes) {
  API_BEGIN
  HostIO io;
  size_t nn = merkle_num_nodes(n);
  DevBuf d(32 * nn, io.s);
  if (n) WHIR_HIP(hipMemcpyAsync(d.ptr, leaves, 32 * n, hipMemcpyHostToDevice, io.s));
  merkle_build_dev(engine, d.as<uint8_t>(), n, io.s);
  io.down(nodes, d.ptr, 32 * nn);
  io.sync();
  API_END
}

int whir_merkle_commit_dev(int engine, uint8_t* nodes, size_t n) {
  API_BEGIN
  merkle_build_dev(engine, nodes, n, thread_stream());
  API_END
}

int whir_matrix_commit(int engine, int D, const uint64_t* matrix, size_t rows, size_t cols, uint8_t* nodes) {
  API_BEGIN
  check_degree(D);
  HostIO io;
  size_t nn = merkle_num_nodes(rows);
  DevBuf m = io.up(matrix, rows * cols * D * 8);
  DevBuf d(32 * nn, io.s);
  if (rows) hash_rows_dev(engine, D, m.as<uint64_t>(), rows, cols, d.as<uint8_t>(), io.s);
  merkle_build_dev(engine, d.as<uint8_t>(), rows, io.s);
  io.down(nodes, d.ptr, 32 * nn);
  io.sync();
  API_END
}

int whir_hash_rows_cols_dev(int engine, int D, const uint64_t* cw, size_t rows, size_t cols, uint8_t* leaves) {
  API_BEGIN
  check_degree(D);
  hash_rows_cols_dev(engine, D, cw, rows, cols, leaves, thread_stream());
  API_END
}

int whir_matrix_commit_cols_dev(int engine, int D, const uint64_t* cw, size_t rows, size_t cols, uint8_t* nodes) {
  API_BEGIN
  check_degree(D);
  commit_cols_dev(engine, D, cw, rows, cols, nodes, thread_stream());
  API_END
}

int whir_merkle_open(const uint8_t* nodes, size_t num_leaves, const size_t* indices, size_t count, uint8_t* hints,
                     size_t* num_hints) {
  API_BEGIN
  size_t layers = merkle_layers(num_leaves);
  std::vector<size_t> idx(indices, indices + count);
  for (size_t i : idx) WHIR_REQUIRE(i < num_leaves, "merkle open: index out of range");
  std::sort(idx.begin(), idx.end());
  idx.erase(std::unique(idx.begin(), idx.end()), idx.end());
  size_t off = 0, width = (size_t)1 << layers, nh = 0;
  while (width > 1) {
    std::vector<size_t> next;
    for (size_t k = 0; k < idx.size();) {
      size_t a = idx[k];
      if (k + 1 < idx.size() && idx[k + 1] == (a ^ 1)) {
        k += 2;
      } else {
        memcpy(hints + 32 * nh, nodes + 32 * (off + (a ^ 1)), 32);
        nh++;
        k += 1;
      }
      next.push_back(a >> 1);
    }
    idx.swap(next);
    off += width;
    width >>= 1;
  }
  *num_hints = nh;
  API_END
}

// ------------------------------------------------------------------ sumcheck / algebra
int whir_sumcheck_polynomial(int D, const uint64_t* a, const uint64_t* b, size_t n, uint64_t* out) {
  API_BEGIN
  check_degree(D);
  HostIO io;
  DevBuf da = io.up(a, n * D * 8), db = io.up(b, n * D * 8), o(2 * D * 8, io.s);
  sumcheck_poly_dev(D, da.as<uint64_t>(), db.as<uint64_t>(), n, o.as<uint64_t>(), io.s);
  io.down(out, o.ptr, 2 * D * 8);
  io.sync();
  API_END
}

int whir_sumcheck_polynomial_dev(int D, const uint64_t* a, const uint64_t* b, size_t n, uint64_t* out) {
  API_BEGIN
  check_degree(D);
  hipStream_t s = thread_stream();
  DevBuf o(2 * D * 8, s);
  sumcheck_poly_dev(D, a, b, n, o.as<uint64_t>(), s);
  WHIR_HIP(hipMemcpyAsync(out, o.ptr, 2 * D * 8, hipMemcpyDeviceToHost, s));
  WHIR_HIP(hipStreamSynchronize(s));
  API_END
}

int whir_fold(int D, uint64_t* v, size_t n, const uint64_t* w, size_t* new_len) {
  API_BEGIN
  check_degree(D);
  HostIO io;
  DevBuf dv = io.up(v, n * D * 8);
  size_t m = fold_dev(D, dv.as<uint64_t>(), n, w, io.s);
  io.down(v, dv.ptr, m * D * 8);
  io.sync();
  *new_len = m;
  API_END
}

int whir_fold_dev(int D, uint64_t* v, size_t n, const uint64_t* w, size_t* new_len) {
  API_BEGIN
  check_degree(D);
  *new_len = fold_dev(D, v, n, w, thread_stream());
  API_END
}

int whir_fold_and_compute_polynomial(int D, uint64_t* a, uint64_t* b, size_t n, const uint64_t* w, uint64_t* out,
                                     size_t* new_len) {
  API_BEGIN
  check_degree(D);
  HostIO io;
  DevBuf da = io.up(a, n * D * 8), db = io.up(b, n * D * 8), o(2 * D * 8, io.s);
  size_t m = fold_and_poly_dev(D, da.as<uint64_t>(), db.as<uint64_t>(), n, w, o.as<uint64_t>(), io.s);
  io.down(a, da.ptr, m * D * 8);
  io.down(b, db.ptr, m * D * 8);
  io.down(out, o.ptr, 2 * D * 8);
  io.sync();
  *new_len = m;
  API_END
}

int whir_fold_and_compute_polynomial_dev(int D, uint64_t* a, uint64_t* b, size_t n, const uint64_t* w, uint64_t* out,
                                         size_t* new_len) {
  API_BEGIN
  check_degree(D);
  hipStream_t s = thread_stream();
  DevBuf o(2 * D * 8, s);
  *new_len = fold_and_poly_dev(D, a, b, n, w, o.as<uint64_t>(), s);
  WHIR_HIP(hipMemcpyAsync(out, o.ptr, 2 * D * 8, hipMemcpyDeviceToHost, s));
  WHIR_HIP(hipStreamSynchronize(s));
  API_END
}

}  // extern "C"

// ------------------------------------------------------------------ allocation-free sumcheck round
// Device scratch of whir_fold_and_compute_polynomial_dev_noalloc: the block partials
// of the fused pass (at most 2048 blocks x 2 values x 3 words) and the 2 D-word
// result.  One copy per GPU, held until the process ends; every call on a GPU shares
// it, so calls that overlap on one GPU race.
static constexpr size_t NOALLOC_PART_WORDS = 2048 * 2 * 3;
static constexpr size_t NOALLOC_WORDS = NOALLOC_PART_WORDS + 2 * 3;
static __device__ __align__(256) uint64_t scratch[NOALLOC_WORDS];

namespace {
std::mutex g_scratch_mu;
std::vector<uint64_t*> g_scratch_ptr;  // per device: the array's device address (null until looked up)

// The device address of `scratch` on the current device, looked up once per device
// with that device current (host code must never use the array's name as an address).
uint64_t* scratch_on_current_device() {
  int dev = 0;
  WHIR_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if ((size_t)dev >= g_scratch_ptr.size()) g_scratch_ptr.resize(dev + 1, nullptr);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    WHIR_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    WHIR_REQUIRE(p, "noalloc: no device address for the scratch array");
    g_scratch_ptr[dev] = (uint64_t*)p;
  }
  return g_scratch_ptr[dev];
}
}  // namespace

extern "C" {

int whir_fold_and_compute_polynomial_dev_noalloc(int D, uint64_t* a, uint64_t* b, size_t n, const uint64_t* w,
                                                 uint64_t* out, size_t* new_len) {
  API_BEGIN
  check_degree(D);
  WHIR_REQUIRE(a && b && w && out && new_len, "fold_and_compute_polynomial_noalloc: null argument");
  WHIR_REQUIRE(n <= ((size_t)1 << 62), "fold_and_compute_polynomial_noalloc: length above 2^62");
  const size_t need = fold_poly_parts_words(D, n);
  WHIR_REQUIRE(need <= NOALLOC_PART_WORDS,
               "fold_and_compute_polynomial_noalloc: the pass needs more block partials than the scratch holds");
  uint64_t* ws = scratch_on_current_device();
  hipStream_t s = thread_stream();
  uint64_t* res = ws + NOALLOC_PART_WORDS;
  *new_len = fold_and_poly_ws_dev(D, a, b, n, w, res, need ? ws : nullptr, s);
  WHIR_HIP(hipMemcpyAsync(out, res, 2 * D * 8, hipMemcpyDeviceToHost, s));
  WHIR_HIP(hipStreamSynchronize(s));
  API_END
}

int whir_eval_eq(int D, uint64_t* acc, const uint64_t* point, size_t k, const uint64_t* scalar) {
  API_BEGIN
  check_degree(D);
  WHIR_REQUIRE(k < 40, "eval_eq: too many variables");
  HostIO io;
  size_t n = (size_t)1 << k;
  DevBuf da = io.up(acc, n * D * 8);
  eval_eq_dev(D, da.as<uint64_t>(), point, k, scalar, io.s);
  io.down(acc, da.ptr, n * D * 8);
  io.sync();
  API_END
}

int whir_eval_eq_dev(int D, uint64_t* acc, const uint64_t* point, size_t k, const uint64_t* scalar) {
  API_BEGIN
  check_degree(D);
  eval_eq_dev(D, acc, point, k, scalar, thread_stream());
  API_END
}

int whir_geometric_accumulate(int D, uint64_t* acc, size_t n, const uint64_t* scalars, const uint64_t* points, size_t m) {
  API_BEGIN
  check_degree(D);
  HostIO io;
  DevBuf da = io.up(acc, n * D * 8);
  geometric_accumulate_dev(D, da.as<uint64_t>(), n, scalars, points, m, io.s);
  io.down(acc, da.ptr, n * D * 8);
  io.sync();
  API_END
}

int whir_geometric_accumulate_dev(int D, uint64_t* acc, size_t n, const uint64_t* scalars, const uint64_t* points,
                                  size_t m) {
  API_BEGIN
  check_degree(D);
  geometric_accumulate_dev(D, acc, n, scalars, points, m, thread_stream());
  API_END
}

int whir_mixed_univariate_evaluate(int Ds, int D, const uint64_t* c, size_t n, const uint64_t* x, uint64_t* out) {
  API_BEGIN
  check_degree(D);
  HostIO io;
  DevBuf dc = io.up(c, n * Ds * 8), o(D * 8, io.s);
  univariate_eval_dev(Ds, D, dc.as<uint64_t>(), n, x, o.as<uint64_t>(), io.s);
  io.down(out, o.ptr, D * 8);
  io.sync();
  API_END
}

int whir_mixed_univariate_evaluate_dev(int Ds, int D, const uint64_t* c, size_t n, const uint64_t* x, uint64_t* out) {
  API_BEGIN
  check_degree(D);
  hipStream_t s = thread_stream();
  DevBuf o(D * 8, s);
  univariate_eval_dev(Ds, D, c, n, x, o.as<uint64_t>(), s);
  WHIR_HIP(hipMemcpyAsync(out, o.ptr, D * 8, hipMemcpyDeviceToHost, s));
  WHIR_HIP(hipStreamSynchronize(s));
  API_END
}

int whir_lift(int D, const uint64_t* src, size_t n, uint64_t* dst) {
  API_BEGIN
  check_degree(D);
  HostIO io;
  DevBuf ds = io.up(src, n * 8), dd(n * D * 8, io.s);
  lift_dev(D, ds.as<uint64_t>(), n, dd.as<uint64_t>(), io.s);
  io.down(dst, dd.ptr, n * D * 8);
  io.sync();
  API_END
}

int whir_lift_dev(int D, const uint64_t* src, size_t n, uint64_t* dst) {
  API_BEGIN
  check_degree(D);
  lift_dev(D, src, n, dst, thread_stream());
  API_END
}

int whir_mixed_scalar_mul_add(int D, uint64_t* acc, const uint64_t* w, const uint64_t* v, size_t n) {
  API_BEGIN
  check_degree(D);
  HostIO io;
  DevBuf da = io.up(acc, n * D * 8), dv = io.up(v, n * 8);
  mixed_scalar_mul_add_dev(D, da.as<uint64_t>(), w, dv.as<uint64_t>(), n, io.s);
  io.down(acc, da.ptr, n * D * 8);
  io.sync();
  API_END
}

int whir_mixed_scalar_mul_add_dev(int D, uint64_t* acc, const uint64_t* w, const uint64_t* v, size_t n) {
  API_BEGIN
  check_degree(D);
  mixed_scalar_mul_add_dev(D, acc, w, v, n, thread_stream());
  API_END
}

int whir_dot(int D, const uint64_t* a, const uint64_t* b, size_t n, uint64_t* out) {
  API_BEGIN
  check_degree(D);
  HostIO io;
  DevBuf da = io.up(a, n * D * 8), db = io.up(b, n * D * 8), o(D * 8, io.s);
  dot_dev(D, da.as<uint64_t>(), db.as<uint64_t>(), n, o.as<uint64_t>(), io.s);
  io.down(out, o.ptr, D * 8);
  io.sync();
  API_END
}

// ------------------------------------------------------------------ PoW
int whir_pow_threshold(double bits, uint64_t* out) {
  API_BEGIN
  WHIR_REQUIRE(bits >= 0.0 && bits <= 60.0, "pow threshold: difficulty must be within [0, 60] bits");
  double t = ceil(exp2(64.0 - bits));
  *out = t >= 18446744073709551615.0 ? UINT64_MAX : (uint64_t)t;
  API_END
}

int whir_pow_grind(int engine, const uint8_t challenge[32], uint64_t threshold, uint64_t* nonce, int* performed) {
  API_BEGIN
  if (threshold == UINT64_MAX) {
    *nonce = 0;
    *performed = 0;
    return 0;
  }
  // Windows of 2^32 nonces, each one bounded launch, scanned in increasing order:
  // the first window with a solution holds the minimal nonce.
  const uint64_t window = 1ull << 32;
  uint64_t n = 0;
  for (uint64_t first = 0;; first += window) {
    if (pow_grind_dev(engine, challenge, threshold, first, window, &n, thread_stream())) break;
    WHIR_REQUIRE(first <= UINT64_MAX - 2 * window, "Proof of Work failed to solve.");
  }
  *nonce = n;
  *performed = 1;
  API_END
}

int whir_pow_grind_bounded(int engine, const uint8_t challenge[32], uint64_t threshold, uint64_t max_nonces,
                           uint64_t* nonce, int* found) {
  API_BEGIN
  uint64_t c = pow_chunk();
  uint64_t count = (max_nonces / c + (max_nonces % c != 0)) * c;
  WHIR_REQUIRE(count > 0, "pow: empty search window");
  uint64_t n = 0;
  bool ok = pow_grind_dev(engine, challenge, threshold, 0, count, &n, thread_stream());
  *found = ok && n < max_nonces;
  *nonce = *found ? n : 0;
  API_END
}

}  // extern "C"
