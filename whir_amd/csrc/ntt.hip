// Reed-Solomon expansion (interleaved coset NTT) for CDNA4.
//
// Replaces NttEngine::interleaved_encode (src/algebra/ntt/cooley_tukey.rs:411-477).
// For message j (length L, zero padded to the coset size C = next_order(L)) and
// codeword index r = c*C + t (coset-major, the reference's default order):
//
//     M[r][j] = sum_i m_j[i] * w_N^{(c + t*Q) i} = DFT_C( m_j[i] * w_N^{c i} )[t],
//
// with w_N = 7^((p-1)/N) and Q = N/C.  The device layout of the codeword is
// column-major (message-major): cw[j*N + r].  Leaf hashing reads it in tiles of
// whole rows, so the row-major matrix only has to be formed for host callers.
//
// DFT_C of length C = 2^n > 2^11 is a four-step transform over C = C1 * C2:
//   pass 1, per column i1: Y[t2][i1] = w_N^{i1 (c + Q t2)} sum_i2 x[i1 + C1 i2] w_N^{c C1 i2} w_C2^{i2 t2}
//   pass 2, per row t2:    X[t2 + C2 t1] = sum_i1 Y[t2][i1] w_C1^{i1 t1}
// The coset twist w_N^{c i} = w_N^{c i1} w_N^{c C1 i2} is split: the column-constant
// factor commutes with the column DFT and merges with the four-step twiddle
// w_C^{i1 t2} into one post factor w_N^{i1 (c + Q t2)}.  Both factors come from
// tables built once per shape (N words each), which every message and coset
// shares, so they stay in L2/MALL while the data streams through HBM.
// Goldilocks multiplication costs ~4x an addition on CDNA4, so the kernels are
// VALU-bound: the in-LDS transforms use radix-8 register butterflies whose inner
// rotations are the shifts w_8 = 2^120 and w_4 = 2^48 (2 has order 192 mod p),
// leaving 7/8 of a generic multiplication per element per three radix-2 stages.
#include "common.h"
#include "gl.h"
#include "ntt.h"

#include <algorithm>
#include <cstdio>
#include <string>
#include <map>
#include <tuple>
#include <mutex>
#include <vector>

namespace whir {

using gl::Ext;

// Threads per block: pass 1 runs 256 (its per-thread post factors and message loop
// favour fewer, fuller threads); pass 2 and the single pass run 512, which holds
// twice the waves per LDS tile (measured 5-10% faster on MI355X than 256).
#ifndef NTT_REM_FIRST
#define NTT_REM_FIRST 1
#endif
static constexpr int P1_THREADS = 256;
static constexpr int NTT_THREADS = 512;
// Pass 2 of Fp3 runs 256-thread blocks (1.46 -> 1.27-1.31 ms for 2^26 outputs on
// MI355X: Fp3 rows fill fewer waves per tile); Fp and Fp2 keep 512 (256 measured 2-14% slower).
template <int D>
constexpr int pass2_threads() { return D == 3 ? 256 : NTT_THREADS; }
static constexpr int MAX_PASS_LOG = 11;   // largest DFT staged in one pass (a 2^11 Fp3
                                          // column is 48 KiB, +16 KiB of twiddles)

// u64 words of data per LDS tile: 4096 (32 KiB) unless WHIR_NTT_TILE_WORDS asks for
// more and the device allows it (gfx950 has 160 KiB of LDS per workgroup).
static int tile_words() {
  static int w = [] {
    int words = 4096;
    if (const char* e = getenv("WHIR_NTT_TILE_WORDS")) words = atoi(e);
    int dev = 0, lds = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess) {
      while (words > 1024 && (size_t)(words + 2048) * 8 > (size_t)lds) words /= 2;
    }
    int p = 1024;
    while (p * 2 <= words) p *= 2;
    return p;
  }();
  return w;
}

template <class K>
static void allow_lds(K kernel, size_t bytes) {
  if (bytes > 65536) WHIR_HIP(hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
}

// --------------------------------------------------------------------------
// Twiddle tables.
// --------------------------------------------------------------------------

namespace {
std::mutex g_tw_mu;
std::map<std::pair<int, int>, TwTable> g_tw;                                   // (device, log n)
std::map<std::pair<int, int>, const uint64_t*> g_flat;                         // (device, log S)
std::map<std::tuple<int, int, int, int, int>, const uint64_t*> g_shape;  // (device, kind, log N, log C, log C1)

const uint64_t* upload_persistent(const std::vector<uint64_t>& v) {
  void* p = nullptr;
  WHIR_HIP(hipMalloc(&p, std::max<size_t>(v.size(), 1) * 8));
  WHIR_HIP(hipMemcpy(p, v.data(), v.size() * 8, hipMemcpyHostToDevice));
  return (const uint64_t*)p;
}
int current_device() {
  int dev = 0;
  WHIR_HIP(hipGetDevice(&dev));
  return dev;
}
}  // namespace

TwTable twiddles(int logn) {
  int dev = current_device();
  std::lock_guard<std::mutex> g(g_tw_mu);
  auto key = std::make_pair(dev, logn);
  auto it = g_tw.find(key);
  if (it != g_tw.end()) return it->second;
  WHIR_REQUIRE(!capture_active(), "ntt: twiddle table built during a stream capture");
  int h = (logn + 1) / 2;
  uint64_t w = gl::root_of_unity(1ull << logn);
  std::vector<uint64_t> lo(1ull << h), hi(1ull << (logn - h));
  uint64_t x = 1;
  for (auto& v : lo) { v = x; x = gl::mul(x, w); }
  uint64_t wh = gl::pow(w, 1ull << h);
  x = 1;
  for (auto& v : hi) { v = x; x = gl::mul(x, wh); }
  TwTable t;
  t.lo = upload_persistent(lo);
  t.hi = upload_persistent(hi);
  t.h = h;
  g_tw[key] = t;
  return t;
}

// w_S^j for every j < S (S <= 2^11): the in-LDS transform reads it from LDS.
static const uint64_t* flat_twiddles(int logS) {
  int dev = current_device();
  std::lock_guard<std::mutex> g(g_tw_mu);
  auto key = std::make_pair(dev, logS);
  auto it = g_flat.find(key);
  if (it != g_flat.end()) return it->second;
  WHIR_REQUIRE(!capture_active(), "ntt: twiddle table built during a stream capture");
  std::vector<uint64_t> v((size_t)1 << logS);
  uint64_t w = gl::root_of_unity(1ull << logS);
  uint64_t x = 1;
  for (auto& e : v) { e = x; x = gl::mul(x, w); }
  const uint64_t* p = upload_persistent(v);
  g_flat[key] = p;
  return p;
}


// kind 0: pre[c*C + i] = w_N^(c i)                        (c < Q, i < C)
// kind 1: post[(c*C2 + t2)*C1 + i1] = w_N^(i1 (c + Q t2))  (pass-1 output order)
// kind 2: row[c*C2 + i2] = w_N^(c C1 i2)                   (Q*C2 entries)
__global__ void k_shape_table(uint64_t* out, TwTable twN, int kind, uint64_t N, uint64_t len, int logC, int logC1) {
  const uint64_t C = 1ull << logC;
  for (uint64_t x = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; x < len; x += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t e;
    if (kind == 0) {
      e = (x >> logC) * (x & (C - 1));
    } else if (kind == 2) {
      const int logC2 = logC - logC1;
      e = ((x >> logC2) * (x & ((1ull << logC2) - 1))) << logC1;
    } else {
      const int logC2 = logC - logC1;
      uint64_t i1 = x & ((1ull << logC1) - 1);
      uint64_t row = x >> logC1;  // c*C2 + t2
      uint64_t t2 = row & ((1ull << logC2) - 1), c = row >> logC2;
      e = i1 * (c + (N >> logC) * t2);
    }
    out[x] = tw_at(twN, e & (N - 1));
  }
}

static const uint64_t* shape_table(int kind, int logN, int logC, int logC1, hipStream_t st) {
  int dev = current_device();
  {
    std::lock_guard<std::mutex> g(g_tw_mu);
    auto it = g_shape.find({dev, kind, logN, logC, logC1});
    if (it != g_shape.end()) return it->second;
  }
  WHIR_REQUIRE(!capture_active(), "ntt: shape table built during a stream capture");
  TwTable twN = twiddles(logN);
  uint64_t N = 1ull << logN;
  uint64_t len = kind == 2 ? (N >> logC) << (logC - logC1) : N;
  void* p = nullptr;
  WHIR_HIP(hipMalloc(&p, len * 8));
  unsigned blocks = (unsigned)std::min<uint64_t>((len + 255) / 256, 16384);
  k_shape_table<<<blocks, 256, 0, st>>>((uint64_t*)p, twN, kind, N, len, logC, logC1);
  WHIR_HIP(hipGetLastError());
  WHIR_HIP(hipStreamSynchronize(st));
  std::lock_guard<std::mutex> g(g_tw_mu);
  auto res = g_shape.emplace(std::make_tuple(dev, kind, logN, logC, logC1), (const uint64_t*)p);
  if (!res.second) (void)hipFree(p);  // another thread built it first
  return res.first->second;
}

// --------------------------------------------------------------------------
// In-LDS DFT: `cols` columns of length S = 2^logS stored s[col*S + idx], natural
// order in, bit-reversed order out (X[k] at index bitrev(k)).  tw[j] = w_S^j, j < S.
//
// A radix-R step (R = 2, 4, 8) on blocks of size S' takes, for each base < S'/R,
// y_m = x[base + m S'/R], runs the R-point DIF in registers (which leaves
// z_{bitrev(m)} in slot m), multiplies slot m by w_S'^(base * bitrev(m)) and
// writes it back in place; that is exactly the state after log2(R) radix-2 DIF
// stages, so the steps compose to the plain bit-reversed DIF.
// --------------------------------------------------------------------------
// Fp3 LDS tiles are padded by one element per 32 (element e lives at e + e/32): the
// strided butterfly accesses, the transposed tile writes and the bit-reversed reads
// then spread over the banks instead of piling onto a few.
//
// With NTT_SWZ, Fp and Fp2 tiles are not padded but XOR-swizzled instead: within each
// aligned run of 16 (Fp, 8-byte elements) or 8 (Fp2, 16-byte elements) elements, the
// slot is XORed with higher index bits, so that any 16 (Fp) or 8 (Fp2) lanes that
// access elements a power-of-two stride apart (the transposed tile loads and stores,
// the 8-consecutive-element groups of a last radix-8 level) land in distinct banks.
#ifndef NTT_SWZ
#define NTT_SWZ 1
#endif
template <int D, bool PAD>
__device__ __forceinline__ uint32_t lidx(uint32_t e) {
  if constexpr (NTT_SWZ && D == 2) return e ^ (((e >> 3) ^ (e >> 6) ^ (e >> 9)) & 7u);
  if constexpr (NTT_SWZ && D == 1) return e ^ (((e >> 4) ^ (e >> 8)) & 15u);
  return PAD ? e + (e >> 5) : e;
}
template <int D, bool PAD>
__host__ __device__ __forceinline__ size_t padded(size_t n) {
  if constexpr (NTT_SWZ && D <= 2) return n;
  return PAD ? n + (n >> 5) : n;
}

template <int D>
__device__ __forceinline__ Ext<D> rot(const Ext<D>& a, unsigned k) {
  Ext<D> r;
#pragma unroll
  for (int d = 0; d < D; d++) r.c[d] = gl::mul_pow2(a.c[d], k);
  return r;
}

template <int D>
__device__ __forceinline__ void bfly(Ext<D>& a, Ext<D>& b) {
  Ext<D> s = gl::eadd<D>(a, b), t = gl::esub<D>(a, b);
  a = s;
  b = t;
}

// (a + b, b - a): a butterfly whose lower output is then multiplied by -2^k is the
// same as this one followed by 2^k (2^96 = -1), which saves the negation.
template <int D>
__device__ __forceinline__ void bfly_neg(Ext<D>& a, Ext<D>& b) {
  Ext<D> s = gl::eadd<D>(a, b), t = gl::esub<D>(b, a);
  a = s;
  b = t;
}

// 8-point DIF on y[0..7] (slot m ends with X[bitrev3(m)]).  w_8 = 2^120 = -2^24,
// w_8^2 = w_4 = 2^48, w_8^3 = 2^168 = -2^72.
template <int D>
__device__ __forceinline__ void dif8(Ext<D>* y) {
  bfly<D>(y[0], y[4]);
  bfly_neg<D>(y[1], y[5]);
  bfly<D>(y[2], y[6]);
  bfly_neg<D>(y[3], y[7]);
  y[5] = rot<D>(y[5], 24);
  y[6] = rot<D>(y[6], 48);
  y[7] = rot<D>(y[7], 72);
  bfly<D>(y[0], y[2]);
  bfly<D>(y[1], y[3]);
  bfly<D>(y[4], y[6]);
  bfly<D>(y[5], y[7]);
  y[3] = rot<D>(y[3], 48);
  y[7] = rot<D>(y[7], 48);
  bfly<D>(y[0], y[1]);
  bfly<D>(y[2], y[3]);
  bfly<D>(y[4], y[5]);
  bfly<D>(y[6], y[7]);
}

// 16-point DIF: one radix-2 level with w_16^m on the differences (w_16 = 2^156, the
// reference's root: powers 2^{0,156,120,84,48,12,168,132}, the negative ones folded
// into bfly_neg), then 8-point DIFs of both halves.  Slot s ends with X[bitrev4(s)]
// (the even outputs come from the first half, the odd ones from the second).
template <int D>
__device__ __forceinline__ void dif16(Ext<D>* y) {
  bfly<D>(y[0], y[8]);
  bfly_neg<D>(y[1], y[9]);
  bfly_neg<D>(y[2], y[10]);
  bfly<D>(y[3], y[11]);
  bfly<D>(y[4], y[12]);
  bfly<D>(y[5], y[13]);
  bfly_neg<D>(y[6], y[14]);
  bfly_neg<D>(y[7], y[15]);
  y[9] = rot<D>(y[9], 60);
  y[10] = rot<D>(y[10], 24);
  y[11] = rot<D>(y[11], 84);
  y[12] = rot<D>(y[12], 48);
  y[13] = rot<D>(y[13], 12);
  y[14] = rot<D>(y[14], 72);
  y[15] = rot<D>(y[15], 36);
  dif8<D>(y);
  dif8<D>(y + 8);
}

template <int D, int R>
__device__ __forceinline__ void dif_regs(Ext<D> (&y)[R]) {
  if constexpr (R == 2) {
    bfly<D>(y[0], y[1]);
  } else if constexpr (R == 4) {
    bfly<D>(y[0], y[2]);
    bfly<D>(y[1], y[3]);
    y[3] = rot<D>(y[3], 48);  // w_4 = 2^48
    bfly<D>(y[0], y[1]);
    bfly<D>(y[2], y[3]);
  } else if constexpr (R == 8) {
    dif8<D>(y);
  } else {
    dif16<D>(y);
  }
}

template <int R>
__device__ __forceinline__ int brev_small(int m) {
  if constexpr (R == 2) return m;
  if constexpr (R == 4) return ((m & 1) << 1) | (m >> 1);
  if constexpr (R == 8) return ((m & 1) << 2) | (m & 2) | (m >> 2);
  return ((m & 1) << 3) | ((m & 2) << 1) | ((m & 4) >> 1) | (m >> 3);
}

// One radix-R DIF level over every sub-block of size S' = 2^logSub (cols columns of
// S = 2^logS in LDS), outputs multiplied by w_S'^(base * bitrev(m)).  TW = false is
// the last level (S' = R), where every such factor is w^0 = 1.
template <int D, bool PAD, int R, int LOGR, bool TW>
__device__ __forceinline__ void dif_step(Ext<D>* s, const uint64_t* tw, int logS, int logSub, int cols) {
  const int S = 1 << logS;
  const int stride = 1 << (logSub - LOGR);                 // S'/R
  const int groups = cols << (logS - LOGR);                // cols * S / R
  const int tw_shift = logS - logSub;                      // w_S' = w_S^(S/S')
#pragma unroll 2
  for (int q = threadIdx.x; q < groups; q += blockDim.x) {
    const int base = q & (stride - 1);
    const int blk = q >> (logSub - LOGR);                  // col * (S/S') + block within column
    const uint32_t e0 = ((uint32_t)blk << logSub) + base;
    Ext<D> y[R];
#pragma unroll
    for (int m = 0; m < R; m++) y[m] = s[lidx<D, PAD>(e0 + m * stride)];
    dif_regs<D, R>(y);
#pragma unroll
    for (int m = 0; m < R; m++) {
      const int e = (base * brev_small<R>(m)) << tw_shift;
      s[lidx<D, PAD>(e0 + m * stride)] = (m == 0 || !TW) ? y[m] : gl::escale<D>(y[m], tw[e & (S - 1)]);
    }
  }
  (void)S;
  (void)tw_shift;
}

template <int D, bool PAD, int R, int LOGR>
__device__ __forceinline__ void dif_level(Ext<D>* s, const uint64_t* tw, int logS, int logSub, int cols) {
  if (logSub == LOGR)
    dif_step<D, PAD, R, LOGR, false>(s, tw, logS, logSub, cols);
  else
    dif_step<D, PAD, R, LOGR, true>(s, tw, logS, logSub, cols);
}

// R16: radix-16 steps first (4 stages per LDS round trip, 15/16 of a generic
// multiply per element per four stages), then radix 8/4/2 for the rest.
template <int D, bool PAD, bool R16>
__device__ __forceinline__ void lds_dft(Ext<D>* s, const uint64_t* tw, int logS, int cols) {
  int logSub = logS;
#if NTT_REM_FIRST
  if (!R16 && logSub > 3 && logSub % 3) {
    // The levels that do not fill a radix-8 step go first, with their twiddles (half or
    // three quarters of a multiply per element), so that the last level is a whole
    // radix-8 step, whose factors are all 1 (instead of a radix-8 step with 7/8 of a
    // multiply and a twiddle-free radix-2 or radix-4 step at the end).
    if (logSub % 3 == 1) {
      dif_step<D, PAD, 2, 1, true>(s, tw, logS, logSub, cols);
      logSub -= 1;
    } else {
      dif_step<D, PAD, 4, 2, true>(s, tw, logS, logSub, cols);
      logSub -= 2;
    }
    __syncthreads();
  }
#endif
  if (R16) {
    while (logSub >= 4) {
      dif_level<D, PAD, 16, 4>(s, tw, logS, logSub, cols);
      logSub -= 4;
      __syncthreads();
    }
  }
  while (logSub >= 3) {
    dif_level<D, PAD, 8, 3>(s, tw, logS, logSub, cols);
    logSub -= 3;
    __syncthreads();
  }
  if (logSub == 2) {
    dif_step<D, PAD, 4, 2, false>(s, tw, logS, logSub, cols);
    logSub -= 2;
    __syncthreads();
  } else if (logSub == 1) {
    dif_step<D, PAD, 2, 1, false>(s, tw, logS, logSub, cols);
    logSub -= 1;
    __syncthreads();
  }
}

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own L2.
// Decode the hardware block index so that consecutive logical blocks (the cosets
// of one column tile, neighbouring tiles sharing cache lines) run on one XCD.
__device__ __forceinline__ bool xcd_block(uint32_t nlogical, uint32_t& logical) {
  constexpr uint32_t NX = 8;
  const uint32_t per = (nlogical + NX - 1) / NX;
  logical = (blockIdx.x % NX) * per + blockIdx.x / NX;
  return logical < nlogical;
}

__device__ __forceinline__ int bitrev(int x, int bits) {
  return bits ? (int)(__brev((unsigned)x) >> (32 - bits)) : 0;
}

// Coefficient i of the masked message: message || mask || zeros.
template <int D>
__device__ __forceinline__ Ext<D> coeff_at(const uint64_t* msg, size_t L, const uint64_t* mask,
                                           size_t ml, size_t i) {
  Ext<D> r;
  if (i < L) {
#pragma unroll
    for (int d = 0; d < D; d++) r.c[d] = msg[i * D + d];
  } else if (i < L + ml) {
#pragma unroll
    for (int d = 0; d < D; d++) r.c[d] = mask[(i - L) * D + d];
  } else {
    r = gl::ext_zero<D>();
  }
  return r;
}

// Message j's coefficients and codeword (see RsArgs).
template <int D>
__device__ __forceinline__ const uint64_t* msg_ptr(const RsArgs& a, size_t j) {
  return a.msg_tab ? a.msg_tab[j / a.mpi] + (j % a.mpi) * a.msg_stride * D : a.msgs + j * a.msg_stride * D;
}
template <int D>
__device__ __forceinline__ uint64_t* out_ptr(const RsArgs& a, size_t j) {
  return a.out_tab ? a.out_tab[j / a.mpi] + (j % a.mpi) * a.N * D : a.out + j * a.N * D;
}

// --------------------------------------------------------------------------
// Single pass: C = 2^logC <= 2^11.  Each block transforms G sequences.
// pre[c*C + i] = w_N^(c i).
// --------------------------------------------------------------------------
template <int D, bool PAD, bool R16>
__global__ __launch_bounds__(NTT_THREADS) void k_ntt_single(RsArgs a, const uint64_t* pre, const uint64_t* twS,
                                                            int G) {
  extern __shared__ uint64_t smem[];
  const int logC = a.logC;
  const int C = 1 << logC;
  Ext<D>* s = (Ext<D>*)smem;
  uint64_t* tws = smem + padded<D, PAD>((size_t)G * C) * D;
  for (int k = threadIdx.x; k < C; k += blockDim.x) tws[k] = twS[k];
  const size_t nseq = (size_t)a.M * a.Q;
  const size_t seq0 = (size_t)blockIdx.x * G;
  for (int e = threadIdx.x; e < G * C; e += blockDim.x) {
    int g = e >> logC, i = e & (C - 1);
    size_t seq = seq0 + g;
    Ext<D> v = gl::ext_zero<D>();
    if (seq < nseq) {
      size_t j = seq / a.Q, c = seq % a.Q;
      v = coeff_at<D>(msg_ptr<D>(a, j), a.L, a.masks ? a.masks + j * a.mask_len * D : nullptr, a.mask_len, i);
      if (c) v = gl::escale<D>(v, pre[c * C + i]);
    }
    s[lidx<D, PAD>(e)] = v;
  }
  __syncthreads();
  lds_dft<D, PAD, R16>(s, tws, logC, G);
  for (int e = threadIdx.x; e < G * C; e += blockDim.x) {
    int g = e >> logC, t = e & (C - 1);
    size_t seq = seq0 + g;
    if (seq >= nseq) continue;
    size_t j = seq / a.Q, c = seq % a.Q;
    Ext<D> v = s[lidx<D, PAD>(((uint32_t)g << logC) + bitrev(t, logC))];
    uint64_t* o = out_ptr<D>(a, j) + (c * C + t) * D;
#pragma unroll
    for (int d = 0; d < D; d++) o[d] = v.c[d];
  }
}

// --------------------------------------------------------------------------
// Pass 1: columns.  Block = (b consecutive i1, coset c, a run of `jn` messages).
// Logical blocks run cosets fastest, then tiles, so blocks whose input rows share
// cache lines run together on one XCD.  The post factors of the block's tile are
// the same for every message: EPT > 0 keeps them in registers across the message
// loop (EPT = elements per thread), EPT = 0 re-reads them (L2 hits) per message.
// row[c*C2 + i2] = w_N^(c C1 i2); post as in k_shape_table.
// Writes Y (same shape as the codeword) at [j][c*C + t2*C1 + i1].
// --------------------------------------------------------------------------
template <int D, int EPT, bool PAD, bool R16>
__global__ __launch_bounds__(P1_THREADS) void k_ntt_pass1(RsArgs a, const uint64_t* row, const uint64_t* post,
                                                           const uint64_t* twS, uint64_t* Y, int logC1, int logb,
                                                           int jn) {
  extern __shared__ uint64_t smem[];
  const int logC2 = a.logC - logC1;
  const int C1 = 1 << logC1, C2 = 1 << logC2, b = 1 << logb;
  const size_t C = (size_t)1 << a.logC;
  const int tile = C2 << logb;
  Ext<D>* s = (Ext<D>*)smem;
  uint64_t* tws = smem + padded<D, PAD>((size_t)tile) * D;
  for (int k = threadIdx.x; k < C2; k += blockDim.x) tws[k] = twS[k];
  const uint32_t ntile = (uint32_t)(C1 >> logb);
  const uint32_t njg = (uint32_t)((a.M + jn - 1) / jn);
  uint32_t lb;
  if (!xcd_block(ntile * (uint32_t)a.Q * njg, lb)) return;
  const size_t c = lb % a.Q;
  const uint32_t rest = (uint32_t)(lb / a.Q);
  const int i1_0 = (int)(rest % ntile) << logb;
  const size_t j0 = (size_t)(rest / ntile) * jn;
  const size_t j1 = j0 + jn < a.M ? j0 + jn : a.M;
  const uint64_t* rowc = row + c * C2;
  const uint64_t* postc = post + c * C;
  uint64_t pf[EPT > 0 ? EPT : 1];
  if constexpr (EPT > 0) {
#pragma unroll
    for (int k = 0; k < EPT; k++) {
      int e = threadIdx.x + k * P1_THREADS;
      if (e < tile) pf[k] = postc[((size_t)(e >> logb) << logC1) + i1_0 + (e & (b - 1))];
    }
  }
  for (size_t j = j0; j < j1; j++) {
    const uint64_t* msg = msg_ptr<D>(a, j);
    const uint64_t* mask = a.masks ? a.masks + j * a.mask_len * D : nullptr;
    for (int e = threadIdx.x; e < tile; e += blockDim.x) {
      int col = e & (b - 1), i2 = e >> logb;
      Ext<D> v = coeff_at<D>(msg, a.L, mask, a.mask_len, (size_t)(i1_0 + col) + ((size_t)i2 << logC1));
      if (c) v = gl::escale<D>(v, rowc[i2]);
      s[lidx<D, PAD>(((uint32_t)col << logC2) + i2)] = v;
    }
    __syncthreads();
    lds_dft<D, PAD, R16>(s, tws, logC2, b);
    uint64_t* yj = Y + (j * a.N + c * C) * D;
    if constexpr (EPT > 0) {
#pragma unroll
      for (int k = 0; k < EPT; k++) {
        int e = threadIdx.x + k * P1_THREADS;
        if (e < tile) {
          int col = e & (b - 1), t2 = e >> logb, li = bitrev(t2, logC2);
          size_t o_idx = ((size_t)t2 << logC1) + i1_0 + col;
          Ext<D> v = gl::escale<D>(s[lidx<D, PAD>(((uint32_t)col << logC2) + li)], pf[k]);
          uint64_t* o = yj + o_idx * D;
#pragma unroll
          for (int d = 0; d < D; d++) o[d] = v.c[d];
        }
      }
    } else {
      for (int e = threadIdx.x; e < tile; e += blockDim.x) {
        int col = e & (b - 1), t2 = e >> logb, li = bitrev(t2, logC2);
        size_t o_idx = ((size_t)t2 << logC1) + i1_0 + col;
        Ext<D> v = gl::escale<D>(s[lidx<D, PAD>(((uint32_t)col << logC2) + li)], postc[o_idx]);
        uint64_t* o = yj + o_idx * D;
#pragma unroll
        for (int d = 0; d < D; d++) o[d] = v.c[d];
      }
    }
    __syncthreads();  // the next message overwrites the tile
  }
}

// Post factors held in registers per thread: 8 for Fp and Fp2 (tiles up to 2048
// elements), measured 7-8% faster on MI355X than reloading them per message; none
// for Fp3, which has no registers to spare.
template <int D>
constexpr int pass1_ept() { return D <= 2 ? 8 : 0; }

// --------------------------------------------------------------------------
// Pass 2: rows.  Block = (b consecutive t2, coset c, message j).
// --------------------------------------------------------------------------
template <int D, bool PAD, bool R16>
__global__ __launch_bounds__(NTT_THREADS) void k_ntt_pass2(RsArgs a, const uint64_t* twS, const uint64_t* Y,
                                                           int logC1, int logb) {
  extern __shared__ uint64_t smem[];
  const int logC2 = a.logC - logC1;
  const int C1 = 1 << logC1, C2 = 1 << logC2, b = 1 << logb;
  const size_t C = (size_t)1 << a.logC;
  Ext<D>* s = (Ext<D>*)smem;
  uint64_t* tws = smem + padded<D, PAD>((size_t)C1 << logb) * D;
  for (int k = threadIdx.x; k < C1; k += blockDim.x) tws[k] = twS[k];
  // logical block = (j * Q + c) * ntile + tile: neighbouring row tiles are adjacent
  const uint32_t ntile = (uint32_t)(C2 >> logb);
  uint32_t lb;
  if (!xcd_block(ntile * (uint32_t)a.Q * (uint32_t)a.M, lb)) return;
  const int t2_0 = (int)(lb % ntile) << logb;
  const size_t c = (lb / ntile) % a.Q, j = (lb / ntile) / a.Q;
  const uint64_t* yj = Y + (j * a.N + c * C) * D;
  const int tile = C1 << logb;
  for (int e = threadIdx.x; e < tile; e += blockDim.x) {
    int row = e >> logC1, i1 = e & (C1 - 1);
    const uint64_t* src = yj + ((size_t)(t2_0 + row) * C1 + i1) * D;
    Ext<D> v;
#pragma unroll
    for (int d = 0; d < D; d++) v.c[d] = src[d];
    s[lidx<D, PAD>(e)] = v;
  }
  __syncthreads();
  lds_dft<D, PAD, R16>(s, tws, logC1, b);
  uint64_t* oj = out_ptr<D>(a, j) + c * C * D;
  for (int e = threadIdx.x; e < tile; e += blockDim.x) {
    int row = e & (b - 1), t1 = e >> logb, li = bitrev(t1, logC1);
    Ext<D> v = s[lidx<D, PAD>(((uint32_t)row << logC1) + li)];
    uint64_t* o = oj + ((size_t)(t2_0 + row) + ((size_t)t1 << logC2)) * D;
#pragma unroll
    for (int d = 0; d < D; d++) o[d] = v.c[d];
  }
  (void)C2;
}

// Column-major [M][N] -> row-major [N][M] (host-facing layout of interleaved_encode).
template <int D>
__global__ void k_cols_to_rows(const uint64_t* __restrict__ cw, uint64_t* __restrict__ out, size_t N, size_t M) {
  __shared__ uint64_t tile[32][33 * D];
  size_t r0 = (size_t)blockIdx.x * 32, j0 = (size_t)blockIdx.y * 32;
  int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  for (int k = ty; k < 32; k += 8) {
    size_t j = j0 + k, r = r0 + tx;
    if (j < M && r < N)
      for (int d = 0; d < D; d++) tile[k][tx * D + d] = cw[(j * N + r) * D + d];
  }
  __syncthreads();
  for (int k = ty; k < 32; k += 8) {
    size_t r = r0 + k, j = j0 + tx;
    if (j < M && r < N)
      for (int d = 0; d < D; d++) out[(r * M + j) * D + d] = tile[tx][k * D + d];
  }
}

// --------------------------------------------------------------------------
// Host drivers.
// --------------------------------------------------------------------------
static int log2_exact(uint64_t n) {
  int k = 0;
  while ((1ull << k) < n) k++;
  return (1ull << k) == n ? k : -1;
}

// Hardware blocks for `nlogical` logical blocks decoded by xcd_block.
static unsigned xcd_grid(uint64_t nlogical) {
  WHIR_REQUIRE(nlogical < (1ull << 31), "interleaved_encode: too many blocks");
  uint64_t per = (nlogical + 7) / 8;
  return (unsigned)(per * 8);
}

static int floor_log2(int x) {
  int k = 0;
  while ((2 << k) <= x) k++;
  return k;
}


// Per-degree variants of the in-LDS transforms, one bit per degree D:
//   pad  tile padded one element per 32
//   r16  radix-16 register steps (else radix 8)
// The defaults are the fastest measured on MI355X; WHIR_NTT_VARIANT = "pad,r16"
// masks overrides them for tuning.
struct NttVariant {
  // Radix 16 is off by default: with 512-thread blocks a 4096-word tile has fewer
  // 16-point groups than threads (the autotuner measured radix 8 3% faster for Fp at
  // 2^26 outputs), and for Fp2/Fp3 it also costs occupancy in registers.
  int pad = 0xC, r16 = 0x0;
};
static const NttVariant& ntt_variant() {
  static NttVariant v = [] {
    NttVariant r;
    if (const char* e = getenv("WHIR_NTT_VARIANT")) sscanf(e, "%i,%i", &r.pad, &r.r16);
    return r;
  }();
  return v;
}
static int variant_bits(int D) { return (((ntt_variant().pad >> D) & 1) << 1) | ((ntt_variant().r16 >> D) & 1); }

// One encode's schedule: the four-step split and, per pass, the in-LDS variant
// (bit 1: padding, bit 0: radix 16) and the LDS tile size in u64 words (which sets
// how many columns / rows / sequences a block transforms).
struct NttSched {
  int logC1 = 0;
  int v1 = 0, v2 = 0;
  int tw1 = 4096, tw2 = 4096;
};

static NttSched default_sched(int D, int logC) {
  NttSched s;
  s.logC1 = logC <= MAX_PASS_LOG ? 0 : (logC + 1) / 2;
  s.v1 = s.v2 = variant_bits(D);
  // Fp pass 1 (post factors in registers) is faster padded (1.13 vs 1.22 ms for both
  // passes at 2^26 outputs on MI355X, as measured by the autotuner); pass 2 is not.
  if (D == 1 && !getenv("WHIR_NTT_VARIANT") && logC > MAX_PASS_LOG) s.v1 = 2;
  s.tw1 = s.tw2 = tile_words();
  return s;
}

template <int D, int EPT, bool PAD, bool R16>
static void pass1_t(RsArgs a, const uint64_t* post, uint64_t* scratch, int logN, int logC1, int tw, hipStream_t st) {
  const int logC = a.logC, logC2 = logC - logC1;
  const int C1 = 1 << logC1, C2 = 1 << logC2;
  int logb = std::min(logC1, floor_log2(std::max(1, tw / (C2 * D))));
  if (EPT > 0) logb = std::min(logb, std::max(0, floor_log2(EPT * P1_THREADS) - logC2));
  const int tile = C2 << logb;
  WHIR_REQUIRE(EPT == 0 || tile <= EPT * P1_THREADS, "ntt pass 1: tile larger than the register budget");
  const uint64_t* row = shape_table(2, logN, logC, logC1, st);
  // Messages per block: as many as keep >= 2048 blocks in flight.
  const uint64_t base_blocks = (uint64_t)(C1 >> logb) * a.Q;
  int jn = 1;
  while (jn * 2 <= (int)a.M && base_blocks * ((a.M + jn * 2 - 1) / (jn * 2)) >= 2048) jn *= 2;
  size_t smem = (padded<D, PAD>((size_t)tile) * D + C2) * 8;
  dim3 grid(xcd_grid(base_blocks * ((a.M + jn - 1) / jn)));
  allow_lds(k_ntt_pass1<D, EPT, PAD, R16>, smem);
  ProfScope prof(kernel_name("k_ntt_pass1", D, EPT), st, 8.0 * D * ((double)a.M * a.L + (double)a.M * a.N));
  k_ntt_pass1<D, EPT, PAD, R16><<<grid, P1_THREADS, smem, st>>>(a, row, post, flat_twiddles(logC2), scratch, logC1,
                                                              logb, jn);
  WHIR_HIP(hipGetLastError());
}

template <int D, int EPT>
static void pass1(RsArgs a, const uint64_t* post, uint64_t* scratch, int logN, int logC1, int v, int tw,
                  hipStream_t st) {
  switch (v) {
    case 0: pass1_t<D, EPT, false, false>(a, post, scratch, logN, logC1, tw, st); break;
    case 1: pass1_t<D, EPT, false, true>(a, post, scratch, logN, logC1, tw, st); break;
    case 2: pass1_t<D, EPT, true, false>(a, post, scratch, logN, logC1, tw, st); break;
    default: pass1_t<D, EPT, true, true>(a, post, scratch, logN, logC1, tw, st); break;
  }
}

template <int D, bool PAD, bool R16>
static void single_t(RsArgs a, const uint64_t* pre, int G, size_t nseq, hipStream_t st) {
  const int C = 1 << a.logC;
  size_t smem = (padded<D, PAD>((size_t)G * C) * D + C) * 8;
  unsigned blocks = (unsigned)((nseq + G - 1) / G);
  allow_lds(k_ntt_single<D, PAD, R16>, smem);
  ProfScope prof(kernel_name("k_ntt_single", D), st, 8.0 * D * ((double)a.M * a.L + (double)a.M * a.N));
  k_ntt_single<D, PAD, R16><<<blocks, NTT_THREADS, smem, st>>>(a, pre, flat_twiddles(a.logC), G);
  WHIR_HIP(hipGetLastError());
}

template <int D>
static void single(RsArgs a, int v, int tw, hipStream_t st) {
  const int logN = log2_exact(a.N);
  const uint64_t* pre = shape_table(0, logN, a.logC, 0, st);
  const int C = 1 << a.logC;
  const int G = 1 << std::max(0, floor_log2(std::max(1, tw / (C * D))));
  const size_t nseq = (size_t)a.M * a.Q;
  switch (v) {
    case 0: single_t<D, false, false>(a, pre, G, nseq, st); break;
    case 1: single_t<D, false, true>(a, pre, G, nseq, st); break;
    case 2: single_t<D, true, false>(a, pre, G, nseq, st); break;
    default: single_t<D, true, true>(a, pre, G, nseq, st); break;
  }
}

template <int D, bool PAD, bool R16>
static void pass2_t(RsArgs a, uint64_t* scratch, int logC1, int tw, hipStream_t st) {
  const int logC2 = a.logC - logC1;
  const int C1 = 1 << logC1, C2 = 1 << logC2;
  int logb = std::min(logC2, floor_log2(std::max(1, tw / (C1 * D))));
  size_t smem = (padded<D, PAD>((size_t)C1 << logb) * D + C1) * 8;
  dim3 grid(xcd_grid((uint64_t)(C2 >> logb) * a.Q * a.M));
  allow_lds(k_ntt_pass2<D, PAD, R16>, smem);
  ProfScope prof(kernel_name("k_ntt_pass2", D), st, 16.0 * D * (double)a.M * a.N);
  k_ntt_pass2<D, PAD, R16><<<grid, pass2_threads<D>(), smem, st>>>(a, flat_twiddles(logC1), scratch, logC1, logb);
  WHIR_HIP(hipGetLastError());
}

template <int D>
static void pass2(RsArgs a, uint64_t* scratch, int logC1, int v, int tw, hipStream_t st) {
  switch (v) {
    case 0: pass2_t<D, false, false>(a, scratch, logC1, tw, st); break;
    case 1: pass2_t<D, false, true>(a, scratch, logC1, tw, st); break;
    case 2: pass2_t<D, true, false>(a, scratch, logC1, tw, st); break;
    default: pass2_t<D, true, true>(a, scratch, logC1, tw, st); break;
  }
}

// ------------------------------------------------------------------ autotuner
// With tuning on (whir_ntt_autotune(1) or WHIR_NTT_AUTOTUNE=1), the first encode of
// each shape (degree, log N, log C, messages) times every candidate schedule on the
// caller's stream, with the caller's data, and keeps the fastest:
//   the split log C1 in {ceil(logC/2), floor(logC/2)} and one level either side of
//   them (two-pass shapes),
//   per pass, the variant (padding x radix 16) and the tile size (4096 or 8192
//   words, where the device's LDS allows it),
// each pass tuned on its own (pass 1, then pass 2 on pass 1's output), each
// candidate timed as the fastest of three runs after a warm one.  A candidate
// displaces the built-in choice only when it is 3% faster.  Every candidate computes the same transform, so the tuned
// encode's output is the untuned one; only its time differs.  Shapes tuned are
// kept for the life of the process (whir_ntt_schedules reports them).
struct SchedKey {
  int D, logN, logC;
  uint64_t M;
  bool operator<(const SchedKey& o) const {
    return std::tie(D, logN, logC, M) < std::tie(o.D, o.logN, o.logC, o.M);
  }
};
struct SchedEntry {
  NttSched s;
  double ms1 = 0, ms2 = 0, ms_default = 0;
  int candidates = 0;
};
static std::mutex g_sched_mu;
static std::map<SchedKey, SchedEntry> g_sched;
static int g_autotune = -1;  // -1: not read yet

static bool autotune_on() {
  std::lock_guard<std::mutex> g(g_sched_mu);
  if (g_autotune < 0) {
    const char* e = getenv("WHIR_NTT_AUTOTUNE");
    g_autotune = (e && atoi(e) > 0) ? 1 : 0;
  }
  return g_autotune == 1;
}

static std::vector<int> tile_candidates() {
  std::vector<int> out{4096};
  int dev = 0, lds = 0;
  if (hipGetDevice(&dev) == hipSuccess &&
      hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess &&
      (size_t)(8192 + 8192 / 32 + 2048) * 8 <= (size_t)lds)
    out.push_back(8192);
  return out;
}

// One warm run, then the fastest of three timed runs, of fn on stream st.
template <class F>
static double time_on(hipStream_t st, F fn) {
  hipEvent_t e0, e1;
  WHIR_HIP(hipEventCreate(&e0));
  WHIR_HIP(hipEventCreate(&e1));
  fn();
  float best = 1e30f;
  for (int r = 0; r < 3; r++) {
    WHIR_HIP(hipEventRecord(e0, st));
    fn();
    WHIR_HIP(hipEventRecord(e1, st));
    WHIR_HIP(hipEventSynchronize(e1));
    float ms = 0;
    WHIR_HIP(hipEventElapsedTime(&ms, e0, e1));
    best = std::min(best, ms);
  }
  WHIR_HIP(hipEventDestroy(e0));
  WHIR_HIP(hipEventDestroy(e1));
  return best;
}

// A candidate replaces the built-in choice only when it is at least this much
// faster, so timing noise does not move a schedule away from the measured default.
static constexpr double kTuneMargin = 0.97;

template <int D>
static SchedEntry tune(RsArgs a, uint64_t* scratch, hipStream_t st) {
  const int logN = log2_exact(a.N), logC = a.logC;
  const NttSched def = default_sched(D, logC);
  const std::vector<int> tws = tile_candidates();
  SchedEntry best;
  best.s = def;
  if (logC <= MAX_PASS_LOG) {
    best.ms_default = time_on(st, [&] { single<D>(a, def.v1, def.tw1, st); });
    best.ms1 = best.ms_default;
    for (int v = 0; v < 4; v++)
      for (int tw : tws) {
        if (v == def.v1 && tw == def.tw1) continue;
        double ms = time_on(st, [&] { single<D>(a, v, tw, st); });
        best.candidates++;
        if (ms < kTuneMargin * best.ms1) {
          best.ms1 = ms;
          best.s.v1 = v;
          best.s.tw1 = tw;
        }
      }
    single<D>(a, best.s.v1, best.s.tw1, st);  // leave the tuned schedule's output behind
    return best;
  }
  constexpr int EPT = pass1_ept<D>();
  double best_total = 1e30;
  // the balanced split(s) first, then one level either side of them
  std::vector<int> splits{def.logC1};
  if (logC % 2) splits.push_back(logC / 2);
  for (int d : {def.logC1 + 1, logC / 2 - 1})
    if (std::find(splits.begin(), splits.end(), d) == splits.end()) splits.push_back(d);
  for (int logC1 : splits) {
    const int logC2 = logC - logC1;
    if (logC1 > MAX_PASS_LOG || logC2 > MAX_PASS_LOG) continue;
    if (EPT > 0 && (1 << logC2) > EPT * P1_THREADS) continue;  // pass-1 tile must fit the registers
    const uint64_t* post = shape_table(1, logN, logC, logC1, st);
    // Each pass starts from the built-in choice and moves only for a clear win.
    NttSched s = def;
    s.logC1 = logC1;
    double ms1 = time_on(st, [&] { pass1<D, EPT>(a, post, scratch, logN, logC1, def.v1, def.tw1, st); });
    for (int v = 0; v < 4; v++)
      for (int tw : tws) {
        if (v == def.v1 && tw == def.tw1) continue;
        double ms = time_on(st, [&] { pass1<D, EPT>(a, post, scratch, logN, logC1, v, tw, st); });
        best.candidates++;
        if (ms < kTuneMargin * ms1) {
          ms1 = ms;
          s.v1 = v;
          s.tw1 = tw;
        }
      }
    pass1<D, EPT>(a, post, scratch, logN, logC1, s.v1, s.tw1, st);
    double ms2 = time_on(st, [&] { pass2<D>(a, scratch, logC1, def.v2, def.tw2, st); });
    for (int v = 0; v < 4; v++)
      for (int tw : tws) {
        if (v == def.v2 && tw == def.tw2) continue;
        double ms = time_on(st, [&] { pass2<D>(a, scratch, logC1, v, tw, st); });
        best.candidates++;
        if (ms < kTuneMargin * ms2) {
          ms2 = ms;
          s.v2 = v;
          s.tw2 = tw;
        }
      }
    if (logC1 == def.logC1)
      best.ms_default = time_on(st, [&] {
        pass1<D, EPT>(a, post, scratch, logN, logC1, def.v1, def.tw1, st);
        pass2<D>(a, scratch, logC1, def.v2, def.tw2, st);
      });
    // The other split must beat the default split's best clearly too.
    if (best_total > 1e29 || ms1 + ms2 < kTuneMargin * best_total) {
      best_total = ms1 + ms2;
      best.s = s;
      best.ms1 = ms1;
      best.ms2 = ms2;
    }
  }
  // Leave the tuned schedule's output behind.
  const uint64_t* post = shape_table(1, logN, logC, best.s.logC1, st);
  pass1<D, EPT>(a, post, scratch, logN, best.s.logC1, best.s.v1, best.s.tw1, st);
  pass2<D>(a, scratch, best.s.logC1, best.s.v2, best.s.tw2, st);
  return best;
}

template <int D>
static void run_sched(RsArgs a, uint64_t* scratch, const NttSched& s, hipStream_t st) {
  if (a.logC <= MAX_PASS_LOG) {
    single<D>(a, s.v1, s.tw1, st);
    return;
  }
  WHIR_REQUIRE(a.logC <= 2 * MAX_PASS_LOG, "interleaved_encode: coset size above 2^22 is not supported");
  const int logN = log2_exact(a.N);
  const uint64_t* post = shape_table(1, logN, a.logC, s.logC1, st);
  pass1<D, pass1_ept<D>()>(a, post, scratch, logN, s.logC1, s.v1, s.tw1, st);
  pass2<D>(a, scratch, s.logC1, s.v2, s.tw2, st);
}

// The schedule recorded for a shape, else the built-in one.
static NttSched sched_for(int D, uint64_t N, int logC, uint64_t M, bool* found = nullptr) {
  const SchedKey key{D, log2_exact(N), logC, M};
  std::lock_guard<std::mutex> g(g_sched_mu);
  auto it = g_sched.find(key);
  if (found) *found = it != g_sched.end();
  return it != g_sched.end() ? it->second.s : default_sched(D, logC);
}

// Messages of one coefficient and no mask: the polynomial is a constant, so every
// evaluation is that coefficient (WHIR's last commitment); a broadcast replaces the
// transform.  Thread e writes element e % N of codeword e / N.
template <int D>
__global__ void k_rs_const(RsArgs a, int logN) {
  const size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (e >= (a.M << logN)) return;
  const size_t j = e >> logN, r = e & (a.N - 1);  // N is a power of two
  const uint64_t* m = msg_ptr<D>(a, j);
  uint64_t* o = out_ptr<D>(a, j) + r * D;
  if constexpr (D == 2) {
    *(uint4*)o = make_uint4((uint32_t)m[0], (uint32_t)(m[0] >> 32), (uint32_t)m[1], (uint32_t)(m[1] >> 32));
  } else {
#pragma unroll
    for (int d = 0; d < D; d++) o[d] = m[d];
  }
}

template <int D>
static void launch(RsArgs a, uint64_t* scratch, hipStream_t st) {
  if (a.L == 1 && a.mask_len == 0) {
    ProfScope prof(kernel_name("k_rs_const", D), st, 8.0 * D * (double)a.M * (double)(a.N + 1));
    k_rs_const<D><<<blocks_for(a.M * a.N, 256), 256, 0, st>>>(a, log2_exact(a.N));
    WHIR_HIP(hipGetLastError());
    return;
  }
  WHIR_REQUIRE(a.logC <= 2 * MAX_PASS_LOG, "interleaved_encode: coset size above 2^22 is not supported");
  bool found = false;
  const NttSched s = sched_for(D, a.N, a.logC, a.M, &found);
  if (found || capture_active()) {  // no tuning inside a stream capture
    run_sched<D>(a, scratch, s, st);
    return;
  }
  if (!autotune_on()) {
    run_sched<D>(a, scratch, default_sched(D, a.logC), st);
    return;
  }
  SchedEntry e = tune<D>(a, scratch, st);  // also leaves the output behind
  std::lock_guard<std::mutex> g(g_sched_mu);
  g_sched.emplace(SchedKey{D, log2_exact(a.N), a.logC, a.M}, e);
}

void rs_set_capturing(bool on) { set_capture_active(on); }

void rs_prepare(int D, uint64_t N, int logC, uint64_t M, hipStream_t st) {
  WHIR_REQUIRE(D >= 1 && D <= 3, "rs_prepare: extension degree must be 1, 2 or 3");
  const int logN = log2_exact(N);
  WHIR_REQUIRE(logN >= 0 && logC <= logN, "rs_prepare: bad shape");
  const NttSched s = sched_for(D, N, logC, M);
  if (logC <= MAX_PASS_LOG) {
    (void)shape_table(0, logN, logC, 0, st);
    (void)flat_twiddles(logC);
  } else {
    (void)shape_table(1, logN, logC, s.logC1, st);
    (void)shape_table(2, logN, logC, s.logC1, st);
    (void)flat_twiddles(logC - s.logC1);
    (void)flat_twiddles(s.logC1);
  }
}

size_t rs_scratch_words(int D, uint64_t M, uint64_t N, uint64_t C) {
  int logC = log2_exact(C);
  return logC > MAX_PASS_LOG ? (size_t)M * N * D : 0;
}

void rs_encode_cols(int D, RsArgs a, uint64_t* scratch, hipStream_t st) {
  WHIR_REQUIRE(a.M > 0, "rs_encode: no messages");
  WHIR_REQUIRE(log2_exact(a.N) >= 0, "rs_encode: the device path needs a power-of-two codeword length");
  WHIR_REQUIRE(((a.N >> a.logC) << a.logC) == a.N, "rs_encode: coset size must divide the codeword length");
  switch (D) {
    case 1: launch<1>(a, scratch, st); break;
    case 2: launch<2>(a, scratch, st); break;
    case 3: launch<3>(a, scratch, st); break;
    default: throw Error("rs_encode: extension degree must be 1, 2 or 3");
  }
}

void ntt_autotune(bool on) {
  std::lock_guard<std::mutex> g(g_sched_mu);
  g_autotune = on ? 1 : 0;
}

void ntt_set_schedule(int D, int logN, int logC, uint64_t M, int logC1, int v1, int tw1, int v2, int tw2) {
  WHIR_REQUIRE(D >= 1 && D <= 3, "ntt schedule: degree must be 1, 2 or 3");
  WHIR_REQUIRE(logC >= 0 && logC <= logN && logC <= 2 * MAX_PASS_LOG, "ntt schedule: bad coset size");
  WHIR_REQUIRE(M > 0, "ntt schedule: no messages");
  WHIR_REQUIRE(v1 >= 0 && v1 < 4 && v2 >= 0 && v2 < 4, "ntt schedule: variant is a 2-bit mask");
  const std::vector<int> tws = tile_candidates();
  auto tile_ok = [&](int tw) { return std::find(tws.begin(), tws.end(), tw) != tws.end(); };
  WHIR_REQUIRE(tile_ok(tw1) && tile_ok(tw2), "ntt schedule: tile size not available on this device");
  if (logC <= MAX_PASS_LOG) {
    WHIR_REQUIRE(logC1 == 0, "ntt schedule: single-pass shapes take log_c1 = 0");
  } else {
    WHIR_REQUIRE(logC1 >= logC - MAX_PASS_LOG && logC1 <= MAX_PASS_LOG, "ntt schedule: split out of range");
  }
  SchedEntry e;
  e.s = NttSched{logC1, v1, v2, tw1, tw2};
  std::lock_guard<std::mutex> g(g_sched_mu);
  g_sched[SchedKey{D, logN, logC, M}] = e;
}

void ntt_clear_schedules() {
  std::lock_guard<std::mutex> g(g_sched_mu);
  g_sched.clear();
}

std::string ntt_schedules_json() {
  std::lock_guard<std::mutex> g(g_sched_mu);
  std::string out = "[";
  char buf[512];
  bool first = true;
  for (const auto& kv : g_sched) {
    const SchedKey& k = kv.first;
    const SchedEntry& e = kv.second;
    snprintf(buf, sizeof buf,
             "%s{\"degree\":%d,\"log_n\":%d,\"log_c\":%d,\"messages\":%llu,\"log_c1\":%d,"
             "\"pass1\":{\"pad\":%d,\"radix16\":%d,\"tile_words\":%d},"
             "\"pass2\":{\"pad\":%d,\"radix16\":%d,\"tile_words\":%d},"
             "\"tuned_ms\":%.4f,\"default_ms\":%.4f,\"candidates\":%d}",
             first ? "" : ",", k.D, k.logN, k.logC, (unsigned long long)k.M, e.s.logC1, (e.s.v1 >> 1) & 1, e.s.v1 & 1,
             e.s.tw1, (e.s.v2 >> 1) & 1, e.s.v2 & 1, e.s.tw2, e.ms1 + e.ms2, e.ms_default, e.candidates);
    out += buf;
    first = false;
  }
  return out + "]";
}

void cols_to_rows(int D, const uint64_t* cw, uint64_t* out, size_t N, size_t M, hipStream_t st) {
  dim3 grid((unsigned)((N + 31) / 32), (unsigned)((M + 31) / 32));
  switch (D) {
    case 1: k_cols_to_rows<1><<<grid, 256, 0, st>>>(cw, out, N, M); break;
    case 2: k_cols_to_rows<2><<<grid, 256, 0, st>>>(cw, out, N, M); break;
    case 3: k_cols_to_rows<3><<<grid, 256, 0, st>>>(cw, out, N, M); break;
  }
  WHIR_HIP(hipGetLastError());
}

}  // namespace whir
