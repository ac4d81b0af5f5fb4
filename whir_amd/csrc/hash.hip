// Leaf hashing, Merkle layers and proof-of-work grinding on CDNA4.
//
// hash_many        HashEngine::hash_many (src/hash/mod.rs:79-85), one message per lane.
// leaves (cols)    matrix_commit::hash_rows (src/protocols/matrix_commit.rs:301-368) over the
//                  column-major codeword: a block stages RB whole rows in LDS with one
//                  coalesced read per column, then each lane hashes one row.  Field
//                  elements are canonical, so the ArkFieldEncoder bytes (canonical
//                  little-endian u64 per base coefficient, matrix_commit.rs:130-156) are
//                  the words as stored.
// merkle           merkle_tree::Config::commit (src/protocols/merkle_tree.rs:82-127): levels
//                  of 2^19 nodes and more lose three levels per launch, each lane reducing
//                  eight nodes in registers; above them each block reduces a 2^k-node
//                  slice of one level through k levels in LDS.  Every node made is
//                  written to its place in the nodes array.
// pow              proof_of_work::Config::prove (src/protocols/proof_of_work.rs:68-166):
//                  persistent search over nonce chunks handed out in increasing order; the
//                  result is the globally minimal nonce.
#include "common.h"
#include "hash.h"

#include <type_traits>
#include "hash_dev.h"
#include "keccak_host.h"

#include <algorithm>
#include <cstring>

namespace whir {

bool engine_supports_size(int engine, size_t size) {
  switch (engine) {
    case ENG_COPY: return size <= 32;
    case ENG_SHA2:
    case ENG_KECCAK:
    case ENG_SHA3:
    case ENG_BLAKE3: return true;
  }
  return false;
}

__device__ __forceinline__ void store_digest(uint8_t* out, const uint32_t h[8]) {
  uint4* o = (uint4*)out;
  o[0] = make_uint4(h[0], h[1], h[2], h[3]);
  o[1] = make_uint4(h[4], h[5], h[6], h[7]);
}

// Batched commits: one launch covers `count` codewords of the same shape, `per`
// blocks each; block b works on instance b / per (codeword cw[b / per], Merkle nodes
// nodes[b / per], root to roots + 32 (b / per)) as block b % per of a single launch.
// A zero table (cw == nullptr) is the single-codeword launch.
struct BatchTab {
  const uint64_t* const* cw = nullptr;
  uint8_t* const* nodes = nullptr;
  uint32_t per = 1;
  uint8_t* roots = nullptr;
};

// Calls f(std::integral_constant<int, E>{}) for the engine E given at run time, so
// each kernel is instantiated per engine (see hash_msg_t).
template <class F>
static void engine_switch(int engine, F&& f) {
  switch (engine) {
    case ENG_COPY: f(std::integral_constant<int, ENG_COPY>{}); break;
    case ENG_SHA2: f(std::integral_constant<int, ENG_SHA2>{}); break;
    case ENG_KECCAK: f(std::integral_constant<int, ENG_KECCAK>{}); break;
    case ENG_SHA3: f(std::integral_constant<int, ENG_SHA3>{}); break;
    case ENG_BLAKE3: f(std::integral_constant<int, ENG_BLAKE3>{}); break;
    default: throw Error("unknown hash engine");
  }
}

// HashEngine::preferred_batch_size (src/hash/mod.rs:67-77): the CPU engines give their
// SIMD width (blake3_engine.rs:79-81) or 1 (digest_engine.rs:118-120, copy_engine.rs:38-40),
// the messages one thread hashes at once.  hash_many here hashes one message per lane,
// so the batch that fills the device in one launch is every CU's four SIMDs holding four
// waves of 64 lanes: 1024 messages per CU (262144 on a 256-CU MI355X).
size_t engine_preferred_batch(int engine) {
  engine_switch(engine, [](auto) {});
  int dev = 0, cus = 0;
  WHIR_HIP(hipGetDevice(&dev));
  WHIR_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  return (size_t)std::max(cus, 1) * 1024;
}

template <int ENG>
__global__ void k_hash_many(size_t size, const uint8_t* __restrict__ in, size_t count, uint8_t* __restrict__ out) {
  size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= count) return;
  hd::MemReader rd{in + i * size, size};
  uint32_t h[8];
  hd::hash_msg_t<ENG>(rd, size, h);
  store_digest(out + 32 * i, h);
}

void hash_many_dev(int engine, size_t size, const uint8_t* in, size_t count, uint8_t* out, hipStream_t st) {
  WHIR_REQUIRE(engine_supports_size(engine, size), "hash_many: engine does not support this message size");
  if (!count) return;
  engine_switch(engine, [&](auto E) {
    k_hash_many<decltype(E)::value><<<blocks_for(count, 128), 128, 0, st>>>(size, in, count, out);
  });
  WHIR_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ leaves
static constexpr int LEAF_THREADS = 128;
static size_t LEAF_TILE_MAX = [] {
  const char* e = getenv("WHIR_LEAF_TILE_MAX");
  return (size_t)(e ? atol(e) : 20480);
}();

template <int D, int ENG>
__global__ __launch_bounds__(LEAF_THREADS) void k_leaves_cols(const uint64_t* __restrict__ cw, size_t N, size_t M,
                                                              uint8_t* __restrict__ leaves, int RB, BatchTab tab) {
  extern __shared__ uint64_t srow[];
  size_t blk = blockIdx.x;
  if (tab.cw) {
    cw = tab.cw[blk / tab.per];
    leaves = tab.nodes[blk / tab.per];
    blk %= tab.per;
  }
  const size_t W = M * D;          // words per row
  const size_t stride = W + 1;     // padded LDS row stride (odd: spreads banks)
  const size_t r0 = blk * RB;
  const size_t nrows = (N - r0) < (size_t)RB ? (N - r0) : (size_t)RB;
  for (size_t e = threadIdx.x; e < nrows * M; e += blockDim.x) {
    size_t r = e % nrows, j = e / nrows;
    const uint64_t* src = cw + (j * N + r0 + r) * D;
#pragma unroll
    for (int d = 0; d < D; d++) srow[r * stride + j * D + d] = src[d];
  }
  __syncthreads();
  for (size_t r = threadIdx.x; r < nrows; r += blockDim.x) {
    hd::WordReader rd{srow + r * stride, 8 * W};
    uint32_t h[8];
    hd::hash_msg_t<ENG>(rd, 8 * W, h);
    store_digest(leaves + 32 * (r0 + r), h);
  }
}

// Rows wider than the LDS tile: one lane per row reading straight from HBM.
template <int D, int ENG>
__global__ void k_leaves_cols_wide(const uint64_t* __restrict__ cw, size_t N, size_t M,
                                   uint8_t* __restrict__ leaves, BatchTab tab) {
  struct ColReader {
    const uint64_t* cw;
    size_t N, r, size;
    __device__ uint64_t w64(size_t off) const {
      if (off >= size) return 0;
      size_t word = off >> 3, j = word / D, d = word % D;
      return cw[(j * N + r) * D + d];
    }
    __device__ uint32_t w32(size_t off) const { return (uint32_t)(w64(off & ~(size_t)7) >> (8 * (off & 7))); }
  };
  size_t blk = blockIdx.x;
  if (tab.cw) {
    cw = tab.cw[blk / tab.per];
    leaves = tab.nodes[blk / tab.per];
    blk %= tab.per;
  }
  size_t r = blk * (size_t)blockDim.x + threadIdx.x;
  if (r >= N) return;
  ColReader rd{cw, N, r, 8 * M * D};
  uint32_t h[8];
  hd::hash_msg_t<ENG>(rd, 8 * M * D, h);
  store_digest(leaves + 32 * r, h);
}

template <int D>
static void leaves_cols(int engine, const uint64_t* cw, size_t N, size_t M, uint8_t* leaves, hipStream_t st,
                        BatchTab tab = {}, size_t count = 1) {
  size_t W = M * D;
  int RB = (int)std::min<size_t>(LEAF_THREADS, 8192 / (W + 1));
  // The LDS-staged kernel coalesces the row reads, but a tile above ~20 KiB per two
  // waves caps residency; wider rows read their columns straight from HBM instead.
  if ((size_t)RB * (W + 1) * 8 > LEAF_TILE_MAX) RB = 0;
  ProfScope prof(kernel_name(RB >= 16 ? "k_leaves_cols" : "k_leaves_cols_wide", D), st,
                 (8.0 * (double)W * N + 32.0 * N) * count);
  engine_switch(engine, [&](auto E) {
    constexpr int ENG = decltype(E)::value;
    if (RB >= 16) {
      size_t smem = (size_t)RB * (W + 1) * 8;
      tab.per = blocks_for(N, RB);
      k_leaves_cols<D, ENG><<<tab.per * (unsigned)count, LEAF_THREADS, smem, st>>>(cw, N, M, leaves, RB, tab);
    } else {
      tab.per = blocks_for(N, 128);
      k_leaves_cols_wide<D, ENG><<<tab.per * (unsigned)count, 128, 0, st>>>(cw, N, M, leaves, tab);
    }
  });
  WHIR_HIP(hipGetLastError());
}

void hash_rows_cols_dev(int engine, int D, const uint64_t* cw, size_t N, size_t M, uint8_t* leaves,
                        hipStream_t st) {
  WHIR_REQUIRE(engine_supports_size(engine, 8 * M * D), "leaf hash: engine does not support the row size");
  if (!N) return;
  switch (D) {
    case 1: leaves_cols<1>(engine, cw, N, M, leaves, st); break;
    case 2: leaves_cols<2>(engine, cw, N, M, leaves, st); break;
    case 3: leaves_cols<3>(engine, cw, N, M, leaves, st); break;
    default: throw Error("leaf hash: extension degree must be 1, 2 or 3");
  }
}

void hash_rows_dev(int engine, int D, const uint64_t* m, size_t rows, size_t M, uint8_t* leaves, hipStream_t st) {
  hash_many_dev(engine, 8 * M * D, (const uint8_t*)m, rows, leaves, st);
}

// ------------------------------------------------------------------ Merkle
size_t merkle_layers(size_t n) {
  size_t l = 0;
  while (((size_t)1 << l) < n) l++;
  return l;
}
size_t merkle_num_nodes(size_t n) { return ((size_t)2 << merkle_layers(n)) - 1; }

__global__ void k_gather_hashes(const uint4* __restrict__ nodes, const uint64_t* __restrict__ idx, size_t k,
                                uint4* __restrict__ out) {
  size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (e >= 2 * k) return;
  size_t q = e >> 1, half = e & 1;
  out[e] = nodes[2 * idx[q] + half];
}

void gather_hashes_dev(const uint8_t* nodes, const uint64_t* idx, size_t k, uint8_t* out, hipStream_t st) {
  if (!k) return;
  k_gather_hashes<<<blocks_for(2 * k, 256), 256, 0, st>>>((const uint4*)nodes, idx, k, (uint4*)out);
  WHIR_HIP(hipGetLastError());
}

static constexpr int MERKLE_THREADS = 256;
static constexpr int MERKLE_SPAN_LOG = 11;  // input nodes per block: 256 lanes x 8

// 64-byte node pair (two 8-word digests) as a hash input.
struct PairReader {
  const uint32_t* l;
  const uint32_t* r;
  __device__ __forceinline__ uint32_t w32(size_t off) const {
    const size_t i = off >> 2;
    return i < 8 ? l[i] : (i < 16 ? r[i - 8] : 0u);
  }
  __device__ __forceinline__ uint64_t w64(size_t off) const {
    return (uint64_t)w32(off) | ((uint64_t)w32(off + 4) << 32);
  }
};

// BLAKE3 of a 64-byte message: one block, flags CHUNK_START | CHUNK_END | ROOT.
__device__ __forceinline__ void b3_pair(const uint32_t l[8], const uint32_t r[8], uint32_t out[8]) {
  uint32_t m[16];
#pragma unroll
  for (int i = 0; i < 8; i++) {
    m[i] = l[i];
    m[8 + i] = r[i];
    out[i] = hd::B3_IV[i];
  }
  hd::b3_compress(out, m, 64, 0, hd::B3_CHUNK_START | hd::B3_CHUNK_END | hd::B3_ROOT);
}

template <int ENG>
__device__ __forceinline__ void hash_pair(const uint32_t l[8], const uint32_t r[8], uint32_t out[8]) {
  if constexpr (ENG == ENG_BLAKE3) {
    b3_pair(l, r, out);
  } else {
    PairReader rd{l, r};
    hd::hash_msg_t<ENG>(rd, 64, out);
  }
}

// Reduce G levels of 2^G consecutive nodes held in h (node index `first` within a
// level of width `width`), storing every node made at its place (the level of width
// x starts at node 2W - 2x); the subtree root ends in h[0].
template <int G, int ENG, int NH>
__device__ __forceinline__ void reduce_levels(uint32_t (&h)[NH][8], uint8_t* nodes, size_t W, size_t width,
                                              size_t first) {
  static_assert((1 << G) <= NH, "register group too small");
#pragma unroll
  for (int lv = 0; lv < G; lv++) {
    width >>= 1;
    first >>= 1;
    const size_t base = 2 * W - 2 * width + first;
#pragma unroll
    for (int k = 0; k < (1 << (G - 1 - lv)); k++) {
      uint32_t o[8];
      hash_pair<ENG>(h[2 * k], h[2 * k + 1], o);
      store_digest(nodes + 32 * (base + k), o);
#pragma unroll
      for (int i = 0; i < 8; i++) h[k][i] = o[i];
    }
  }
}

// merkle_tree::Config::commit (src/protocols/merkle_tree.rs:82-127) from a level of
// width w: block b reduces input nodes [b 2^L, (b+1) 2^L) through L levels.  The
// first phase has lane i reduce 2^G1 consecutive nodes (G1 = 3: seven hashes per
// lane, every lane busy) to one in registers; after it, one level per phase (lane i
// hashes the pair 2i, 2i + 1 from LDS), which keeps the chain of dependent hashes
// short once few nodes are left.
// Largest first-phase group (2^G nodes per lane).  Eight for every engine: for the
// Keccak engines four-node groups (122 instead of 180 registers) measured 4% slower
// on MI355X and pairs 40% slower, the longer chains of dependent hashes costing more
// than the occupancy gained.
template <int ENG>
constexpr int merkle_gmax() { return 3; }
static int merkle_gmax_rt(int) { return 3; }

template <int ENG>
__global__ __launch_bounds__(MERKLE_THREADS) void k_merkle_sub(uint8_t* __restrict__ nodes, size_t W, size_t w,
                                                               int L, int G1, uint32_t* root_out, BatchTab tab,
                                                               MailboxSig sig) {
  constexpr int NH = 1 << merkle_gmax<ENG>();
  __shared__ uint32_t buf[2][MERKLE_THREADS][8];
  size_t blk = blockIdx.x;
  if (tab.nodes) {
    nodes = tab.nodes[blk / tab.per];
    if (root_out) root_out = (uint32_t*)(tab.roots + 32 * (blk / tab.per));
    blk %= tab.per;
  }
  size_t width = w;                                 // width of the current level
  size_t first = blk << L;                          // block's first node within the current level
  size_t cnt = (size_t)1 << L;                      // block's nodes in the current level
  int left = L, cur = 0;
  bool from_global = true;
  while (left > 0) {
    const int G = from_global ? (left < G1 ? left : G1) : 1;
    const size_t lanes = cnt >> G;
    if (threadIdx.x < lanes) {
      uint32_t h[NH][8];
      const size_t f = first + ((size_t)threadIdx.x << G);
#pragma unroll
      for (int k = 0; k < NH; k++) {
        if (k < (1 << G)) {
          if (from_global) {
            const uint4* src = (const uint4*)(nodes + 32 * (2 * W - 2 * width + f + k));
            const uint4 a = src[0], b = src[1];
            h[k][0] = a.x; h[k][1] = a.y; h[k][2] = a.z; h[k][3] = a.w;
            h[k][4] = b.x; h[k][5] = b.y; h[k][6] = b.z; h[k][7] = b.w;
          } else {
#pragma unroll
            for (int i = 0; i < 8; i++) h[k][i] = buf[cur][(threadIdx.x << G) + k][i];
          }
        }
      }
      if constexpr (NH >= 8) {
        if (G == 3) reduce_levels<3, ENG, NH>(h, nodes, W, width, f);
        else if (G == 2) reduce_levels<2, ENG, NH>(h, nodes, W, width, f);
        else reduce_levels<1, ENG, NH>(h, nodes, W, width, f);
      } else if constexpr (NH >= 4) {
        if (G == 2) reduce_levels<2, ENG, NH>(h, nodes, W, width, f);
        else reduce_levels<1, ENG, NH>(h, nodes, W, width, f);
      } else {
        reduce_levels<1, ENG, NH>(h, nodes, W, width, f);
      }
#pragma unroll
      for (int i = 0; i < 8; i++) buf[cur ^ 1][threadIdx.x][i] = h[0][i];
    }
    __syncthreads();
    cur ^= 1;
    from_global = false;
    width >>= G;
    first >>= G;
    cnt = lanes;
    left -= G;
  }
  // The launch that makes the root also hands it to the caller (e.g. mapped host memory).
  if (root_out && threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < 8; i++) root_out[i] = buf[cur][0][i];
    mailbox_raise(sig);
  }
}

// ------------------------------------------------------------------ BLAKE3, quad-split
// Below 2^16 nodes a Merkle level holds less than one wave per SIMD, so each level
// costs the latency of one wave running a whole compression.  Here the four lanes of a
// quad share one compression: lane q holds column q of the state (v[q], v[4+q],
// v[8+q], v[12+q]) and runs that column's G; for the diagonal step rows 1-3 rotate by
// 1, 2 and 3 lanes inside the quad (DPP quad permutes), lane q runs diagonal q's G, and
// the rows rotate back.  Lane q then owns chaining words q and 4+q.  The message words
// each lane needs in each round are fixed by the schedule, so every lane reads its own
// 28 words from LDS at offsets computed once per kernel.
struct B3QSched {
  uint16_t p[28];  // per (round, word slot): the message word of lane q in bits 4q..4q+3
};
constexpr B3QSched b3q_sched() {
  constexpr int perm[16] = {2, 6, 3, 10, 7, 0, 4, 13, 1, 11, 12, 5, 9, 14, 15, 8};
  int s[16] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15};  // round word j = m[s[j]]
  B3QSched t{};
  for (int r = 0; r < 7; r++) {
    for (int q = 0; q < 4; q++) {
      t.p[4 * r + 0] |= (uint16_t)(s[2 * q] << (4 * q));
      t.p[4 * r + 1] |= (uint16_t)(s[2 * q + 1] << (4 * q));
      t.p[4 * r + 2] |= (uint16_t)(s[8 + 2 * q] << (4 * q));
      t.p[4 * r + 3] |= (uint16_t)(s[9 + 2 * q] << (4 * q));
    }
    int n[16] = {};
    for (int i = 0; i < 16; i++) n[i] = s[perm[i]];
    for (int i = 0; i < 16; i++) s[i] = n[i];
  }
  return t;
}

// Lane q of the quad receives v from lane ((CTRL >> 2q) & 3) of the same quad.
template <int CTRL>
__device__ __forceinline__ uint32_t quad_perm(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, CTRL, 0xF, 0xF, false);
}

// BLAKE3 of the 64-byte message msg[0..15] (LDS, the same for the quad's four lanes),
// flags CHUNK_START | CHUNK_END | ROOT.  All four lanes of the quad must be active.
__device__ __forceinline__ void b3q_pair(const uint32_t* msg, const uint32_t (&mo)[28], uint32_t iva, uint32_t ivb,
                                         uint32_t dinit, uint32_t& lo, uint32_t& hi) {
  using hd::rotr32;
  uint32_t x[28];
#pragma unroll
  for (int j = 0; j < 28; j++) x[j] = msg[mo[j]];
  uint32_t a = iva, b = ivb, c = iva, d = dinit;
#pragma unroll
  for (int r = 0; r < 7; r++) {
    B3G(a, b, c, d, x[4 * r], x[4 * r + 1]);
    b = quad_perm<0x39>(b);  // lanes q+1, q+2, q+3
    c = quad_perm<0x4E>(c);
    d = quad_perm<0x93>(d);
    B3G(a, b, c, d, x[4 * r + 2], x[4 * r + 3]);
    b = quad_perm<0x93>(b);  // back: lanes q-1, q-2, q-3
    c = quad_perm<0x4E>(c);
    d = quad_perm<0x39>(d);
  }
  lo = a ^ c;
  hi = b ^ d;
}

// k_merkle_sub for BLAKE3 with a quad per node: quad i takes the place of lane i (the
// same phases, groups and node indices), four times the threads.
constexpr int B3Q_THREADS = 4 * MERKLE_THREADS;

__global__ __launch_bounds__(B3Q_THREADS) void k_merkle_sub_b3q(uint8_t* __restrict__ nodes, size_t W, size_t w,
                                                                 int L, int G1, uint32_t* root_out, BatchTab tab,
                                                                 MailboxSig sig) {
  __shared__ uint32_t buf[2][MERKLE_THREADS][8];
  __shared__ uint32_t grp[MERKLE_THREADS][8][8];
  size_t blk = blockIdx.x;
  if (tab.nodes) {
    nodes = tab.nodes[blk / tab.per];
    if (root_out) root_out = (uint32_t*)(tab.roots + 32 * (blk / tab.per));
    blk %= tab.per;
  }
  uint32_t* nodes32 = (uint32_t*)nodes;
  const uint32_t q = threadIdx.x & 3, qi = threadIdx.x >> 2;
  constexpr B3QSched S = b3q_sched();
  uint32_t mo[28];
#pragma unroll
  for (int j = 0; j < 28; j++) mo[j] = (S.p[j] >> (4 * q)) & 15u;
  const uint32_t iva = hd::B3_IV[q], ivb = hd::B3_IV[4 + q];
  const uint32_t dinit = q == 2 ? 64u : (q == 3 ? (uint32_t)(hd::B3_CHUNK_START | hd::B3_CHUNK_END | hd::B3_ROOT) : 0u);
  size_t width = w;
  size_t first = blk << L;
  size_t cnt = (size_t)1 << L;
  int left = L, cur = 0;
  bool from_global = true;
  while (left > 0) {
    const int G = from_global ? (left < G1 ? left : G1) : 1;
    const size_t quads = cnt >> G;
    if (qi < quads) {
      const size_t f = first + ((size_t)qi << G);
      uint32_t lo = 0, hi = 0;
      if (from_global) {
        for (int k = (int)q; k < (1 << G); k += 4) {
          const uint4* src = (const uint4*)(nodes + 32 * (2 * W - 2 * width + f + k));
          uint4* dst = (uint4*)&grp[qi][k][0];
          dst[0] = src[0];
          dst[1] = src[1];
        }
        __builtin_amdgcn_wave_barrier();
        size_t wd = width, fl = f;
        for (int lv = 0; lv < G; lv++) {
          wd >>= 1;
          fl >>= 1;
          const size_t base = 2 * W - 2 * wd + fl;
          for (int k = 0; k < (1 << (G - 1 - lv)); k++) {
            b3q_pair(&grp[qi][2 * k][0], mo, iva, ivb, dinit, lo, hi);
            __builtin_amdgcn_wave_barrier();
            grp[qi][k][q] = lo;
            grp[qi][k][4 + q] = hi;
            nodes32[8 * (base + k) + q] = lo;
            nodes32[8 * (base + k) + 4 + q] = hi;
            __builtin_amdgcn_wave_barrier();
          }
        }
      } else {
        b3q_pair(&buf[cur][2 * qi][0], mo, iva, ivb, dinit, lo, hi);
        const size_t base = 2 * W - 2 * (width >> 1) + (f >> 1);
        nodes32[8 * base + q] = lo;
        nodes32[8 * base + 4 + q] = hi;
      }
      buf[cur ^ 1][qi][q] = lo;
      buf[cur ^ 1][qi][4 + q] = hi;
    }
    __syncthreads();
    cur ^= 1;
    from_global = false;
    width >>= G;
    first >>= G;
    cnt = quads;
    left -= G;
  }
  if (root_out && threadIdx.x == 0) {
#pragma unroll
    for (int i = 0; i < 8; i++) root_out[i] = buf[cur][0][i];
    mailbox_raise(sig);
  }
}

// Quad-split BLAKE3 levels on (WHIR_MERKLE_QUAD=0 keeps one lane per node, for comparison).
static bool b3q_enabled() {
  static const bool on = [] {
    const char* e = getenv("WHIR_MERKLE_QUAD");
    return !e || atoi(e) != 0;
  }();
  return on;
}
// A launch takes the quad kernel while its first phase has at most this many lanes
// (one wave per SIMD): there each level costs one compression's latency, which the
// quads cut; on more lanes the launch is bound by instruction issue, and the quads'
// extra instructions (DPP moves, per-lane message reads, four lanes of G) cost more
// than they save.  Measured on MI355X per launch, pairs per lane (one lane / quad):
// one block 16.0 / 10.1 us, 2^16 nodes (128 blocks) 19.2 / 12.9, 2^17 nodes 19.4 /
// 13.6, 2^18 nodes (2^17 lanes) 25.1 / 26.0; eight nodes per lane from 2^19 nodes
// (which k_merkle_wide takes first anyway) 38.4 us as quads against 32.2 us for
// k_merkle_wide and the quad launch above it.
static constexpr size_t B3Q_MAX_LANES = (size_t)1 << 16;

// Three Merkle levels above a level of width w (w % 2^11 == 0): lane g reduces nodes
// 8g..8g+7 to one in registers and stores the seven nodes it makes.  No lane idles
// and nothing waits at a barrier, which matters on levels wide enough to keep every
// SIMD busy: there the block-wide reduction's later phases (128, 64, ... lanes of a
// 256-lane block) hold registers that waiting blocks could use.
template <int ENG>
__global__ __launch_bounds__(MERKLE_THREADS) void k_merkle_wide(uint8_t* __restrict__ nodes, size_t W, size_t w,
                                                                BatchTab tab) {
  size_t blk = blockIdx.x;
  if (tab.nodes) {
    nodes = tab.nodes[blk / tab.per];
    blk %= tab.per;
  }
  const size_t f = (blk * MERKLE_THREADS + threadIdx.x) << 3;
  uint32_t h[8][8];
#pragma unroll
  for (int k = 0; k < 8; k++) {
    const uint4* src = (const uint4*)(nodes + 32 * (2 * W - 2 * w + f + k));
    const uint4 a = src[0], b = src[1];
    h[k][0] = a.x; h[k][1] = a.y; h[k][2] = a.z; h[k][3] = a.w;
    h[k][4] = b.x; h[k][5] = b.y; h[k][6] = b.z; h[k][7] = b.w;
  }
  reduce_levels<3, ENG, 8>(h, nodes, W, w, f);
}

// Levels at least this wide go through k_merkle_wide (2^16 lanes and more: a wave or
// more per SIMD).  Measured on MI355X, d = 26 SHA3 proofs: 20.6 -> 19.2 ms with 2^19
// or 2^18, 19.6 ms with 2^20; BLAKE3 trees, latency-bound at these widths, unchanged.
static constexpr size_t MERKLE_WIDE_MIN = (size_t)1 << 19;

// Merkle levels above a level of width w (a power of two) already in place.  The
// root also goes to root_out (32 bytes) when given.  With a batch table, the same
// levels of `count` trees (root i to tab.roots + 32 i).  With stop > 1 the launches end
// at a level of width <= stop (never below it unless w already is), whose width is
// returned; the caller makes the rest.
static size_t merkle_from(int engine, uint8_t* nodes, size_t W, size_t w, hipStream_t st, uint8_t* root_out,
                          BatchTab tab = {}, size_t count = 1, size_t stop = 1) {
  if (w <= stop) {
    if (stop > 1) return w;
    WHIR_REQUIRE(!tab.nodes, "merkle: a batch of one-leaf trees is not supported");
    if (root_out) WHIR_HIP(hipMemcpyAsync(root_out, nodes + 32 * (2 * W - 2), 32, hipMemcpyDefault, st));
    return w;
  }
  while (w >= MERKLE_WIDE_MIN) {
    const size_t blocks = w / ((size_t)MERKLE_THREADS << 3);
    ProfScope prof("k_merkle_wide", st, 32.0 * (double)(w + (w >> 1) + (w >> 2) + (w >> 3)) * count);
    tab.per = (uint32_t)blocks;
    engine_switch(engine, [&](auto E) {
      k_merkle_wide<decltype(E)::value><<<(unsigned)(blocks * count), MERKLE_THREADS, 0, st>>>(nodes, W, w, tab);
    });
    WHIR_HIP(hipGetLastError());
    w >>= 3;
  }
  while (w > stop) {
    // Wide levels: eight nodes per lane first (2^11 per block); otherwise pairs
    // (2^9 per block), which has the shortest chain of dependent hashes.
    const int G1 = w >= ((size_t)MERKLE_THREADS << MERKLE_SPAN_LOG) ? merkle_gmax_rt(engine) : 1;
    const int maxL = MERKLE_SPAN_LOG - 3 + G1;
    int L = 0;
    while (L < maxL && ((size_t)2 << L) <= w && (w >> (L + 1)) >= stop) L++;
    const size_t blocks = w >> L;
    const size_t lanes = ((size_t)1 << L) >> std::min(L, G1);  // nodes made by the first phase
    const bool quad = engine == ENG_BLAKE3 && b3q_enabled() && blocks * count * lanes <= B3Q_MAX_LANES;
    const unsigned threads = quad ? (unsigned)std::max<size_t>(64, std::min<size_t>(B3Q_THREADS, 4 * lanes))
                                  : (unsigned)std::max<size_t>(64, std::min<size_t>(MERKLE_THREADS, lanes));
    ProfScope prof(quad ? "k_merkle_sub_b3q" : "k_merkle_sub", st, 32.0 * (double)(2 * w - (w >> L)) * count);
    tab.per = (uint32_t)blocks;
    uint32_t* ro = blocks == 1 ? (tab.nodes ? (uint32_t*)tab.roots : (uint32_t*)root_out) : nullptr;
    // The launch that makes a single tree's root is the last of the commit.
    const MailboxSig sig = blocks == 1 && !tab.nodes && ro ? mailbox_take(ro) : MailboxSig{};
    if (quad) {
      k_merkle_sub_b3q<<<(unsigned)(blocks * count), threads, 0, st>>>(nodes, W, w, L, G1, ro, tab, sig);
    } else {
      engine_switch(engine, [&](auto E) {
        k_merkle_sub<decltype(E)::value><<<(unsigned)(blocks * count), threads, 0, st>>>(nodes, W, w, L, G1, ro,
                                                                                         tab, sig);
      });
    }
    WHIR_HIP(hipGetLastError());
    w >>= L;
  }
  return w;
}

void merkle_build_dev(int engine, uint8_t* nodes, size_t num_leaves, hipStream_t st, uint8_t* root_out) {
  WHIR_REQUIRE(engine_supports_size(engine, 64), "merkle: engine does not hash 64-byte nodes");
  size_t layers = merkle_layers(num_leaves);
  size_t W = (size_t)1 << layers;
  if (W > num_leaves) fill_zero(nodes + 32 * num_leaves, 32 * (W - num_leaves), st);
  merkle_from(engine, nodes, W, W, st, root_out);
}

// BLAKE3 leaves of a column-major codeword whose rows are single-chunk messages of
// a whole number of 64-byte blocks (W = M D words, W % 8 == 0, W <= 128), fused
// with the two Merkle levels above them: lane g hashes rows 4g..4g+3, their two
// parents and the grandparent, all in registers, reading each row's words straight
// from HBM (for a fixed word, consecutive lanes read consecutive rows).
static constexpr int TREE_THREADS = 256;

template <int D>
__device__ __forceinline__ void row_block(const uint64_t* __restrict__ cw, size_t N, size_t r, size_t b,
                                          uint32_t m[16]) {
  if constexpr (D == 2) {
    const uint4* c4 = (const uint4*)cw;
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const uint4 v = c4[(4 * b + q) * N + r];  // element (column 4b + q, row r)
      m[4 * q] = v.x; m[4 * q + 1] = v.y; m[4 * q + 2] = v.z; m[4 * q + 3] = v.w;
    }
  } else {
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const size_t w = 8 * b + q, j = w / D, d = w % D;
      const uint64_t v = cw[(j * N + r) * D + d];
      m[2 * q] = (uint32_t)v;
      m[2 * q + 1] = (uint32_t)(v >> 32);
    }
  }
}

template <int D>
__global__ __launch_bounds__(TREE_THREADS) void k_leaves_tree_b3(const uint64_t* __restrict__ cw, size_t N, size_t M,
                                                                 uint8_t* __restrict__ nodes, BatchTab tab) {
  size_t blk = blockIdx.x;
  if (tab.cw) {
    cw = tab.cw[blk / tab.per];
    nodes = tab.nodes[blk / tab.per];
    blk %= tab.per;
  }
  const size_t g = blk * (size_t)blockDim.x + threadIdx.x;
  if (4 * g >= N) return;
  const size_t nb = M * D / 8;
  uint32_t left[8], par[8];
#pragma unroll 1
  for (int k = 0; k < 4; k++) {
    const size_t r = 4 * g + k;
    uint32_t cv[8];
#pragma unroll
    for (int i = 0; i < 8; i++) cv[i] = hd::B3_IV[i];
#pragma unroll 1
    for (size_t b = 0; b < nb; b++) {
      uint32_t m[16];
      row_block<D>(cw, N, r, b, m);
      const uint32_t flags =
          (b == 0 ? hd::B3_CHUNK_START : 0u) | (b + 1 == nb ? (hd::B3_CHUNK_END | hd::B3_ROOT) : 0u);
      hd::b3_compress(cv, m, 64, 0, flags);
    }
    store_digest(nodes + 32 * r, cv);
    if ((k & 1) == 0) {
#pragma unroll
      for (int i = 0; i < 8; i++) left[i] = cv[i];
    } else {
      uint32_t p[8];
      b3_pair(left, cv, p);
      store_digest(nodes + 32 * (N + (r >> 1)), p);
      if (k == 1) {
#pragma unroll
        for (int i = 0; i < 8; i++) par[i] = p[i];
      } else {
        uint32_t gp[8];
        b3_pair(par, p, gp);
        store_digest(nodes + 32 * (N + N / 2 + g), gp);
      }
    }
  }
}

// Wave-cooperative variant for N >= 256: a wave takes 256 consecutive rows, lane l
// hashing rows R0 + 64k + l (k < 4), so every row load is coalesced across the wave.
// The two levels above are then shared out so every lane keeps working: lane l
// makes the level-1 nodes of pair l/2 in groups l%2 and l%2 + 2, and the level-2
// node of group l%4, quad m = l/4; the children come from lanes l^1 and l^2
// through DPP quad permutations (no LDS).
__device__ __forceinline__ void dpp_xor1(const uint32_t in[8], uint32_t out[8]) {
#pragma unroll
  for (int i = 0; i < 8; i++) out[i] = (uint32_t)__builtin_amdgcn_mov_dpp((int)in[i], 0xB1, 0xF, 0xF, false);
}
__device__ __forceinline__ void dpp_xor2(const uint32_t in[8], uint32_t out[8]) {
#pragma unroll
  for (int i = 0; i < 8; i++) out[i] = (uint32_t)__builtin_amdgcn_mov_dpp((int)in[i], 0x4E, 0xF, 0xF, false);
}
__device__ __forceinline__ void sel8(bool c, const uint32_t a[8], const uint32_t b[8], uint32_t out[8]) {
#pragma unroll
  for (int i = 0; i < 8; i++) out[i] = c ? a[i] : b[i];
}

template <int D>
__global__ __launch_bounds__(TREE_THREADS) void k_leaves_tree_b3w(const uint64_t* __restrict__ cw, size_t N, size_t M,
                                                                  uint8_t* __restrict__ nodes, BatchTab tab) {
  size_t blk = blockIdx.x;
  if (tab.cw) {
    cw = tab.cw[blk / tab.per];
    nodes = tab.nodes[blk / tab.per];
    blk %= tab.per;
  }
  const size_t wave = (blk * (size_t)blockDim.x + threadIdx.x) >> 6;
  const size_t R0 = wave * 256;
  if (R0 >= N) return;  // whole waves only (N is a multiple of 256)
  const int l = threadIdx.x & 63;
  const size_t nb = M * D / 8;
  uint32_t leaf[4][8];
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const size_t r = R0 + 64 * k + l;
    uint32_t cv[8];
#pragma unroll
    for (int i = 0; i < 8; i++) cv[i] = hd::B3_IV[i];
#pragma unroll 1
    for (size_t b = 0; b < nb; b++) {
      uint32_t m[16];
      row_block<D>(cw, N, r, b, m);
      const uint32_t flags =
          (b == 0 ? hd::B3_CHUNK_START : 0u) | (b + 1 == nb ? (hd::B3_CHUNK_END | hd::B3_ROOT) : 0u);
      hd::b3_compress(cv, m, 64, 0, flags);
    }
    store_digest(nodes + 32 * r, cv);
#pragma unroll
    for (int i = 0; i < 8; i++) leaf[k][i] = cv[i];
  }
  // Level 1: pair i = l/2 of group k is (lane 2i, lane 2i + 1).  Even lanes make
  // groups 0 and 2 (own leaf on the left), odd lanes groups 1 and 3 (own on the right).
  const bool odd = l & 1;
  uint32_t send[8], x1[8], x2[8], pa[8], pb[8];
  sel8(odd, leaf[0], leaf[1], send);
  dpp_xor1(send, x1);  // even: partner's leaf[0]; odd: partner's leaf[1]
  sel8(odd, leaf[2], leaf[3], send);
  dpp_xor1(send, x2);  // even: partner's leaf[2]; odd: partner's leaf[3]
  {
    // Operands are selected per lane and every lane runs the same two compressions
    // (a branch on `odd` would make the wave run both sides).
    uint32_t own_a[8], own_b[8], lhs[8], rhs[8];
    sel8(odd, leaf[1], leaf[0], own_a);
    sel8(odd, leaf[3], leaf[2], own_b);
    sel8(odd, x1, own_a, lhs);
    sel8(odd, own_a, x1, rhs);
    b3_pair(lhs, rhs, pa);
    sel8(odd, x2, own_b, lhs);
    sel8(odd, own_b, x2, rhs);
    b3_pair(lhs, rhs, pb);
  }
  const size_t i1 = l >> 1, ka = odd ? 1 : 0;
  store_digest(nodes + 32 * (N + R0 / 2 + 32 * ka + i1), pa);
  store_digest(nodes + 32 * (N + R0 / 2 + 32 * (ka + 2) + i1), pb);
  // Level 2: lane 4m + q makes group q's node m from the level-1 nodes (q, 2m) and
  // (q, 2m + 1), held by lanes 4m + (q & 1) and 4m + 2 + (q & 1) (pa for q < 2, pb else).
  const bool hi = (l >> 1) & 1;
  uint32_t y[8], gp[8];
  sel8(hi, pa, pb, send);
  dpp_xor2(send, y);  // lanes with bit 1 clear receive the partner's pa, set: partner's pb
  {
    uint32_t lhs[8], rhs[8];
    sel8(hi, y, pa, lhs);
    sel8(hi, pb, y, rhs);
    b3_pair(lhs, rhs, gp);
  }
  store_digest(nodes + 32 * (N + N / 2 + R0 / 4 + 16 * (l & 3) + (l >> 2)), gp);
}

static bool fused_tree_ok(int engine, size_t N, size_t W) {
  return engine == ENG_BLAKE3 && N >= 4 && (N & (N - 1)) == 0 && W % 8 == 0 && W > 0 && W <= 128;
}

template <int D>
static void commit_cols_t(int engine, const uint64_t* cw, size_t N, size_t M, uint8_t* nodes, hipStream_t st,
                          BatchTab tab = {}, size_t count = 1) {
  const size_t W = M * D;
  tab.per = blocks_for(N / 4, TREE_THREADS);
  const unsigned grid = tab.per * (unsigned)count;
  const double bytes = (8.0 * (double)W * N + 32.0 * (N + N / 2 + N / 4)) * count;
  if (N >= 256) {
    ProfScope prof(kernel_name("k_leaves_tree_b3w", D), st, bytes);
    k_leaves_tree_b3w<D><<<grid, TREE_THREADS, 0, st>>>(cw, N, M, nodes, tab);
  } else {
    ProfScope prof(kernel_name("k_leaves_tree_b3", D), st, bytes);
    k_leaves_tree_b3<D><<<grid, TREE_THREADS, 0, st>>>(cw, N, M, nodes, tab);
  }
  WHIR_HIP(hipGetLastError());
}

void commit_cols_dev(int engine, int D, const uint64_t* cw, size_t N, size_t M, uint8_t* nodes, hipStream_t st,
                     uint8_t* root_out) {
  WHIR_REQUIRE(engine_supports_size(engine, 8 * M * D), "leaf hash: engine does not support the row size");
  WHIR_REQUIRE(N > 0, "commit: empty codeword");
  if (!fused_tree_ok(engine, N, M * D)) {
    hash_rows_cols_dev(engine, D, cw, N, M, nodes, st);
    merkle_build_dev(engine, nodes, N, st, root_out);
    return;
  }
  switch (D) {
    case 1: commit_cols_t<1>(engine, cw, N, M, nodes, st); break;
    case 2: commit_cols_t<2>(engine, cw, N, M, nodes, st); break;
    case 3: commit_cols_t<3>(engine, cw, N, M, nodes, st); break;
    default: throw Error("leaf hash: extension degree must be 1, 2 or 3");
  }
  merkle_from(engine, nodes, N, N / 4, st, root_out);
}

// SHA3-256 / Keccak-256 trees with their top on the host.  The narrow levels of a Keccak
// tree are one lone-wave Keccak-f each (≈ 15 us per level on the GPU, the permutation
// being ~4600 dependent-issue instructions), against ≈ 0.2-0.7 us for a host
// permutation: the launches stop at a level of at most KECCAK_HOST_TOP nodes, which comes
// to the host through the mailbox (the wait the root needed anyway); the host hashes the
// levels above it and copies them back (they are contiguous at the end of the array).
static constexpr size_t KECCAK_HOST_TOP = 64;

// One node: the sponge of the 64-byte message l || r (rate 136 bytes: one permutation,
// the domain byte `pad` right after the message and 0x80 in the last byte of the rate).
static void keccak_pair_host(uint8_t pad, const uint8_t* l, const uint8_t* r, uint8_t* out) {
  uint64_t A[25] = {};
  memcpy(A, l, 32);
  memcpy(A + 4, r, 32);
  A[8] ^= pad;
  A[16] ^= 0x80ull << 56;
  keccak_f1600(A);
  memcpy(out, A, 32);
}

bool host_top_ok(int engine) { return engine == ENG_SHA3 || engine == ENG_KECCAK; }

void commit_cols_host_top(int engine, int D, const uint64_t* cw, size_t N, size_t M, uint8_t* nodes, hipStream_t st,
                          uint8_t root[32]) {
  WHIR_REQUIRE(host_top_ok(engine), "commit: the host tree top is for the Keccak engines");
  WHIR_REQUIRE(engine_supports_size(engine, 8 * M * D), "leaf hash: engine does not support the row size");
  WHIR_REQUIRE(N > 0, "commit: empty codeword");
  hash_rows_cols_dev(engine, D, cw, N, M, nodes, st);
  const size_t W = (size_t)1 << merkle_layers(N);
  if (W > N) fill_zero(nodes + 32 * N, 32 * (W - N), st);
  const size_t w = merkle_from(engine, nodes, W, W, st, nullptr, {}, 1, KECCAK_HOST_TOP);
  uint8_t* mb = (uint8_t*)result_mailbox(st);
  WHIR_REQUIRE(32 * (2 * w - 1) <= 8 * RESULT_MAILBOX_WORDS, "commit: tree top too wide for the mailbox");
  mailbox_arm(st);
  WHIR_HIP(hipMemcpyAsync(mb, nodes + 32 * (2 * W - 2 * w), 32 * w, hipMemcpyDefault, st));
  mailbox_wait(st);
  const uint8_t pad = engine == ENG_SHA3 ? 0x06 : 0x01;
  const uint8_t* cur = mb;
  uint8_t* out = mb + 32 * w;
  for (size_t width = w; width > 1; width >>= 1) {
    for (size_t i = 0; i < width / 2; i++) keccak_pair_host(pad, cur + 64 * i, cur + 64 * i + 32, out + 32 * i);
    cur = out;
    out += 32 * (width / 2);
  }
  memcpy(root, cur, 32);
  // (ordered before any later launch that writes results to the mailbox)
  if (w > 1) WHIR_HIP(hipMemcpyAsync(nodes + 32 * (2 * W - w), mb + 32 * w, 32 * (w - 1), hipMemcpyDefault, st));
}

void commit_cols_many_dev(int engine, int D, const uint64_t* const* cw_tab, uint8_t* const* nodes_tab, size_t count,
                          size_t N, size_t M, uint8_t* roots, hipStream_t st) {
  WHIR_REQUIRE(engine_supports_size(engine, 8 * M * D), "leaf hash: engine does not support the row size");
  WHIR_REQUIRE(N >= 8 && (N & (N - 1)) == 0, "batched commit: codewords of 8 or more rows, a power of two");
  if (!count) return;
  BatchTab tab;
  tab.cw = cw_tab;
  tab.nodes = nodes_tab;
  tab.roots = roots;
  if (!fused_tree_ok(engine, N, M * D)) {
    switch (D) {
      case 1: leaves_cols<1>(engine, nullptr, N, M, nullptr, st, tab, count); break;
      case 2: leaves_cols<2>(engine, nullptr, N, M, nullptr, st, tab, count); break;
      case 3: leaves_cols<3>(engine, nullptr, N, M, nullptr, st, tab, count); break;
      default: throw Error("leaf hash: extension degree must be 1, 2 or 3");
    }
    merkle_from(engine, nullptr, N, N, st, nullptr, tab, count);
    return;
  }
  switch (D) {
    case 1: commit_cols_t<1>(engine, nullptr, N, M, nullptr, st, tab, count); break;
    case 2: commit_cols_t<2>(engine, nullptr, N, M, nullptr, st, tab, count); break;
    case 3: commit_cols_t<3>(engine, nullptr, N, M, nullptr, st, tab, count); break;
    default: throw Error("leaf hash: extension degree must be 1, 2 or 3");
  }
  merkle_from(engine, nullptr, N, N / 4, st, nullptr, tab, count);
}

// ------------------------------------------------------------------ constant rows
// Every Merkle level of a codeword with equal rows holds one node value: one lane hashes
// row 0 and then pairs each level's node with itself up to the root (dig[l] = the node
// of level l, leaves at l = 0); a fill launch writes them over the levels.
template <int D, int ENG>
__global__ void k_const_chain(const uint64_t* __restrict__ cw, size_t N, size_t M, int levels, uint32_t* dig,
                              uint32_t* root_out, MailboxSig sig) {
  struct Row0 {
    const uint64_t* cw;
    size_t N, size;
    __device__ uint64_t w64(size_t off) const {
      if (off >= size) return 0;
      const size_t word = off >> 3, j = word / D, d = word % D;
      return cw[j * N * D + d];
    }
    __device__ uint32_t w32(size_t off) const { return (uint32_t)(w64(off & ~(size_t)7) >> (8 * (off & 7))); }
  };
  if (threadIdx.x != 0) return;
  uint32_t h[8];
  hd::hash_msg_t<ENG>(Row0{cw, N, 8 * M * D}, 8 * M * D, h);
#pragma unroll
  for (int i = 0; i < 8; i++) dig[i] = h[i];
  for (int l = 1; l <= levels; l++) {
    uint32_t o[8];
    hash_pair<ENG>(h, h, o);
#pragma unroll
    for (int i = 0; i < 8; i++) dig[8 * l + i] = h[i] = o[i];
  }
  if (root_out) {
#pragma unroll
    for (int i = 0; i < 8; i++) root_out[i] = h[i];
    mailbox_raise(sig);
  }
}

// Node e of a tree over W leaves (W a power of two; level widths W, W/2, ..., 1 laid
// out one after another) gets its level's digest.
__global__ void k_const_fill(uint8_t* __restrict__ nodes, size_t W, const uint32_t* __restrict__ dig) {
  const size_t total = 2 * W - 1;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
    size_t base = 0, width = W;  // level l covers [base, base + width)
    int l = 0;
    while (e >= base + width) {
      base += width;
      width >>= 1;
      l++;
    }
    uint32_t h[8];
#pragma unroll
    for (int i = 0; i < 8; i++) h[i] = dig[8 * l + i];
    store_digest(nodes + 32 * e, h);
  }
}

void commit_const_rows_dev(int engine, int D, const uint64_t* cw, size_t N, size_t M, uint8_t* nodes,
                           hipStream_t st, uint8_t* root_out) {
  WHIR_REQUIRE(engine_supports_size(engine, 8 * M * D) && engine_supports_size(engine, 64),
               "commit: engine does not support the row or node size");
  WHIR_REQUIRE(N > 0 && (N & (N - 1)) == 0, "commit: constant rows need a power-of-two codeword length");
  const int levels = (int)merkle_layers(N);
  DevBuf dig((size_t)(levels + 1) * 32, st);
  {
    ProfScope prof("k_const_chain", st, 8.0 * M * D);
    const MailboxSig sig = root_out ? mailbox_take(root_out) : MailboxSig{};
    engine_switch(engine, [&](auto E) {
      constexpr int ENG = decltype(E)::value;
      switch (D) {
        case 1: k_const_chain<1, ENG><<<1, 64, 0, st>>>(cw, N, M, levels, dig.as<uint32_t>(), (uint32_t*)root_out, sig); break;
        case 2: k_const_chain<2, ENG><<<1, 64, 0, st>>>(cw, N, M, levels, dig.as<uint32_t>(), (uint32_t*)root_out, sig); break;
        case 3: k_const_chain<3, ENG><<<1, 64, 0, st>>>(cw, N, M, levels, dig.as<uint32_t>(), (uint32_t*)root_out, sig); break;
        default: throw Error("leaf hash: extension degree must be 1, 2 or 3");
      }
    });
    WHIR_HIP(hipGetLastError());
  }
  ProfScope prof("k_const_fill", st, 32.0 * (2.0 * N - 1));
  k_const_fill<<<(unsigned)std::min<size_t>(blocks_for(2 * N - 1, 256), 4096), 256, 0, st>>>(nodes, N,
                                                                                             dig.as<uint32_t>());
  WHIR_HIP(hipGetLastError());
}

// Constant rows with the root made on the host (BLAKE3, rows of one chunk): the tree's
// digests are a chain of ~20 dependent compressions, which one lane takes ~1.7 us each to
// issue but the host runs in well under a microsecond each.  Row 0 comes to the host
// through the stream's mailbox (one wait, which the caller needs for the root anyway);
// the node array is then filled from the digests passed with the launch.
static constexpr int CONST_MAX_LEVELS = 40;
struct ConstDigests {
  uint32_t w[CONST_MAX_LEVELS + 1][8];
};

__global__ void k_const_fill_arg(uint8_t* __restrict__ nodes, size_t W, ConstDigests dig) {
  const size_t total = 2 * W - 1;
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
    size_t base = 0, width = W;  // level l covers [base, base + width)
    int l = 0;
    while (e >= base + width) {
      base += width;
      width >>= 1;
      l++;
    }
    store_digest(nodes + 32 * e, dig.w[l]);
  }
}

// Row 0 of the column-major codeword (word t = cw[(t / D) N D + t % D]) into the mailbox.
__global__ void k_row0_mailbox(const uint64_t* __restrict__ cw, size_t N, int D, size_t words, uint64_t* out,
                               MailboxSig sig) {
  for (size_t t = threadIdx.x; t < words; t += blockDim.x) out[t] = cw[(t / D) * N * D + t % D];
  __threadfence_system();  // every writer's stores reach host memory before the flag
  __syncthreads();
  if (threadIdx.x == 0) mailbox_raise(sig);
}

// BLAKE3 compression on the host (the public specification: 7 rounds of the G mixing
// function over columns then diagonals, message words permuted between rounds).
static void b3_compress_host(uint32_t cv[8], const uint32_t msg[16], uint32_t block_len, uint32_t flags) {
  static const uint32_t IV[8] = {0x6A09E667u, 0xBB67AE85u, 0x3C6EF372u, 0xA54FF53Au,
                                 0x510E527Fu, 0x9B05688Cu, 0x1F83D9ABu, 0x5BE0CD19u};
  static const int PERM[16] = {2, 6, 3, 10, 7, 0, 4, 13, 1, 11, 12, 5, 9, 14, 15, 8};
  auto rotr = [](uint32_t x, int n) { return (x >> n) | (x << (32 - n)); };
  uint32_t v[16] = {cv[0], cv[1], cv[2], cv[3], cv[4], cv[5], cv[6], cv[7],
                    IV[0], IV[1], IV[2], IV[3], 0u, 0u, block_len, flags};
  uint32_t m[16];
  std::copy(msg, msg + 16, m);
  auto g = [&](int a, int b, int c, int d, uint32_t x, uint32_t y) {
    v[a] = v[a] + v[b] + x;
    v[d] = rotr(v[d] ^ v[a], 16);
    v[c] = v[c] + v[d];
    v[b] = rotr(v[b] ^ v[c], 12);
    v[a] = v[a] + v[b] + y;
    v[d] = rotr(v[d] ^ v[a], 8);
    v[c] = v[c] + v[d];
    v[b] = rotr(v[b] ^ v[c], 7);
  };
  for (int r = 0; r < 7; r++) {
    g(0, 4, 8, 12, m[0], m[1]);
    g(1, 5, 9, 13, m[2], m[3]);
    g(2, 6, 10, 14, m[4], m[5]);
    g(3, 7, 11, 15, m[6], m[7]);
    g(0, 5, 10, 15, m[8], m[9]);
    g(1, 6, 11, 12, m[10], m[11]);
    g(2, 7, 8, 13, m[12], m[13]);
    g(3, 4, 9, 14, m[14], m[15]);
    if (r < 6) {
      uint32_t t[16];
      for (int i = 0; i < 16; i++) t[i] = m[PERM[i]];
      std::copy(t, t + 16, m);
    }
  }
  for (int i = 0; i < 8; i++) cv[i] = v[i] ^ v[i + 8];
}

bool const_rows_on_host(int engine, size_t row_bytes, size_t N) {
  return engine == ENG_BLAKE3 && row_bytes > 0 && row_bytes <= 1024 && N > 0 && (N & (N - 1)) == 0 &&
         (int)merkle_layers(N) <= CONST_MAX_LEVELS;
}

void commit_const_rows_host(int D, const uint64_t* cw, size_t N, size_t M, uint8_t* nodes, hipStream_t st,
                            uint8_t root[32]) {
  const size_t words = M * D, bytes = 8 * words;
  WHIR_REQUIRE(const_rows_on_host(ENG_BLAKE3, bytes, N), "commit: constant rows on the host need BLAKE3 rows of one chunk");
  uint64_t* mbox = result_mailbox(st);
  mailbox_arm(st);
  k_row0_mailbox<<<1, 128, 0, st>>>(cw, N, D, words, mbox, mailbox_take(mbox));
  WHIR_HIP(hipGetLastError());
  mailbox_wait(st);
  // Leaf: BLAKE3 of the row's bytes (one chunk: CHUNK_START on its first block,
  // CHUNK_END | ROOT on its last), words little-endian.
  static const uint32_t IV[8] = {0x6A09E667u, 0xBB67AE85u, 0x3C6EF372u, 0xA54FF53Au,
                                 0x510E527Fu, 0x9B05688Cu, 0x1F83D9ABu, 0x5BE0CD19u};
  ConstDigests dig{};
  uint32_t h[8];
  std::copy(IV, IV + 8, h);
  const size_t nblocks = (bytes + 63) / 64;
  for (size_t b = 0; b < nblocks; b++) {
    uint32_t m[16] = {};
    const size_t len = std::min<size_t>(64, bytes - 64 * b);
    for (size_t i = 0; i < len / 8; i++) {
      const uint64_t w = mbox[8 * b + i];
      m[2 * i] = (uint32_t)w;
      m[2 * i + 1] = (uint32_t)(w >> 32);
    }
    const uint32_t flags = (b == 0 ? 1u : 0u) | (b + 1 == nblocks ? 2u | 8u : 0u);
    b3_compress_host(h, m, (uint32_t)len, flags);
  }
  std::copy(h, h + 8, dig.w[0]);
  // Level l's nodes all hash the 64-byte pair (d, d) of level l - 1's digest d.
  const int levels = (int)merkle_layers(N);
  for (int l = 1; l <= levels; l++) {
    uint32_t m[16];
    for (int i = 0; i < 8; i++) m[i] = m[8 + i] = h[i];
    std::copy(IV, IV + 8, h);
    b3_compress_host(h, m, 64, 1u | 2u | 8u);
    std::copy(h, h + 8, dig.w[l]);
  }
  memcpy(root, h, 32);
  ProfScope prof("k_const_fill", st, 32.0 * (2.0 * N - 1));
  k_const_fill_arg<<<(unsigned)std::min<size_t>(blocks_for(2 * N - 1, 256), 4096), 256, 0, st>>>(nodes, N, dig);
  WHIR_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ PoW
static constexpr int POW_THREADS = 256;
static constexpr int POW_PER_LANE = 16;
static constexpr uint64_t POW_CHUNK = (uint64_t)POW_THREADS * POW_PER_LANE;
static constexpr int POW_BLOCKS_PER_CU = 8;  // 16 and 32 measured no faster on MI355X

// state[0] = minimal nonce found (init UINT64_MAX), state[1] = next chunk index.
// A chunk is POW_THREADS * per_lane consecutive nonces (per_lane <= POW_PER_LANE, a
// power of two, so chunks tile the POW_CHUNK-aligned window).
template <int ENG>
__global__ __launch_bounds__(POW_THREADS) void k_pow(uint4 ch_lo, uint4 ch_hi, uint64_t threshold,
                                                     uint64_t first, uint64_t num_chunks, int per_lane,
                                                     unsigned long long* state) {
  const uint64_t chunk_len = (uint64_t)POW_THREADS * per_lane;
  __shared__ unsigned long long s_chunk;
  __shared__ int s_stop;
  const uint32_t chw[8] = {ch_lo.x, ch_lo.y, ch_lo.z, ch_lo.w, ch_hi.x, ch_hi.y, ch_hi.z, ch_hi.w};
  for (;;) {
    if (threadIdx.x == 0) {
      unsigned long long c = atomicAdd(&state[1], 1ull);
      unsigned long long best = __hip_atomic_load(&state[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      // Stop when the chunks run out or every remaining chunk lies above the current minimum.
      s_stop = (c >= num_chunks) || (first + c * chunk_len > best);
      s_chunk = c;
    }
    __syncthreads();
    const unsigned long long chunk = s_chunk;
    const int stop = s_stop;
    __syncthreads();
    if (stop) break;
    uint64_t start = first + chunk * chunk_len;
    for (int k = 0; k < per_lane; k++) {
      uint64_t nonce = start + (uint64_t)k * POW_THREADS + threadIdx.x;
      uint32_t h[8];
      if constexpr (ENG == ENG_BLAKE3) {
        uint32_t m[16];
#pragma unroll
        for (int i = 0; i < 8; i++) m[i] = chw[i];
        m[8] = (uint32_t)nonce;
        m[9] = (uint32_t)(nonce >> 32);
#pragma unroll
        for (int i = 10; i < 16; i++) m[i] = 0;
#pragma unroll
        for (int i = 0; i < 8; i++) h[i] = hd::B3_IV[i];
        hd::b3_compress(h, m, 64, 0, hd::B3_CHUNK_START | hd::B3_CHUNK_END | hd::B3_ROOT);
      } else {
        uint64_t words[8];
#pragma unroll
        for (int i = 0; i < 4; i++) words[i] = (uint64_t)chw[2 * i] | ((uint64_t)chw[2 * i + 1] << 32);
        words[4] = nonce;
        words[5] = words[6] = words[7] = 0;
        hd::WordReader rd{words, 64};
        hd::hash_msg_t<ENG>(rd, 64, h);
      }
      uint64_t v = (uint64_t)h[0] | ((uint64_t)h[1] << 32);
      if (v <= threshold) atomicMin(&state[0], (unsigned long long)nonce);
    }
  }
}

__global__ void k_pow_init(unsigned long long* state) {
  if (threadIdx.x == 0) {
    state[0] = ~0ull;
    state[1] = 0ull;
  }
}

bool pow_grind_dev(int engine, const uint8_t challenge[32], uint64_t threshold, uint64_t first, uint64_t count,
                   uint64_t* nonce, hipStream_t st) {
  WHIR_REQUIRE(engine_supports_size(engine, 64), "pow: engine does not hash 64-byte messages");
  WHIR_REQUIRE(count > 0 && count % POW_CHUNK == 0 && first % POW_CHUNK == 0,
               "pow: window must be whole chunks");
  uint32_t w[8];
  for (int i = 0; i < 8; i++)
    w[i] = (uint32_t)challenge[4 * i] | ((uint32_t)challenge[4 * i + 1] << 8) |
           ((uint32_t)challenge[4 * i + 2] << 16) | ((uint32_t)challenge[4 * i + 3] << 24);
  DevBuf state(16, st);
  k_pow_init<<<1, 64, 0, st>>>(state.as<unsigned long long>());
  WHIR_HIP(hipGetLastError());
  int dev = 0, cus = 0;
  WHIR_HIP(hipGetDevice(&dev));
  WHIR_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  // Every block takes at least one chunk before the first minimum is known, so the
  // nonces in flight are hashed whatever the difficulty.  Size them to the expected
  // work (2^64 / (threshold + 1) hashes): a low-difficulty search (the query-phase
  // PoW, 16-20 bits) runs one-nonce-per-lane chunks on fewer blocks instead of
  // hashing millions of nonces it does not need.
  const double expected = 18446744073709551616.0 / ((double)threshold + 1.0);
  int per_lane = POW_PER_LANE;
  uint64_t max_blocks = (uint64_t)cus * POW_BLOCKS_PER_CU;
  while (per_lane > 1 && (double)max_blocks * POW_THREADS * per_lane > expected / 2) per_lane >>= 1;
  while (max_blocks > (uint64_t)cus && (double)max_blocks * POW_THREADS * per_lane > expected / 2) max_blocks >>= 1;
  const uint64_t chunk_len = (uint64_t)POW_THREADS * per_lane;
  uint64_t num_chunks = count / chunk_len;
  unsigned blocks = (unsigned)std::min<uint64_t>(max_blocks, num_chunks);
  ProfScope prof("pow", st, 0.0);
  engine_switch(engine, [&](auto E) {
    k_pow<decltype(E)::value><<<blocks, POW_THREADS, 0, st>>>(make_uint4(w[0], w[1], w[2], w[3]),
                                                             make_uint4(w[4], w[5], w[6], w[7]), threshold, first,
                                                             num_chunks, per_lane, state.as<unsigned long long>());
  });
  WHIR_HIP(hipGetLastError());
  unsigned long long host = 0;
  download_sync(&host, state.ptr, 8, st);
  if (host == ~0ull) return false;
  *nonce = host;
  return true;
}

uint64_t pow_chunk() { return POW_CHUNK; }

}  // namespace whir
