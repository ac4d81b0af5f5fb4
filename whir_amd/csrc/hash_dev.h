// Per-lane hash functions for CDNA4 (one message per lane).
//
// BLAKE3 (the reference wraps blake3 1.8.3, src/hash/blake3_engine.rs:88-160; for
// sizes that are multiples of 64 up to 1024 bytes it is a single chunk with
// CHUNK_START on the first block and CHUNK_END|ROOT on the last, counter 0),
// SHA3-256 / Keccak-256 (sha3 0.10.8, digest_engine.rs:22-23) and SHA-256
// (digest_engine.rs:21), written from the public specifications.
//
// A Reader supplies message bytes: w32(off) / w64(off) return the little-endian
// word at byte offset `off`, zero past the message end.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace whir {
namespace hd {

__device__ __forceinline__ uint32_t rotr32(uint32_t x, int n) { return __builtin_amdgcn_alignbit(x, x, n); }
__device__ __forceinline__ uint64_t rotl64(uint64_t x, int n) { return (x << n) | (x >> ((64 - n) & 63)); }

// 64-bit rotate left by a compile-time amount as two 32-bit funnel shifts
// (v_alignbit_b32, full rate) instead of 64-bit shifts and ors.
template <int N>
__device__ __forceinline__ uint64_t rotl64c(uint64_t x) {
  static_assert(N >= 0 && N < 64, "rotate amount");
  uint32_t lo = (uint32_t)x, hi = (uint32_t)(x >> 32);
  if constexpr (N >= 32) {
    const uint32_t t = lo;
    lo = hi;
    hi = t;
  }
  constexpr int n = N & 31;
  if constexpr (n == 0) {
    return ((uint64_t)hi << 32) | lo;
  } else {
    const uint32_t nhi = __builtin_amdgcn_alignbit(hi, lo, 32 - n);  // (hi << n) | (lo >> (32 - n))
    const uint32_t nlo = __builtin_amdgcn_alignbit(lo, hi, 32 - n);  // (lo << n) | (hi >> (32 - n))
    return ((uint64_t)nhi << 32) | nlo;
  }
}

// Reader over global memory (any alignment, any size).
struct MemReader {
  const uint8_t* p;
  size_t size;
  __device__ __forceinline__ uint64_t w64(size_t off) const {
    if (off + 8 <= size && ((((uintptr_t)p) + off) & 7) == 0) return *(const uint64_t*)(p + off);
    uint64_t v = 0;
    for (int b = 0; b < 8; b++)
      if (off + b < size) v |= (uint64_t)p[off + b] << (8 * b);
    return v;
  }
  __device__ __forceinline__ uint32_t w32(size_t off) const {
    if (off + 4 <= size && ((((uintptr_t)p) + off) & 3) == 0) return *(const uint32_t*)(p + off);
    uint32_t v = 0;
    for (int b = 0; b < 4; b++)
      if (off + b < size) v |= (uint32_t)p[off + b] << (8 * b);
    return v;
  }
};

// Reader over a row of u64 words (size a multiple of 8 bytes), e.g. in LDS.
struct WordReader {
  const uint64_t* w;
  size_t size;
  __device__ __forceinline__ uint64_t w64(size_t off) const { return off < size ? w[off >> 3] : 0; }
  __device__ __forceinline__ uint32_t w32(size_t off) const {
    return off < size ? (uint32_t)(w[off >> 3] >> (8 * (off & 7))) : 0u;
  }
};

// ------------------------------------------------------------------ BLAKE3
__constant__ static const uint32_t B3_IV[8] = {0x6A09E667u, 0xBB67AE85u, 0x3C6EF372u, 0xA54FF53Au,
                                                0x510E527Fu, 0x9B05688Cu, 0x1F83D9ABu, 0x5BE0CD19u};
enum { B3_CHUNK_START = 1, B3_CHUNK_END = 2, B3_PARENT = 4, B3_ROOT = 8 };

#define B3G(a, b, c, d, x, y)          \
  do {                                 \
    a = a + b + (x);                   \
    d = rotr32(d ^ a, 16);             \
    c = c + d;                         \
    b = rotr32(b ^ c, 12);             \
    a = a + b + (y);                   \
    d = rotr32(d ^ a, 8);              \
    c = c + d;                         \
    b = rotr32(b ^ c, 7);              \
  } while (0)

// Compression; writes the 8-word chaining value to cv (in place).
__device__ __forceinline__ void b3_compress(uint32_t cv[8], const uint32_t mm[16], uint32_t block_len,
                                            uint64_t counter, uint32_t flags) {
  uint32_t s0 = cv[0], s1 = cv[1], s2 = cv[2], s3 = cv[3], s4 = cv[4], s5 = cv[5], s6 = cv[6], s7 = cv[7];
  uint32_t s8 = 0x6A09E667u, s9 = 0xBB67AE85u, s10 = 0x3C6EF372u, s11 = 0xA54FF53Au;
  uint32_t s12 = (uint32_t)counter, s13 = (uint32_t)(counter >> 32), s14 = block_len, s15 = flags;
  uint32_t m[16];
#pragma unroll
  for (int i = 0; i < 16; i++) m[i] = mm[i];
#pragma unroll
  for (int r = 0; r < 7; r++) {
    B3G(s0, s4, s8, s12, m[0], m[1]);
    B3G(s1, s5, s9, s13, m[2], m[3]);
    B3G(s2, s6, s10, s14, m[4], m[5]);
    B3G(s3, s7, s11, s15, m[6], m[7]);
    B3G(s0, s5, s10, s15, m[8], m[9]);
    B3G(s1, s6, s11, s12, m[10], m[11]);
    B3G(s2, s7, s8, s13, m[12], m[13]);
    B3G(s3, s4, s9, s14, m[14], m[15]);
    if (r < 6) {
      // message permutation (2,6,3,10,7,0,4,13,1,11,12,5,9,14,15,8)
      uint32_t t[16] = {m[2], m[6], m[3], m[10], m[7], m[0], m[4], m[13],
                        m[1], m[11], m[12], m[5], m[9], m[14], m[15], m[8]};
#pragma unroll
      for (int i = 0; i < 16; i++) m[i] = t[i];
    }
  }
  cv[0] = s0 ^ s8; cv[1] = s1 ^ s9; cv[2] = s2 ^ s10; cv[3] = s3 ^ s11;
  cv[4] = s4 ^ s12; cv[5] = s5 ^ s13; cv[6] = s6 ^ s14; cv[7] = s7 ^ s15;
}

template <class R>
__device__ __forceinline__ void b3_chunk_cv(const R& rd, size_t off, size_t len, uint64_t chunk, bool root,
                                            uint32_t cv[8]) {
#pragma unroll
  for (int i = 0; i < 8; i++) cv[i] = B3_IV[i];
  size_t nblocks = len == 0 ? 1 : (len + 63) / 64;
  for (size_t b = 0; b < nblocks; b++) {
    uint32_t m[16];
    size_t bo = off + 64 * b;
    size_t bl = len == 0 ? 0 : ((len - 64 * b) < 64 ? (len - 64 * b) : 64);
#pragma unroll
    for (int i = 0; i < 16; i++) m[i] = (4 * (size_t)i < bl) ? rd.w32(bo + 4 * i) : 0u;
    if (bl & 3) {  // partial trailing word: mask bytes past the block end
      int wi = (int)(bl >> 2);
      uint32_t keep = (1u << (8 * (bl & 3))) - 1;
      m[wi] &= keep;
    }
    uint32_t flags = (b == 0 ? B3_CHUNK_START : 0) | (b == nblocks - 1 ? (B3_CHUNK_END | (root ? B3_ROOT : 0)) : 0);
    b3_compress(cv, m, (uint32_t)bl, chunk, flags);
  }
}

__device__ __forceinline__ void b3_parent(const uint32_t l[8], const uint32_t r[8], bool root, uint32_t out[8]) {
  uint32_t m[16];
#pragma unroll
  for (int i = 0; i < 8; i++) { m[i] = l[i]; m[8 + i] = r[i]; out[i] = B3_IV[i]; }
  b3_compress(out, m, 64, 0, B3_PARENT | (root ? B3_ROOT : 0));
}

// Tree mode for inputs above one chunk (cold path, kept out of line).
template <class R>
__device__ __noinline__ void blake3_multichunk(const R rd, size_t size, uint32_t* out) {
  uint32_t stack[54][8];
  int sp = 0;
  size_t nchunks = (size + 1023) / 1024;
  for (size_t c = 0; c + 1 < nchunks; c++) {
    uint32_t cv[8];
    b3_chunk_cv(rd, c * 1024, 1024, c, false, cv);
    uint64_t total = c + 1;
    while ((total & 1) == 0) {
      uint32_t p[8];
      b3_parent(stack[sp - 1], cv, false, p);
      sp--;
#pragma unroll
      for (int i = 0; i < 8; i++) cv[i] = p[i];
      total >>= 1;
    }
#pragma unroll
    for (int i = 0; i < 8; i++) stack[sp][i] = cv[i];
    sp++;
  }
  uint32_t cv[8];
  size_t last = nchunks - 1;
  b3_chunk_cv(rd, last * 1024, size - last * 1024, last, false, cv);
  while (sp > 0) {
    uint32_t p[8];
    b3_parent(stack[sp - 1], cv, sp == 1, p);
    sp--;
#pragma unroll
    for (int i = 0; i < 8; i++) cv[i] = p[i];
  }
#pragma unroll
  for (int i = 0; i < 8; i++) out[i] = cv[i];
}

// Full BLAKE3-256 of `size` bytes.
template <class R>
__device__ __forceinline__ void blake3(const R& rd, size_t size, uint32_t out[8]) {
  if (size <= 1024) {
    b3_chunk_cv(rd, 0, size, 0, true, out);
    return;
  }
  blake3_multichunk(rd, size, out);
}

// ------------------------------------------------------------------ Keccak
__constant__ static const uint64_t KRC[24] = {
    0x0000000000000001ull, 0x0000000000008082ull, 0x800000000000808aull, 0x8000000080008000ull,
    0x000000000000808bull, 0x0000000080000001ull, 0x8000000080008081ull, 0x8000000000008009ull,
    0x000000000000008aull, 0x0000000000000088ull, 0x0000000080008009ull, 0x000000008000000aull,
    0x000000008000808bull, 0x800000000000008bull, 0x8000000000008089ull, 0x8000000000008003ull,
    0x8000000000008002ull, 0x8000000000000080ull, 0x000000000000800aull, 0x800000008000000aull,
    0x8000000080008081ull, 0x8000000000008080ull, 0x0000000080000001ull, 0x8000000080008008ull};

// Three-input XOR in one VALU instruction: v_bitop3_b32 with the parity truth table
// 0x96 (symmetric, so the operand order does not matter); the compiler emits two
// two-input XORs for a ^ b ^ c.
__device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) {
  uint32_t r;
  asm("v_bitop3_b32 %0, %1, %2, %3 bitop3:0x96" : "=v"(r) : "v"(a), "v"(b), "v"(c));
  return r;
}

// 64-bit lane (lo, hi) rotated left by a compile-time amount: two funnel shifts.
template <int N>
__device__ __forceinline__ void rot2(uint32_t lo, uint32_t hi, uint32_t& olo, uint32_t& ohi) {
  static_assert(N >= 0 && N < 64, "rotate amount");
  if constexpr (N >= 32) {
    const uint32_t t = lo;
    lo = hi;
    hi = t;
  }
  constexpr int n = N & 31;
  if constexpr (n == 0) {
    olo = lo;
    ohi = hi;
  } else {
    ohi = __builtin_amdgcn_alignbit(hi, lo, 32 - n);
    olo = __builtin_amdgcn_alignbit(lo, hi, 32 - n);
  }
}

// Keccak-f[1600] on the state as 32-bit halves (FIPS 202 round function).  Written
// on 32-bit words so that chi, a ^ (~b & c), is one v_bitop3_b32 per half (on
// 64-bit values the compiler splits it into a bit-field insert and an XOR) and the
// column parities use three-input XORs.
__device__ __forceinline__ void keccak_f(uint64_t A[25]) {
  uint32_t L[25], H[25];
#pragma unroll
  for (int i = 0; i < 25; i++) {
    L[i] = (uint32_t)A[i];
    H[i] = (uint32_t)(A[i] >> 32);
  }
  for (int r = 0; r < 24; r++) {
    uint32_t CL[5], CH[5], DL[5], DH[5];
#pragma unroll
    for (int x = 0; x < 5; x++) {
      CL[x] = xor3(xor3(L[x], L[x + 5], L[x + 10]), L[x + 15], L[x + 20]);
      CH[x] = xor3(xor3(H[x], H[x + 5], H[x + 10]), H[x + 15], H[x + 20]);
    }
#pragma unroll
    for (int x = 0; x < 5; x++) {
      uint32_t rl, rh;
      rot2<1>(CL[(x + 1) % 5], CH[(x + 1) % 5], rl, rh);
      DL[x] = CL[(x + 4) % 5] ^ rl;
      DH[x] = CH[(x + 4) % 5] ^ rh;
    }
    // theta + rho + pi: B[y, 2x+3y] = rot(A[x, y] ^ D[x], r[x, y])
    uint32_t BL[25], BH[25];
#define KECCAK_RP(src, dst, rot) rot2<rot>(L[src] ^ DL[(src) % 5], H[src] ^ DH[(src) % 5], BL[dst], BH[dst]);
    KECCAK_RP(0, 0, 0)
    KECCAK_RP(1, 10, 1)
    KECCAK_RP(2, 20, 62)
    KECCAK_RP(3, 5, 28)
    KECCAK_RP(4, 15, 27)
    KECCAK_RP(5, 16, 36)
    KECCAK_RP(6, 1, 44)
    KECCAK_RP(7, 11, 6)
    KECCAK_RP(8, 21, 55)
    KECCAK_RP(9, 6, 20)
    KECCAK_RP(10, 7, 3)
    KECCAK_RP(11, 17, 10)
    KECCAK_RP(12, 2, 43)
    KECCAK_RP(13, 12, 25)
    KECCAK_RP(14, 22, 39)
    KECCAK_RP(15, 23, 41)
    KECCAK_RP(16, 8, 45)
    KECCAK_RP(17, 18, 15)
    KECCAK_RP(18, 3, 21)
    KECCAK_RP(19, 13, 8)
    KECCAK_RP(20, 14, 18)
    KECCAK_RP(21, 24, 2)
    KECCAK_RP(22, 9, 61)
    KECCAK_RP(23, 19, 56)
    KECCAK_RP(24, 4, 14)
#undef KECCAK_RP
    // chi
#pragma unroll
    for (int y = 0; y < 25; y += 5) {
#pragma unroll
      for (int x = 0; x < 5; x++) {
        L[y + x] = BL[y + x] ^ (~BL[y + (x + 1) % 5] & BL[y + (x + 2) % 5]);
        H[y + x] = BH[y + x] ^ (~BH[y + (x + 1) % 5] & BH[y + (x + 2) % 5]);
      }
    }
    // iota
    const uint64_t rc = KRC[r];
    L[0] ^= (uint32_t)rc;
    H[0] ^= (uint32_t)(rc >> 32);
  }
#pragma unroll
  for (int i = 0; i < 25; i++) A[i] = ((uint64_t)H[i] << 32) | L[i];
}

// Keccak sponge with rate 136 and domain byte `pad` (0x06 SHA3-256, 0x01 Keccak-256).
template <class R>
__device__ void keccak256(const R& rd, size_t size, uint8_t pad, uint32_t out[8]) {
  uint64_t A[25];
#pragma unroll
  for (int i = 0; i < 25; i++) A[i] = 0;
  size_t nblocks = size / 136 + 1;
  for (size_t b = 0; b < nblocks; b++) {
    size_t bo = 136 * b;
#pragma unroll
    for (int i = 0; i < 17; i++) {
      size_t off = bo + 8 * i;
      uint64_t v = off < size ? rd.w64(off) : 0;
      if (off < size && off + 8 > size) v &= (1ull << (8 * (size - off))) - 1;  // bytes past the end
      A[i] ^= v;
    }
    if (b == nblocks - 1) {
      // Padding: the domain byte at position pp of the block.  The word is chosen by
      // comparing against every static index, so the state stays in registers (a
      // runtime index into A would put it in scratch memory).
      const size_t pp = size - bo;
      const uint64_t padw = (uint64_t)pad << (8 * (pp & 7));
#pragma unroll
      for (int i = 0; i < 17; i++)
        if ((size_t)i == (pp >> 3)) A[i] ^= padw;
      A[16] ^= 0x80ull << 56;
    }
    keccak_f(A);
  }
#pragma unroll
  for (int i = 0; i < 4; i++) {
    out[2 * i] = (uint32_t)A[i];
    out[2 * i + 1] = (uint32_t)(A[i] >> 32);
  }
}

// ------------------------------------------------------------------ SHA-256
__constant__ static const uint32_t SHA_K[64] = {
    0x428a2f98u, 0x71374491u, 0xb5c0fbcfu, 0xe9b5dba5u, 0x3956c25bu, 0x59f111f1u, 0x923f82a4u, 0xab1c5ed5u,
    0xd807aa98u, 0x12835b01u, 0x243185beu, 0x550c7dc3u, 0x72be5d74u, 0x80deb1feu, 0x9bdc06a7u, 0xc19bf174u,
    0xe49b69c1u, 0xefbe4786u, 0x0fc19dc6u, 0x240ca1ccu, 0x2de92c6fu, 0x4a7484aau, 0x5cb0a9dcu, 0x76f988dau,
    0x983e5152u, 0xa831c66du, 0xb00327c8u, 0xbf597fc7u, 0xc6e00bf3u, 0xd5a79147u, 0x06ca6351u, 0x14292967u,
    0x27b70a85u, 0x2e1b2138u, 0x4d2c6dfcu, 0x53380d13u, 0x650a7354u, 0x766a0abbu, 0x81c2c92eu, 0x92722c85u,
    0xa2bfe8a1u, 0xa81a664bu, 0xc24b8b70u, 0xc76c51a3u, 0xd192e819u, 0xd6990624u, 0xf40e3585u, 0x106aa070u,
    0x19a4c116u, 0x1e376c08u, 0x2748774cu, 0x34b0bcb5u, 0x391c0cb3u, 0x4ed8aa4au, 0x5b9cca4fu, 0x682e6ff3u,
    0x748f82eeu, 0x78a5636fu, 0x84c87814u, 0x8cc70208u, 0x90befffau, 0xa4506cebu, 0xbef9a3f7u, 0xc67178f2u};

__device__ __forceinline__ uint32_t bswap32(uint32_t x) { return __builtin_bswap32(x); }

__device__ __forceinline__ void sha256_block(uint32_t H[8], const uint32_t Win[16]) {
  uint32_t W[16];
#pragma unroll
  for (int i = 0; i < 16; i++) W[i] = Win[i];
  uint32_t a = H[0], b = H[1], c = H[2], d = H[3], e = H[4], f = H[5], g = H[6], h = H[7];
#pragma unroll
  for (int i = 0; i < 64; i++) {
    uint32_t w;
    if (i < 16) {
      w = W[i];
    } else {
      uint32_t w15 = W[(i - 15) & 15], w2 = W[(i - 2) & 15];
      uint32_t s0 = rotr32(w15, 7) ^ rotr32(w15, 18) ^ (w15 >> 3);
      uint32_t s1 = rotr32(w2, 17) ^ rotr32(w2, 19) ^ (w2 >> 10);
      w = W[i & 15] + s0 + W[(i - 7) & 15] + s1;
      W[i & 15] = w;
    }
    uint32_t S1 = rotr32(e, 6) ^ rotr32(e, 11) ^ rotr32(e, 25);
    uint32_t ch = (e & f) ^ (~e & g);
    uint32_t t1 = h + S1 + ch + SHA_K[i] + w;
    uint32_t S0 = rotr32(a, 2) ^ rotr32(a, 13) ^ rotr32(a, 22);
    uint32_t mj = (a & b) ^ (a & c) ^ (b & c);
    uint32_t t2 = S0 + mj;
    h = g; g = f; f = e; e = d + t1; d = c; c = b; b = a; a = t1 + t2;
  }
  H[0] += a; H[1] += b; H[2] += c; H[3] += d; H[4] += e; H[5] += f; H[6] += g; H[7] += h;
}

template <class R>
__device__ void sha256(const R& rd, size_t size, uint32_t out[8]) {
  uint32_t H[8] = {0x6a09e667u, 0xbb67ae85u, 0x3c6ef372u, 0xa54ff53au, 0x510e527fu, 0x9b05688cu, 0x1f83d9abu, 0x5be0cd19u};
  size_t total = ((size + 9 + 63) / 64) * 64;
  for (size_t bo = 0; bo < total; bo += 64) {
    uint32_t W[16];
#pragma unroll
    for (int i = 0; i < 16; i++) {
      size_t off = bo + 4 * i;
      uint32_t v = 0;  // little-endian word of message bytes off..off+3, then byte-swapped
      if (off + 4 <= size) {
        v = rd.w32(off);
      } else if (off < size) {
        v = rd.w32(off) & ((1u << (8 * (size - off))) - 1);
      }
      if (off <= size && size < off + 4) v |= 0x80u << (8 * (size - off));
      W[i] = bswap32(v);
    }
    if (bo + 64 == total) {
      uint64_t bits = (uint64_t)size * 8;
      W[14] = (uint32_t)(bits >> 32);
      W[15] = (uint32_t)bits;
    }
    sha256_block(H, W);
  }
#pragma unroll
  for (int i = 0; i < 8; i++) out[i] = bswap32(H[i]);
}

// ------------------------------------------------------------------ dispatch
// 32-byte digest as 8 little-endian words, engine fixed at compile time: a kernel
// instantiated per engine holds one hash's code and is allocated its registers and
// private memory only (the BLAKE3 multi-chunk stack, for one, is private memory).
template <int ENG, class R>
__device__ __forceinline__ void hash_msg_t(const R& rd, size_t size, uint32_t out[8]) {
  if constexpr (ENG == 4) {
    blake3(rd, size, out);
  } else if constexpr (ENG == 3) {
    keccak256(rd, size, 0x06, out);
  } else if constexpr (ENG == 2) {
    keccak256(rd, size, 0x01, out);
  } else if constexpr (ENG == 1) {
    sha256(rd, size, out);
  } else {  // Copy: zero-padded input of at most 32 bytes (src/hash/copy_engine.rs:25-60)
#pragma unroll
    for (int i = 0; i < 8; i++) {
      uint32_t v = (4 * (size_t)i < size) ? rd.w32(4 * i) : 0u;
      if (4 * (size_t)i < size && size < 4 * (size_t)i + 4) v &= (1u << (8 * (size & 3))) - 1;
      out[i] = v;
    }
  }
}

// Same with the engine chosen at run time.
template <class R>
__device__ __forceinline__ void hash_msg(int engine, const R& rd, size_t size, uint32_t out[8]) {
  switch (engine) {
    case 4: blake3(rd, size, out); break;
    case 3: keccak256(rd, size, 0x06, out); break;
    case 2: keccak256(rd, size, 0x01, out); break;
    case 1: sha256(rd, size, out); break;
    default: {  // Copy: zero-padded input of at most 32 bytes (src/hash/copy_engine.rs:25-60)
#pragma unroll
      for (int i = 0; i < 8; i++) out[i] = (4 * (size_t)i < size) ? rd.w32(4 * i) : 0u;
      if (size < 32 && (size & 3)) out[size >> 2] &= (1u << (8 * (size & 3))) - 1;
    }
  }
}

}  // namespace hd
}  // namespace whir
