#include "keccak_host.h"

#include <string.h>

namespace whir {

static const uint64_t RC[24] = {
    0x0000000000000001ull, 0x0000000000008082ull, 0x800000000000808aull, 0x8000000080008000ull,
    0x000000000000808bull, 0x0000000080000001ull, 0x8000000080008081ull, 0x8000000000008009ull,
    0x000000000000008aull, 0x0000000000000088ull, 0x0000000080008009ull, 0x000000008000000aull,
    0x000000008000808bull, 0x800000000000008bull, 0x8000000000008089ull, 0x8000000000008003ull,
    0x8000000000008002ull, 0x8000000000000080ull, 0x000000000000800aull, 0x800000008000000aull,
    0x8000000080008081ull, 0x8000000000008080ull, 0x0000000080000001ull, 0x8000000080008008ull};

// Rotation offsets r[x + 5y] and the pi lane permutation, from FIPS 202 section 3.2.
static const int ROT[25] = {0, 1, 62, 28, 27, 36, 44, 6, 55, 20, 3, 10, 43, 25, 39, 41, 45, 15, 21, 8, 18, 2, 61, 56, 14};

static inline uint64_t rol(uint64_t x, int n) { return n ? (x << n) | (x >> (64 - n)) : x; }

void keccak_f1600(uint64_t A[25]) {
  for (int round = 0; round < 24; round++) {
    uint64_t C[5], B[25];
    for (int x = 0; x < 5; x++) C[x] = A[x] ^ A[x + 5] ^ A[x + 10] ^ A[x + 15] ^ A[x + 20];
    for (int x = 0; x < 5; x++) {
      uint64_t d = C[(x + 4) % 5] ^ rol(C[(x + 1) % 5], 1);
      for (int y = 0; y < 25; y += 5) A[x + y] ^= d;
    }
    for (int x = 0; x < 5; x++)
      for (int y = 0; y < 5; y++) B[y + 5 * ((2 * x + 3 * y) % 5)] = rol(A[x + 5 * y], ROT[x + 5 * y]);
    for (int y = 0; y < 25; y += 5)
      for (int x = 0; x < 5; x++) A[x + y] = B[x + y] ^ (~B[(x + 1) % 5 + y] & B[(x + 2) % 5 + y]);
    A[0] ^= RC[round];
  }
}

std::vector<uint8_t> keccak_digest(const uint8_t* in, size_t len, size_t rate, uint8_t pad, size_t out_len) {
  uint64_t A[25];
  memset(A, 0, sizeof A);
  uint8_t block[200];
  size_t off = 0;
  for (;;) {
    size_t take = len - off < rate ? len - off : rate;
    memset(block, 0, rate);
    if (take) memcpy(block, in + off, take);
    bool last = take < rate;
    if (last) {
      block[take] ^= pad;
      block[rate - 1] ^= 0x80;
    }
    for (size_t i = 0; i < rate / 8; i++) {
      uint64_t w = 0;
      for (int b = 0; b < 8; b++) w |= (uint64_t)block[8 * i + b] << (8 * b);
      A[i] ^= w;
    }
    keccak_f1600(A);
    off += take;
    if (last) break;
  }
  std::vector<uint8_t> out(out_len);
  for (size_t i = 0; i < out_len; i++) out[i] = (uint8_t)(A[i / 8] >> (8 * (i % 8)));
  return out;
}

}  // namespace whir
