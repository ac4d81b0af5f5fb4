// Host Keccak-f[1600] and SHA3-256 / SHA3-512 (FIPS 202) for engine ids, the
// protocol id and the transcript sponge.
#pragma once
#include <stddef.h>
#include <stdint.h>

#include <array>
#include <string>
#include <vector>

namespace whir {

void keccak_f1600(uint64_t A[25]);

// Generic Keccak sponge digest: rate in bytes, domain padding byte, output length.
std::vector<uint8_t> keccak_digest(const uint8_t* in, size_t len, size_t rate, uint8_t pad, size_t out_len);

inline std::array<uint8_t, 32> sha3_256(const uint8_t* in, size_t len) {
  auto v = keccak_digest(in, len, 136, 0x06, 32);
  std::array<uint8_t, 32> r;
  for (int i = 0; i < 32; i++) r[i] = v[i];
  return r;
}
inline std::array<uint8_t, 64> sha3_512(const uint8_t* in, size_t len) {
  auto v = keccak_digest(in, len, 72, 0x06, 64);
  std::array<uint8_t, 64> r;
  for (int i = 0; i < 64; i++) r[i] = v[i];
  return r;
}

}  // namespace whir
