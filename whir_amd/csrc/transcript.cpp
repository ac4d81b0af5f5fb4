#include "transcript.h"

#include <string.h>

#include <algorithm>

#include "common.h"
#include "keccak_host.h"

namespace whir {

static constexpr size_t RATE = 136;
static constexpr uint64_t P = 0xFFFFFFFF00000001ull;

Duplex::Duplex(const uint8_t protocol_id[64], const uint8_t session_id[32], const uint8_t* instance, size_t len) {
  // A d = 24 proof carries ~120 KiB of hints and ~5 KiB of prover messages; growing
  // from empty would copy them several times while the proof is being made.
  narg_.reserve(16 << 10);
  hints_.reserve(256 << 10);
  absorb(protocol_id, 64);
  absorb(session_id, 32);
  if (len) absorb(instance, len);
}

// Byte by byte, the permutation runs when a byte arrives at a full rate (not when the
// rate fills); the loops below take the bytes up to the end of the rate at a time.
void Duplex::absorb(const uint8_t* p, size_t n) {
  squeezing_ = false;
  uint8_t* s = reinterpret_cast<uint8_t*>(st_);
  while (n) {
    if (pos_ == RATE) {
      keccak_f1600(st_);
      pos_ = 0;
    }
    const size_t take = std::min(n, RATE - pos_);
    for (size_t i = 0; i < take; i++) s[pos_ + i] ^= p[i];
    pos_ += take;
    p += take;
    n -= take;
  }
}

void Duplex::squeeze(uint8_t* out, size_t n) {
  uint8_t* s = reinterpret_cast<uint8_t*>(st_);
  if (!squeezing_) {
    keccak_f1600(st_);
    pos_ = 0;
    squeezing_ = true;
  }
  while (n) {
    if (pos_ == RATE) {
      keccak_f1600(st_);
      pos_ = 0;
    }
    const size_t take = std::min(n, RATE - pos_);
    memcpy(out, s + pos_, take);
    pos_ += take;
    out += take;
    n -= take;
  }
}

static void le_bytes(uint64_t v, uint8_t* b) {
  for (int i = 0; i < 8; i++) b[i] = (uint8_t)(v >> (8 * i));
}

// n words as little-endian bytes.
static void le_words(const uint64_t* v, size_t n, uint8_t* b) {
#if defined(__BYTE_ORDER__) && __BYTE_ORDER__ == __ORDER_LITTLE_ENDIAN__
  memcpy(b, v, 8 * n);
#else
  for (size_t i = 0; i < n; i++) le_bytes(v[i], b + 8 * i);
#endif
}

void Duplex::prover_field(int degree, const uint64_t* elems, size_t count) {
  const size_t words = (size_t)degree * count, at = narg_.size();
  narg_.resize(at + 8 * words);
  le_words(elems, words, &narg_[at]);
  absorb(&narg_[at], 8 * words);
}

void Duplex::verifier_field(int degree, uint64_t* out) {
  for (int d = 0; d < degree; d++) {
    uint8_t b[24];
    squeeze(b, 24);
    // The 24 bytes as a little-endian integer mod p: Horner over its three 64-bit limbs
    // (the same value as byte by byte, with 3 reductions instead of 24).
    unsigned __int128 acc = 0;
    for (int k = 2; k >= 0; k--) {
      uint64_t limb = 0;
      for (int i = 7; i >= 0; i--) limb = (limb << 8) | b[8 * k + i];
      acc = ((acc << 64) | limb) % P;
    }
    out[d] = (uint64_t)acc;
  }
}

void Duplex::verifier_bytes(uint8_t* out, size_t n) { squeeze(out, n); }

void Duplex::prover_hash(const uint8_t* h32) {
  narg_.insert(narg_.end(), h32, h32 + 32);
  absorb(h32, 32);
}

void Duplex::prover_u64(uint64_t v) {
  uint8_t b[8];
  le_bytes(v, b);
  narg_.insert(narg_.end(), b, b + 8);
  absorb(b, 8);
}

void Duplex::hint_hash(const uint8_t* h32) { hints_.insert(hints_.end(), h32, h32 + 32); }

void Duplex::hint_hashes(const uint8_t* h, size_t n) { hints_.insert(hints_.end(), h, h + 32 * n); }

void Duplex::hint_field_vec(int degree, const uint64_t* elems, size_t count) {
  const size_t words = (size_t)degree * count, at = hints_.size();
  hints_.resize(at + 8 + 8 * words);
  le_bytes((uint64_t)count, &hints_[at]);
  le_words(elems, words, &hints_[at + 8]);
}

#define VT_CALL(expr, what) WHIR_REQUIRE((expr) == 0, "transcript callback failed: " what)

void VtableTranscript::prover_field(int degree, const uint64_t* elems, size_t count) {
  VT_CALL(vt_.prover_field(vt_.ctx, degree, elems, count), "prover_field");
}
void VtableTranscript::verifier_field(int degree, uint64_t* out) {
  VT_CALL(vt_.verifier_field(vt_.ctx, degree, out), "verifier_field");
}
void VtableTranscript::verifier_bytes(uint8_t* out, size_t n) {
  VT_CALL(vt_.verifier_bytes(vt_.ctx, out, n), "verifier_bytes");
}
void VtableTranscript::prover_hash(const uint8_t* h32) { VT_CALL(vt_.prover_hash(vt_.ctx, h32), "prover_hash"); }
void VtableTranscript::prover_u64(uint64_t v) { VT_CALL(vt_.prover_u64(vt_.ctx, v), "prover_u64"); }
void VtableTranscript::hint_hash(const uint8_t* h32) { VT_CALL(vt_.hint_hash(vt_.ctx, h32), "hint_hash"); }
void VtableTranscript::hint_field_vec(int degree, const uint64_t* elems, size_t count) {
  VT_CALL(vt_.hint_field_vec(vt_.ctx, degree, elems, count), "hint_field_vec");
}

}  // namespace whir
