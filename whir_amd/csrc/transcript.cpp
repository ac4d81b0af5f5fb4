#include "transcript.h"

#include <string.h>

#include "common.h"
#include "keccak_host.h"

namespace whir {

static constexpr size_t RATE = 136;
static constexpr uint64_t P = 0xFFFFFFFF00000001ull;

Duplex::Duplex(const uint8_t protocol_id[64], const uint8_t session_id[32], const uint8_t* instance, size_t len) {
  absorb(protocol_id, 64);
  absorb(session_id, 32);
  if (len) absorb(instance, len);
}

void Duplex::absorb(const uint8_t* p, size_t n) {
  squeezing_ = false;
  uint8_t* s = reinterpret_cast<uint8_t*>(st_);
  for (size_t i = 0; i < n; i++) {
    if (pos_ == RATE) {
      keccak_f1600(st_);
      pos_ = 0;
    }
    s[pos_++] ^= p[i];
  }
}

void Duplex::squeeze(uint8_t* out, size_t n) {
  uint8_t* s = reinterpret_cast<uint8_t*>(st_);
  if (!squeezing_) {
    keccak_f1600(st_);
    pos_ = 0;
    squeezing_ = true;
  }
  for (size_t i = 0; i < n; i++) {
    if (pos_ == RATE) {
      keccak_f1600(st_);
      pos_ = 0;
    }
    out[i] = s[pos_++];
  }
}

static void le_bytes(uint64_t v, uint8_t* b) {
  for (int i = 0; i < 8; i++) b[i] = (uint8_t)(v >> (8 * i));
}

void Duplex::prover_field(int degree, const uint64_t* elems, size_t count) {
  std::vector<uint8_t> b(8 * degree * count);
  for (size_t i = 0; i < (size_t)degree * count; i++) le_bytes(elems[i], &b[8 * i]);
  narg_.insert(narg_.end(), b.begin(), b.end());
  absorb(b.data(), b.size());
}

void Duplex::verifier_field(int degree, uint64_t* out) {
  for (int d = 0; d < degree; d++) {
    uint8_t b[24];
    squeeze(b, 24);
    unsigned __int128 acc = 0;
    for (int k = 23; k >= 0; k--) acc = ((acc << 8) | b[k]) % P;
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

void Duplex::hint_field_vec(int degree, const uint64_t* elems, size_t count) {
  uint8_t b[8];
  le_bytes((uint64_t)count, b);
  hints_.insert(hints_.end(), b, b + 8);
  for (size_t i = 0; i < (size_t)degree * count; i++) {
    le_bytes(elems[i], b);
    hints_.insert(hints_.end(), b, b + 8);
  }
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
