// Fiat-Shamir transcript interface of the GPU prover.
//
// The reference drives spongefish through ProverState (src/transcript/mod.rs:183-267).
// The prover only ever needs the typed operations below, so a host binding can
// plug its own transcript in (whir_transcript_from_vtable) and get proofs in its
// own format.  The built-in Duplex is this project's Keccak-f[1600] sponge.
#pragma once
#include <stddef.h>
#include <stdint.h>

#include <array>
#include <vector>

#include "../../include/whir_amd.h"

namespace whir {

struct Transcript {
  virtual ~Transcript() = default;
  // prover_message of `count` field elements of `degree` words each
  virtual void prover_field(int degree, const uint64_t* elems, size_t count) = 0;
  // verifier_message::<F>() of one element
  virtual void verifier_field(int degree, uint64_t* out) = 0;
  // `n` single-byte verifier messages (u8 challenges, or a [u8; n] array)
  virtual void verifier_bytes(uint8_t* out, size_t n) = 0;
  virtual void prover_hash(const uint8_t* h32) = 0;
  virtual void prover_u64(uint64_t v) = 0;
  virtual void hint_hash(const uint8_t* h32) = 0;
  // n hashes of 32 bytes, one hint_hash each
  virtual void hint_hashes(const uint8_t* h, size_t n) {
    for (size_t i = 0; i < n; i++) hint_hash(h + 32 * i);
  }
  // prover_hint_ark(&Vec<F>)
  virtual void hint_field_vec(int degree, const uint64_t* elems, size_t count) = 0;
};

// Keccak-f[1600] duplex, rate 136: absorb XORs into the rate (permuting when it
// is full); the first squeeze after absorbing permutes, later squeezes continue
// in the rate.  Field challenge = 24 squeezed bytes (LE) mod p per coefficient.
class Duplex final : public Transcript {
 public:
  Duplex(const uint8_t protocol_id[64], const uint8_t session_id[32], const uint8_t* instance, size_t len);
  void prover_field(int degree, const uint64_t* elems, size_t count) override;
  void verifier_field(int degree, uint64_t* out) override;
  void verifier_bytes(uint8_t* out, size_t n) override;
  void prover_hash(const uint8_t* h32) override;
  void prover_u64(uint64_t v) override;
  void hint_hash(const uint8_t* h32) override;
  void hint_hashes(const uint8_t* h, size_t n) override;
  void hint_field_vec(int degree, const uint64_t* elems, size_t count) override;

  const std::vector<uint8_t>& narg() const { return narg_; }
  const std::vector<uint8_t>& hints() const { return hints_; }

 private:
  void absorb(const uint8_t* p, size_t n);
  void squeeze(uint8_t* out, size_t n);
  uint64_t st_[25] = {};
  size_t pos_ = 0;
  bool squeezing_ = false;
  std::vector<uint8_t> narg_, hints_;
};

// Adaptor over caller-supplied callbacks (whir_transcript_vtable).
class VtableTranscript final : public Transcript {
 public:
  explicit VtableTranscript(const whir_transcript_vtable& vt) : vt_(vt) {}
  void prover_field(int degree, const uint64_t* elems, size_t count) override;
  void verifier_field(int degree, uint64_t* out) override;
  void verifier_bytes(uint8_t* out, size_t n) override;
  void prover_hash(const uint8_t* h32) override;
  void prover_u64(uint64_t v) override;
  void hint_hash(const uint8_t* h32) override;
  void hint_field_vec(int degree, const uint64_t* elems, size_t count) override;

 private:
  whir_transcript_vtable vt_;
};

}  // namespace whir
