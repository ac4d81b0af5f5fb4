// Drives the library's host Duplex transcript (whir_amd/csrc/transcript.cpp) through a
// script of operations read from stdin, one per line, and prints every challenge, then
// the prover messages (narg) and the hints, in hex; tests/test_transcript_host.py runs
// the same script on the oracle's transcript and compares.
//   pf D n w...   prover_field of n elements of D words
//   vf D          verifier_field (prints D words)
//   vb n          verifier_bytes (prints n bytes)
//   ph hex        prover_hash (32 bytes)
//   pu v          prover_u64
//   hh n hex      hint_hashes of n hashes (32 n bytes)
//   hv D n w...   hint_field_vec
#include <stdint.h>
#include <stdio.h>

#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "transcript.h"

static std::vector<uint8_t> unhex(const std::string& s) {
  std::vector<uint8_t> out(s.size() / 2);
  for (size_t i = 0; i < out.size(); i++) out[i] = (uint8_t)std::stoul(s.substr(2 * i, 2), nullptr, 16);
  return out;
}

static void hex(const uint8_t* p, size_t n) {
  for (size_t i = 0; i < n; i++) printf("%02x", p[i]);
  printf("\n");
}

int main() {
  std::string line;
  std::getline(std::cin, line);  // first line: protocol id (64 bytes) and session id (32 bytes) in hex
  std::istringstream head(line);
  std::string pid_hex, sid_hex;
  head >> pid_hex >> sid_hex;
  const std::vector<uint8_t> pid = unhex(pid_hex), sid = unhex(sid_hex);
  whir::Duplex t(pid.data(), sid.data(), nullptr, 0);
  while (std::getline(std::cin, line)) {
    std::istringstream in(line);
    std::string op;
    if (!(in >> op)) continue;
    if (op == "pf" || op == "hv") {
      int D;
      size_t n;
      in >> D >> n;
      std::vector<uint64_t> w(D * n);
      for (auto& x : w) in >> x;
      if (op == "pf") t.prover_field(D, w.data(), n);
      else t.hint_field_vec(D, w.data(), n);
    } else if (op == "vf") {
      int D;
      in >> D;
      uint64_t out[3];
      t.verifier_field(D, out);
      for (int d = 0; d < D; d++) printf("%llu ", (unsigned long long)out[d]);
      printf("\n");
    } else if (op == "vb") {
      size_t n;
      in >> n;
      std::vector<uint8_t> b(n);
      t.verifier_bytes(b.data(), n);
      hex(b.data(), n);
    } else if (op == "ph") {
      std::string h;
      in >> h;
      t.prover_hash(unhex(h).data());
    } else if (op == "pu") {
      unsigned long long v;
      in >> v;
      t.prover_u64(v);
    } else if (op == "hh") {
      size_t n;
      std::string h;
      in >> n >> h;
      t.hint_hashes(unhex(h).data(), n);
    }
  }
  hex(t.narg().data(), t.narg().size());
  hex(t.hints().data(), t.hints().size());
  return 0;
}
