// WHIR parameter selection (host), restating src/protocols/whir/config.rs:16-152 and
// src/protocols/irs_commit.rs:125-288, 579-630.  All float expressions keep the
// reference's operation order so the derived integers (sample counts, PoW
// thresholds) come out identical.
#pragma once
#include <stddef.h>
#include <stdint.h>

#include <string>
#include <vector>

namespace whir {

struct IrsCfg {
  size_t num_vectors = 1;
  size_t vector_size = 0;
  size_t codeword_length = 0;
  size_t interleaving_depth = 1;
  double johnson_slack = 0.0;
  size_t in_domain_samples = 0;
  size_t out_domain_samples = 0;
  int target_degree = 1;
  size_t mask_length = 0;
  bool deduplicate_in_domain = false;

  static IrsCfg make(double security_target, bool unique_decoding, size_t num_vectors, size_t vector_size,
                     size_t interleaving_depth, double rate, int target_degree);
  size_t num_messages() const { return interleaving_depth * num_vectors; }
  size_t message_length() const { return vector_size / interleaving_depth; }
  size_t masked_message_length() const { return message_length() + mask_length; }
  double rate() const;
  bool unique_decoding() const;
  double list_size() const;
  double rbr_queries() const;
  double rbr_soundness_fold_prox_gaps() const;
};

struct SumcheckCfg {
  size_t initial_size = 0;
  uint64_t round_pow = UINT64_MAX;  // threshold
  size_t num_rounds = 0;
  size_t final_size() const;
};

struct RoundCfg {
  IrsCfg irs;
  SumcheckCfg sumcheck;
  uint64_t pow = UINT64_MAX;
};

struct WhirParams {
  int degree = 2;
  int security_level = 100;
  int pow_bits = 20;
  int initial_folding_factor = 4;
  int folding_factor = 4;
  bool unique_decoding = false;
  int starting_log_inv_rate = 1;
  int batch_size = 1;
  int hash_engine = 4;
};

struct WhirCfg {
  int D = 2;
  int engine = 4;
  IrsCfg initial;
  SumcheckCfg initial_sumcheck;
  uint64_t initial_skip_pow = UINT64_MAX;
  std::vector<RoundCfg> rounds;
  SumcheckCfg final_sumcheck;
  uint64_t final_pow = UINT64_MAX;

  static WhirCfg make(size_t size, const WhirParams& p);
  void disable_pow();
  std::string describe() const;
};

double field_size_bits(int degree);
uint64_t pow_threshold(double bits);

}  // namespace whir
