#include "config.h"

#include <math.h>

#include <algorithm>
#include <sstream>

#include "common.h"

namespace whir {

static const double LOG2_10 = 3.321928094887362347870319429489390175864831393;

double field_size_bits(int degree) {
  // FieldWithSize::field_size_bits, src/algebra/fields.rs:14-30
  return log2((double)0xFFFFFFFF00000001ull) * (double)degree;
}

uint64_t pow_threshold(double bits) {
  // proof_of_work::threshold, src/protocols/proof_of_work.rs:29-39
  WHIR_REQUIRE(bits >= 0.0 && bits <= 60.0, "pow threshold: difficulty must be within [0, 60] bits");
  double t = ceil(exp2(64.0 - bits));
  if (t >= 18446744073709551615.0) return UINT64_MAX;
  return (uint64_t)t;
}

static size_t num_ood_samples(bool unique_decoding, double security_target, double fsb, double list_size,
                              size_t degree) {
  // src/protocols/irs_commit.rs:584-601
  if (unique_decoding) return 0;
  double l_choose_2 = list_size * (list_size - 1.0) / 2.0;
  double log_per_sample = fsb - log2((double)(degree - 1));
  WHIR_REQUIRE(log_per_sample > 0.0, "ood samples: field too small for the degree");
  return (size_t)std::max(ceil((security_target + log2(l_choose_2)) / log_per_sample), 1.0);
}

IrsCfg IrsCfg::make(double security_target, bool unique_decoding, size_t num_vectors, size_t vector_size,
                    size_t interleaving_depth, double rate, int target_degree) {
  // irs_commit::Config::new, src/protocols/irs_commit.rs:125-196
  WHIR_REQUIRE(vector_size % interleaving_depth == 0, "irs: vector size not a multiple of the interleaving");
  WHIR_REQUIRE(rate > 0.0 && rate <= 1.0, "irs: rate must be in (0, 1]");
  IrsCfg c;
  size_t message_length = vector_size / interleaving_depth;
  c.codeword_length = (size_t)ceil((double)message_length / rate);
  double r = (double)message_length / (double)c.codeword_length;
  c.johnson_slack = unique_decoding ? 0.0 : sqrt(r) / 20.0;
  if (!unique_decoding) {
    double list_size = 1.0 / (2.0 * c.johnson_slack * sqrt(r));
    c.out_domain_samples =
        num_ood_samples(false, security_target, field_size_bits(target_degree), list_size, vector_size);
  }
  double per_sample = unique_decoding ? (1.0 + r) / 2.0 : sqrt(r) + c.johnson_slack;
  c.in_domain_samples = (size_t)ceil(security_target / (-log2(per_sample)));
  c.num_vectors = num_vectors;
  c.vector_size = vector_size;
  c.interleaving_depth = interleaving_depth;
  c.target_degree = target_degree;
  return c;
}

double IrsCfg::rate() const { return (double)masked_message_length() / (double)codeword_length; }
bool IrsCfg::unique_decoding() const { return out_domain_samples == 0 && johnson_slack == 0.0; }
double IrsCfg::list_size() const {
  if (unique_decoding()) return 1.0;
  return 1.0 / (2.0 * johnson_slack * sqrt(rate()));
}
double IrsCfg::rbr_queries() const {
  double per_sample = unique_decoding() ? (1.0 + rate()) / 2.0 : sqrt(rate()) + johnson_slack;
  return (double)in_domain_samples * -log2(per_sample);
}
double IrsCfg::rbr_soundness_fold_prox_gaps() const {
  double log_field = field_size_bits(target_degree);
  double log_inv_rate = -log2(rate());
  double log_k = log2((double)masked_message_length());
  double error = unique_decoding() ? log_k + log_inv_rate : 7.0 * LOG2_10 + 3.5 * log_inv_rate + 2.0 * log_k;
  return log_field - error;
}

size_t SumcheckCfg::final_size() const {
  if (initial_size == 0 || num_rounds == 0) return initial_size;
  size_t p = 1;
  while (p < initial_size) p <<= 1;
  return p >> num_rounds;
}

WhirCfg WhirCfg::make(size_t size, const WhirParams& p) {
  // whir::Config::new, src/protocols/whir/config.rs:16-152
  WHIR_REQUIRE(size && (size & (size - 1)) == 0, "Only powers of two size are supported at the moment.");
  WHIR_REQUIRE(p.degree >= 1 && p.degree <= 3, "degree must be 1, 2 or 3");
  WHIR_REQUIRE(p.batch_size >= 1, "batch_size must be at least 1");
  WhirCfg c;
  c.D = p.degree;
  c.engine = p.hash_engine;
  const double security_level = (double)p.security_level;
  const double protocol_security_level = (double)std::max(p.security_level - p.pow_bits, 0);
  const double fsb = field_size_bits(p.degree);
  int log_inv_rate = p.starting_log_inv_rate;
  int num_variables = 0;
  while (((size_t)1 << num_variables) < size) num_variables++;
  WHIR_REQUIRE(num_variables >= p.initial_folding_factor, "folding factor larger than the number of variables");
  c.initial = IrsCfg::make(protocol_security_level, p.unique_decoding, p.batch_size, size,
                           (size_t)1 << p.initial_folding_factor, ldexp(1.0, -p.starting_log_inv_rate), p.degree);
  double prox = c.initial.rbr_soundness_fold_prox_gaps();
  double sumcheck_error = fsb - log2(c.initial.list_size()) - 1.0;
  double starting_folding_pow_bits = std::max(security_level - std::min(prox, sumcheck_error), 0.0);
  double initial_skip_pow_bits = std::max(security_level - (prox + log2((double)p.initial_folding_factor)), 0.0);
  int round = 0;
  size_t in_domain_samples = c.initial.in_domain_samples;
  double query_error = c.initial.rbr_queries();
  num_variables -= p.initial_folding_factor;
  while (num_variables >= p.folding_factor) {
    int rff = round == 0 ? p.initial_folding_factor : p.folding_factor;
    int next_rate = log_inv_rate + (rff - 1);
    RoundCfg rc;
    rc.irs = IrsCfg::make(protocol_security_level, p.unique_decoding, 1, (size_t)1 << num_variables,
                          (size_t)1 << p.folding_factor, ldexp(1.0, -next_rate), p.degree);
    double log_list = log2(rc.irs.list_size());
    size_t count = rc.irs.out_domain_samples + in_domain_samples;
    double combination_error = fsb - (log2((double)count) + log_list + 1.0);
    double pow_bits = std::max(0.0, security_level - std::min(query_error, combination_error));
    double fprox = rc.irs.rbr_soundness_fold_prox_gaps();
    double fsum = fsb - (log_list + 1.0);
    double folding_pow_bits = std::max(security_level - std::min(fprox, fsum), 0.0);
    rc.sumcheck.initial_size = (size_t)1 << num_variables;
    rc.sumcheck.round_pow = pow_threshold(folding_pow_bits);
    rc.sumcheck.num_rounds = (size_t)p.folding_factor;
    rc.pow = pow_threshold(pow_bits);
    round++;
    num_variables -= p.folding_factor;
    log_inv_rate = next_rate;
    in_domain_samples = rc.irs.in_domain_samples;
    query_error = rc.irs.rbr_queries();
    c.rounds.push_back(rc);
  }
  double rbr_error = c.rounds.empty() ? c.initial.rbr_queries() : c.rounds.back().irs.rbr_queries();
  double final_pow_bits = std::max(0.0, security_level - rbr_error);
  double final_folding_pow_bits = std::max(0.0, security_level - fsb + 1.0);
  c.initial_sumcheck.initial_size = size;
  c.initial_sumcheck.round_pow = pow_threshold(starting_folding_pow_bits);
  c.initial_sumcheck.num_rounds = (size_t)p.initial_folding_factor;
  c.initial_skip_pow = pow_threshold(initial_skip_pow_bits);
  c.final_sumcheck.initial_size = (size_t)1 << num_variables;
  c.final_sumcheck.round_pow = pow_threshold(final_folding_pow_bits);
  c.final_sumcheck.num_rounds = (size_t)num_variables;
  c.final_pow = pow_threshold(final_pow_bits);
  return c;
}

void WhirCfg::disable_pow() {
  initial_sumcheck.round_pow = UINT64_MAX;
  initial_skip_pow = UINT64_MAX;
  for (auto& r : rounds) {
    r.sumcheck.round_pow = UINT64_MAX;
    r.pow = UINT64_MAX;
  }
  final_sumcheck.round_pow = UINT64_MAX;
  final_pow = UINT64_MAX;
}

static void irs_json(std::ostringstream& o, const IrsCfg& c) {
  o << "{\"vector_size\": " << c.vector_size << ", \"codeword_length\": " << c.codeword_length
    << ", \"interleaving_depth\": " << c.interleaving_depth << ", \"num_vectors\": " << c.num_vectors
    << ", \"in_domain_samples\": " << c.in_domain_samples << ", \"out_domain_samples\": " << c.out_domain_samples
    << "}";
}
static void sc_json(std::ostringstream& o, const SumcheckCfg& s) {
  o << "[" << s.initial_size << ", " << s.num_rounds << ", " << s.round_pow << "]";
}

std::string WhirCfg::describe() const {
  std::ostringstream o;
  o << "{\"degree\": " << D << ", \"hash\": " << engine << ", \"initial\": ";
  irs_json(o, initial);
  o << ", \"initial_sumcheck\": ";
  sc_json(o, initial_sumcheck);
  o << ", \"initial_skip_pow\": " << initial_skip_pow << ", \"rounds\": [";
  for (size_t i = 0; i < rounds.size(); i++) {
    if (i) o << ", ";
    o << "{\"irs\": ";
    irs_json(o, rounds[i].irs);
    o << ", \"sumcheck\": ";
    sc_json(o, rounds[i].sumcheck);
    o << ", \"pow\": " << rounds[i].pow << "}";
  }
  o << "], \"final_sumcheck\": ";
  sc_json(o, final_sumcheck);
  o << ", \"final_pow\": " << final_pow << "}";
  return o.str();
}

}  // namespace whir
