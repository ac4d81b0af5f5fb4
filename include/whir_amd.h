/*
 * whir_amd.h -- C ABI of libwhir_amd, the MI355X (gfx950) WHIR prover hot path.
 *
 * Every entry point returns 0 on success and a nonzero status on failure; the
 * message is then available from whir_last_error() (per calling thread).  No
 * C++ exception crosses this boundary.  A binding turns a nonzero status into
 * the reference's panic (the reference asserts on contract violations, e.g.
 * src/algebra/ntt/cooley_tukey.rs:412-423).
 *
 * Field elements are canonical (non-Montgomery) little-endian uint64 words,
 * `degree` words per element, lowest coefficient first: degree 1 = Field64,
 * 2 = Field64_2 (u^2 = 7), 3 = Field64_3 (u^3 = 2) (src/algebra/fields.rs:85-143).
 * A binding over ark-ff converts with into_bigint/from_bigint at the boundary.
 *
 * Host-pointer entry points copy in, run on the GPU and copy out before they
 * return; nothing retains a host pointer.  `_dev` entry points take device
 * pointers (from whir_dev_alloc or any HIP allocation), run on the calling
 * thread's stream and return once the work is enqueued unless they return a
 * host value; whir_synchronize() waits for the calling thread's stream.
 * Every entry point is thread-safe; each host thread has its own stream.
 *
 * Hash engines (src/hash/mod.rs:28-36): 0 Copy, 1 Sha2 (SHA-256), 2 Keccak
 * (Keccak-256), 3 Sha3 (SHA3-256), 4 Blake3.
 */
#ifndef WHIR_AMD_H
#define WHIR_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  WHIR_HASH_COPY = 0,
  WHIR_HASH_SHA2 = 1,
  WHIR_HASH_KECCAK = 2,
  WHIR_HASH_SHA3 = 3,
  WHIR_HASH_BLAKE3 = 4
};

/* ----------------------------------------------------------------- runtime */

/* Message of the last failure on this thread ("" if none). */
const char *whir_last_error(void);
/* ABI version of this library. */
int whir_abi_version(void);
/* Number of visible GPUs (0 when no HIP device is present). */
int whir_device_count(int *count);
/* Select the device for the calling thread. */
int whir_set_device(int device);

int whir_dev_alloc(size_t bytes, void **ptr);
int whir_dev_free(void *ptr);
/* Pinned (page-locked) host memory, for host-memory inputs that overlap their upload
 * with device work (whir_commit_host). */
int whir_host_alloc(size_t bytes, void **ptr);
int whir_host_free(void *ptr);
int whir_memcpy_h2d(void *dst, const void *src, size_t bytes);
int whir_memcpy_d2h(void *dst, const void *src, size_t bytes);
int whir_memcpy_d2d(void *dst, const void *src, size_t bytes);
/* Wait for all work enqueued by the calling thread. */
int whir_synchronize(void);

/* Kernel timing with HIP events recorded on the launching stream (no reference
 * counterpart; the reference's `tracing` spans play this role).  When enabled,
 * every kernel family records launches, elapsed ms and algorithmic HBM bytes;
 * whir_profile_report writes them as JSON ({"name": {"launches", "total_ms",
 * "bytes"}}); pass buf = NULL to query the length. */
int whir_profile_enable(int on);
int whir_profile_reset(void);
int whir_profile_report(char *buf, size_t cap, size_t *len);

/* -------------------------------------------- ReedSolomon<F> (interleaved RS)
 * Replaces the trait ReedSolomon<F> (src/algebra/ntt/mod.rs:65-100) and the free
 * functions next_order / generator / evaluation_points / interleaved_rs_encode
 * (src/algebra/ntt/mod.rs:104-134), implemented by NttEngine
 * (src/algebra/ntt/cooley_tukey.rs:363-477). */

/* next_order (cooley_tukey.rs:364-368): smallest divisor of 3*2^32 >= size.
 * Returns a nonzero status when size exceeds the largest supported order. */
int whir_rs_next_order(uint64_t size, uint64_t *out);

/* generator (cooley_tukey.rs:370-373): 7^((p-1)/codeword_length) (base field). */
int whir_rs_generator(uint64_t codeword_length, uint64_t *out);

/* evaluation_points (cooley_tukey.rs:375-402), default (permuted) order:
 * point of index r = c*C + t is generator^(c + t*Q). */
int whir_rs_evaluation_points(uint64_t masked_message_length, uint64_t codeword_length,
                              const uint64_t *indices, size_t count, uint64_t *out);

/* interleaved_encode (cooley_tukey.rs:411-477): messages[j] points to
 * message_length elements; masks is num_messages x mask_length elements (may be
 * NULL when mask_length == 0); out receives codeword_length x num_messages
 * elements, row-major (row r holds the r-th evaluation of every message).
 * The GPU path needs power-of-two codeword_length and coset size. */
int whir_rs_interleaved_encode(int degree, const uint64_t *const *messages, size_t num_messages,
                               size_t message_length, const uint64_t *masks, size_t mask_length,
                               size_t codeword_length, uint64_t *out);

/* Device variant.  messages: num_messages contiguous messages of message_length
 * elements; out: codeword_length x num_messages elements, column-major
 * (out[j*codeword_length + r]) when row_major == 0, row-major otherwise. */
int whir_rs_interleaved_encode_dev(int degree, const uint64_t *messages, size_t num_messages,
                                   size_t message_length, const uint64_t *masks, size_t mask_length,
                                   size_t codeword_length, uint64_t *out, int row_major);

/* NTT schedule autotuner.  No counterpart in the reference (its NttEngine picks
 * one fixed recursion, cooley_tukey.rs:240-360); for callers that encode the same
 * shapes many times.  With tuning on, the first encode of each shape (degree,
 * log2 N, log2 coset size, number of messages) times every candidate schedule
 * (four-step split, per-pass LDS padding / radix-16 / tile size) on the caller's
 * stream and keeps the fastest for the rest of the process; outputs do not depend
 * on the schedule.  Pinned schedules are used whether or not tuning is on.
 * Variants are 2-bit masks (bit 1: LDS padding, bit 0: radix 16); tiles are u64
 * words of LDS (4096, or 8192 where the device allows it); log_c1 = 0 for shapes
 * with log2 C <= 11 (one pass). */
int whir_ntt_autotune(int enable);
int whir_ntt_set_schedule(int degree, size_t log_n, size_t log_c, size_t num_messages, int log_c1, int variant1,
                          int tile1, int variant2, int tile2);
int whir_ntt_clear_schedules(void);
/* JSON array of the schedule table.  Pass buf = NULL to query the length. */
int whir_ntt_schedules(char *buf, size_t cap, size_t *len);

/* ------------------------------------------------------------- HashEngine
 * Replaces HashEngine (src/hash/mod.rs:58-86) for the engines of the registry
 * (src/hash/mod.rs:28-36): Blake3 (src/hash/blake3_engine.rs:83-160), Sha2 / Sha3 /
 * Keccak (src/hash/digest_engine.rs:102-142), Copy (src/hash/copy_engine.rs). */

/* engine_id (src/hash/mod.rs:88-103): SHA3-256("whir::hash" || oid or name). */
int whir_hash_engine_id(int engine, uint8_t out[32]);
/* supports_size (src/hash/mod.rs:65-66).  *out = 1 if supported. */
int whir_hash_supports_size(int engine, size_t size, int *out);
/* preferred_batch_size (src/hash/mod.rs:67-77): messages per whir_hash_many call that
 * fill the current device in one launch (one message per lane: 1024 per CU). */
int whir_hash_preferred_batch_size(int engine, size_t *out);
/* hash_many (src/hash/mod.rs:79-85): `count` messages of `size` bytes, concatenated. */
int whir_hash_many(int engine, size_t size, const uint8_t *input, size_t count, uint8_t *output);
int whir_hash_many_dev(int engine, size_t size, const uint8_t *input, size_t count, uint8_t *output);

/* --------------------------------------------------- matrix / Merkle commit
 * Replace the bodies of matrix_commit::Config::commit (src/protocols/matrix_commit.rs:211-232,
 * leaf hashing hash_rows 301-368) and merkle_tree::Config::commit / open
 * (src/protocols/merkle_tree.rs:82-186). */

/* Number of hashes in the witness of a tree over num_leaves leaves:
 * 2^(layers_for_size(num_leaves)+1) - 1 (merkle_tree.rs:77-79, 281-283). */
int whir_merkle_num_nodes(size_t num_leaves, size_t *out);

/* merkle_tree::Config::commit: leaves = num_leaves x 32 bytes; nodes receives
 * whir_merkle_num_nodes hashes laid out [padded leaves][layer]...[root]. */
int whir_merkle_commit(int engine, const uint8_t *leaves, size_t num_leaves, uint8_t *nodes);
int whir_merkle_commit_dev(int engine, uint8_t *nodes, size_t num_leaves);

/* matrix_commit::Config::commit: hashes each row of a row-major rows x cols
 * matrix of degree-`degree` elements (ArkFieldEncoder bytes) and builds the tree. */
int whir_matrix_commit(int engine, int degree, const uint64_t *matrix, size_t rows, size_t cols,
                       uint8_t *nodes);

/* Leaf hashes of a device codeword in column-major layout (cw[j*rows + r]). */
int whir_hash_rows_cols_dev(int engine, int degree, const uint64_t *cw, size_t rows, size_t cols,
                            uint8_t *leaves);

/* matrix_commit::Config::commit (src/protocols/matrix_commit.rs:211-232) of a device
 * codeword in column-major layout: leaf hashes and every Merkle layer into the device
 * array nodes (whir_merkle_num_nodes(rows) hashes), as irs_commit.rs:292-350 commits
 * the interleaved codeword. */
int whir_matrix_commit_cols_dev(int engine, int degree, const uint64_t *cw, size_t rows, size_t cols,
                                uint8_t *nodes);

/* merkle_tree::Config::open (merkle_tree.rs:145-186): writes the sibling hints
 * (32 bytes each) in transcript order; *num_hints receives their count.  hints
 * must hold count * layers_for_size(num_leaves) hashes. */
int whir_merkle_open(const uint8_t *nodes, size_t num_leaves, const size_t *indices, size_t count,
                     uint8_t *hints, size_t *num_hints);

/* --------------------------------------------------- sumcheck and covectors
 * Replace src/algebra/sumcheck.rs:12-111, src/algebra/multilinear.rs:170-190 and
 * src/algebra/mod.rs:36-177.  Vectors hold n elements of `degree` words. */

/* compute_sumcheck_polynomial (sumcheck.rs:12-63): out = (c0, c2), 2*degree words. */
int whir_sumcheck_polynomial(int degree, const uint64_t *a, const uint64_t *b, size_t n, uint64_t *out);
int whir_sumcheck_polynomial_dev(int degree, const uint64_t *a, const uint64_t *b, size_t n, uint64_t *out);
/* fold (sumcheck.rs:69-104), in place; *new_len = next_pow2(n)/2 (n if n <= 1). */
int whir_fold(int degree, uint64_t *values, size_t n, const uint64_t *weight, size_t *new_len);
int whir_fold_dev(int degree, uint64_t *values, size_t n, const uint64_t *weight, size_t *new_len);
/* fold_and_compute_polynomial (sumcheck.rs:106-111), one fused pass. */
int whir_fold_and_compute_polynomial(int degree, uint64_t *a, uint64_t *b, size_t n, const uint64_t *weight,
                                     uint64_t *out, size_t *new_len);
int whir_fold_and_compute_polynomial_dev(int degree, uint64_t *a, uint64_t *b, size_t n,
                                         const uint64_t *weight, uint64_t *out, size_t *new_len);
/* Allocation-free variant of whir_fold_and_compute_polynomial_dev.  No counterpart
 * in the reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, results and errors; it calls no hipMalloc / hipFree / allocating
 * routine.  Its kernels keep their block partials and result in one static device
 * array per GPU (98,352 bytes: at most 2048 partials of 2 x 3 words plus the result),
 * held until the process ends; the partial count is capped by the kernel grid, so
 * every length up to 2^62 fits (longer ones are rejected).  Every call on a GPU
 * shares that array: calls that overlap on one GPU race. */
int whir_fold_and_compute_polynomial_dev_noalloc(int degree, uint64_t *a, uint64_t *b, size_t n,
                                                 const uint64_t *weight, uint64_t *out, size_t *new_len);
/* eval_eq (multilinear.rs:170-190): acc has 2^k elements, point k elements. */
int whir_eval_eq(int degree, uint64_t *acc, const uint64_t *point, size_t k, const uint64_t *scalar);
int whir_eval_eq_dev(int degree, uint64_t *acc, const uint64_t *point, size_t k, const uint64_t *scalar);
/* geometric_accumulate (algebra/mod.rs:154-177): acc[i] += sum_j scalars[j] points[j]^i. */
int whir_geometric_accumulate(int degree, uint64_t *acc, size_t n, const uint64_t *scalars,
                              const uint64_t *points, size_t m);
int whir_geometric_accumulate_dev(int degree, uint64_t *acc, size_t n, const uint64_t *scalars,
                                  const uint64_t *points, size_t m);
/* mixed_univariate_evaluate (algebra/mod.rs:105-129): coefficients of degree
 * src_degree (1 = base field, or = degree), point of `degree` words. */
int whir_mixed_univariate_evaluate(int src_degree, int degree, const uint64_t *coeffs, size_t n,
                                   const uint64_t *point, uint64_t *out);
int whir_mixed_univariate_evaluate_dev(int src_degree, int degree, const uint64_t *coeffs, size_t n,
                                       const uint64_t *point, uint64_t *out);
/* lift (algebra/mod.rs:51-59): base-field src -> degree-`degree` dst. */
int whir_lift(int degree, const uint64_t *src, size_t n, uint64_t *dst);
int whir_lift_dev(int degree, const uint64_t *src, size_t n, uint64_t *dst);
/* mixed_scalar_mul_add (algebra/mod.rs:88-98): acc[i] += weight * v[i], v base field. */
int whir_mixed_scalar_mul_add(int degree, uint64_t *acc, const uint64_t *weight, const uint64_t *v, size_t n);
int whir_mixed_scalar_mul_add_dev(int degree, uint64_t *acc, const uint64_t *weight, const uint64_t *v,
                                  size_t n);
/* dot (algebra/mod.rs:36). */
int whir_dot(int degree, const uint64_t *a, const uint64_t *b, size_t n, uint64_t *out);

/* ------------------------------------------------------------ proof of work
 * Replaces the grinding loop of proof_of_work::Config::prove
 * (src/protocols/proof_of_work.rs:86-163). */

/* threshold (proof_of_work.rs:29-39). */
int whir_pow_threshold(double difficulty_bits, uint64_t *out);
/* Minimal nonce with u64_le(H(challenge || nonce_le || 0^24)[0..8]) <= threshold.
 * threshold == UINT64_MAX is the reference's no-op: *nonce = 0, *performed = 0. */
int whir_pow_grind(int engine, const uint8_t challenge[32], uint64_t threshold, uint64_t *nonce,
                   int *performed);
/* Same search restricted to nonces below max_nonces (one bounded launch);
 * *found = 0 when none of them meets the threshold. */
int whir_pow_grind_bounded(int engine, const uint8_t challenge[32], uint64_t threshold, uint64_t max_nonces,
                           uint64_t *nonce, int *found);

/* ------------------------------------------------------------ WHIR prover
 * Replaces whir::Config::new (src/protocols/whir/config.rs:16-152),
 * whir::Config::commit (src/protocols/whir/mod.rs:79-92 -> irs_commit.rs:292-350)
 * and whir::Config::prove (src/protocols/whir/prover.rs:50-316) for base-field
 * vectors (Basefield<Field64_D> embedding, or Identity for degree 1) committed
 * batch_size at a time, one or more commitments proven together, and
 * MultilinearExtension linear forms (evaluation points). */

/* ProtocolParameters (src/parameters.rs:8-24) plus the field degree. */
typedef struct whir_params {
  int degree;                 /* 1, 2 or 3: Goldilocks1/2/3 (src/cmdline_utils.rs:7-15) */
  int security_level;
  int pow_bits;
  int initial_folding_factor;
  int folding_factor;
  int unique_decoding;
  int starting_log_inv_rate;
  int batch_size;             /* vectors per commitment (irs_commit.rs:58) */
  int hash_engine;            /* WHIR_HASH_* */
} whir_params;

typedef struct whir_config whir_config;
typedef struct whir_transcript whir_transcript;
typedef struct whir_witness whir_witness;

/* Config::new for a vector of `size` (a power of two) coefficients. */
int whir_config_new(const whir_params *params, size_t size, whir_config **out);
/* Sets every PoW threshold to u64::MAX (the reference's disable_pow, whir/mod.rs:107-118). */
int whir_config_disable_pow(whir_config *cfg);
/* JSON description of the derived round shapes, samples and PoW thresholds. */
int whir_config_describe(const whir_config *cfg, char *buf, size_t cap, size_t *len);
void whir_config_free(whir_config *cfg);

/* Typed transcript operations (ProverState, src/transcript/mod.rs:183-267); each
 * callback returns 0 on success.  A Rust binding implements them over spongefish. */
typedef struct whir_transcript_vtable {
  void *ctx;
  int (*prover_field)(void *ctx, int degree, const uint64_t *elems, size_t count);
  int (*verifier_field)(void *ctx, int degree, uint64_t *out);
  int (*verifier_bytes)(void *ctx, uint8_t *out, size_t count);
  int (*prover_hash)(void *ctx, const uint8_t *hash32);
  int (*prover_u64)(void *ctx, uint64_t value);
  int (*hint_hash)(void *ctx, const uint8_t *hash32);
  int (*hint_field_vec)(void *ctx, int degree, const uint64_t *elems, size_t count);
} whir_transcript_vtable;

/* Built-in transcript: Keccak-f[1600] duplex keyed by a 64-byte protocol id and a
 * 32-byte session id (DomainSeparator, src/transcript/mod.rs:110-137). */
int whir_transcript_new(const uint8_t protocol_id[64], const uint8_t session_id[32], const uint8_t *instance,
                        size_t instance_len, whir_transcript **out);
int whir_transcript_from_vtable(const whir_transcript_vtable *vtable, whir_transcript **out);
/* Proof bytes of a built-in transcript (Proof{narg_string, hints}, transcript/mod.rs:63-70).
 * Pass NULL buffers to query the lengths. */
int whir_transcript_proof(const whir_transcript *t, uint8_t *narg, size_t narg_cap, size_t *narg_len,
                          uint8_t *hints, size_t hints_cap, size_t *hints_len);
void whir_transcript_free(whir_transcript *t);

/* whir::Config::commit of a device-resident base-field vector of `size` words.
 * The witness (codeword and Merkle tree) stays on the device. */
int whir_commit_dev(const whir_config *cfg, whir_transcript *t, const uint64_t *vector, whir_witness **out);
/* whir::Config::commit of a vector in host memory (irs_commit::Config::commit takes
 * host slices, src/protocols/irs_commit.rs:292-350): the coefficients are uploaded into
 * dev_vector (vector size words, allocated by the caller and kept for whir_prove_dev)
 * a chunk of messages at a time on a second stream, each chunk's Reed-Solomon encode
 * starting as soon as its copy has landed.  host_vector in pinned memory
 * (whir_host_alloc) overlaps best.  Same witness and transcript as uploading the
 * vector and calling whir_commit_dev. */
int whir_commit_host(const whir_config *cfg, whir_transcript *t, const uint64_t *host_vector, uint64_t *dev_vector,
                     whir_witness **out);
/* whir::Config::commit of `count` independent instances (each its own configuration
 * with batch_size 1, transcript and device vector), with results identical to `count`
 * calls of whir_commit_dev.  No counterpart in the reference (which commits one
 * instance per call); for callers that commit many instances.  The instances are
 * grouped by shape and each group's Reed-Solomon encodes, leaf hashes and Merkle levels
 * run as shared launches over device tables of the instances' buffers; the launch
 * sequence of a batch layout (the list of shapes) is recorded as a HIP graph the first
 * time and replayed afterwards.  The calling thread keeps the last few layouts' plans
 * (tables, NTT scratch of the largest group) until it ends.  out receives count
 * witnesses (free each with whir_witness_free). */
int whir_commit_many_dev(const whir_config *const *cfgs, whir_transcript *const *transcripts,
                         const uint64_t *const *vectors, size_t count, whir_witness **out);
/* whir::Config::prove: `points` holds num_points points of num_variables
 * elements (degree words each), `evaluations` the claimed MLE values.
 * evaluation_point (optional) receives the FinalClaim point (num_variables elements). */
int whir_prove_dev(const whir_config *cfg, whir_transcript *t, const uint64_t *vector, const whir_witness *w,
                   const uint64_t *points, size_t num_points, const uint64_t *evaluations,
                   uint64_t *evaluation_point);
/* whir::Config::commit of batch_size device-resident base-field vectors in one
 * commitment (irs_commit.rs:292-350: vector b fills messages [b*2^k, (b+1)*2^k)
 * of the committed matrix, OOD values point-major).  whir_commit_dev is the
 * batch of one. */
int whir_commit_batch_dev(const whir_config *cfg, whir_transcript *t, const uint64_t *const *vectors,
                          size_t num_vectors, whir_witness **out);
/* whir::Config::prove over num_witnesses commitments holding num_vectors =
 * num_witnesses * batch_size vectors, in commitment order (prover.rs:50-316:
 * OOD cross terms, vector RLC, all initial witnesses opened at the first round).
 * evaluations: num_points x num_vectors values (row-major), vector j at point k
 * at (k * num_vectors + j) * degree.  whir_prove_dev is the case of one vector. */
int whir_prove_batch_dev(const whir_config *cfg, whir_transcript *t, const uint64_t *const *vectors,
                         size_t num_vectors, const whir_witness *const *witnesses, size_t num_witnesses,
                         const uint64_t *points, size_t num_points, const uint64_t *evaluations,
                         uint64_t *evaluation_point);
/* Linear forms of whir::Config::prove (src/algebra/linear_form/): a
 * MultilinearExtension at a point (multilinear_extension.rs), an explicit Covector
 * (covector.rs) or a UnivariateEvaluation at a point (univariate_evaluation.rs). */
#define WHIR_FORM_MLE 0
#define WHIR_FORM_COVECTOR 1
#define WHIR_FORM_UNIVARIATE 2
/* whir_prove_batch_dev with linear forms of any kind, in the order the reference
 * receives them (kinds[k] for form k; kinds NULL = all MLE).  `points` holds, in form
 * order, num_variables elements per MLE form and one element per univariate form;
 * each Covector form takes the next device vector of `covectors` (vector_size
 * elements of degree words).  evaluations: num_forms x num_vectors values. */
int whir_prove_forms_dev(const whir_config *cfg, whir_transcript *t, const uint64_t *const *vectors,
                         size_t num_vectors, const whir_witness *const *witnesses, size_t num_witnesses,
                         const int *kinds, size_t num_forms, const uint64_t *points,
                         const uint64_t *const *covectors, const uint64_t *evaluations, uint64_t *evaluation_point);
void whir_witness_free(whir_witness *w);
/* MultilinearExtension::evaluate (mixed_multilinear_extend, src/algebra/multilinear.rs:15-151)
 * of a device base-field vector of 2^num_variables words at a degree-`degree` point. */
int whir_mle_evaluate_dev(int degree, const uint64_t *vector, size_t num_variables, const uint64_t *point,
                          uint64_t *out);

/* ---- Batched and served proving.  No counterpart in the reference, which proves
 * one instance per call (src/bin/main.rs:152-190); for callers that prove many
 * instances or serve a stream of requests.  A pool of worker threads, each on its
 * own HIP stream, runs whir_commit_dev + whir_prove_dev for queued jobs, starting
 * the first queued job whose estimated peak device memory fits what the budget has
 * left.  A job's proof is byte-identical to the same job run alone. */
typedef struct whir_executor whir_executor;
/* Estimated peak device bytes of one commit + prove under this configuration. */
int whir_config_memory(const whir_config *cfg, size_t *bytes);
/* hbm_budget 0: 80% of the device memory free at creation. */
int whir_executor_new(int num_workers, size_t hbm_budget, whir_executor **out);
/* Waits for every queued and running job, then stops the workers. */
void whir_executor_free(whir_executor *e);
/* kind 0: commit + prove (points and evaluations as for whir_prove_dev); kind 1:
 * commit only.  cfg, t and the device vector must stay valid until the job has been
 * waited for; points and evaluations are copied. */
int whir_executor_submit(whir_executor *e, int kind, const whir_config *cfg, whir_transcript *t,
                         const uint64_t *vector, const uint64_t *points, size_t num_points,
                         const uint64_t *evaluations, uint64_t *job_id);
/* A job proving num_vectors vectors (a multiple of batch_size) together: they are
 * committed batch_size at a time in order (whir_commit_batch_dev), then proven with
 * whir_prove_batch_dev; evaluations are num_points x num_vectors values.  The
 * device vectors must stay valid until the job has been waited for. */
int whir_executor_submit_batch(whir_executor *e, const whir_config *cfg, whir_transcript *t,
                               const uint64_t *const *vectors, size_t num_vectors, const uint64_t *points,
                               size_t num_points, const uint64_t *evaluations, uint64_t *job_id);
/* kind 2 (prove only): whir_prove_dev for a vector already committed to in `t`
 * (witness from whir_commit_dev or whir_commit_many_dev; it must stay valid until the
 * job has been waited for). */
int whir_executor_submit_prove(whir_executor *e, const whir_config *cfg, whir_transcript *t,
                               const uint64_t *vector, const whir_witness *witness, const uint64_t *points,
                               size_t num_points, const uint64_t *evaluations, uint64_t *job_id);
/* Blocks until the job has finished and returns its status.  times (optional)
 * receives {submitted, started, finished} on the executor's clock (seconds);
 * eval_point (optional, num_variables * degree words) the folding randomness. */
int whir_executor_wait(whir_executor *e, uint64_t job_id, double times[3], uint64_t *eval_point);
/* Seconds since the executor was created (the clock of whir_executor_wait). */
int whir_executor_clock(whir_executor *e, double *seconds);

#ifdef __cplusplus
}
#endif

#endif /* WHIR_AMD_H */
