// WHIR prover orchestration on the GPU (host control, device data).
//
// Restates the transcript-ordered control flow of the reference prover:
//   commit   whir::Config::commit -> irs_commit::Config::commit
//            (src/protocols/whir/mod.rs:79-92, src/protocols/irs_commit.rs:292-350)
//   prove    whir::Config::prove (src/protocols/whir/prover.rs:50-316)
//   open     irs_commit::Config::open (irs_commit.rs:385-432), merkle_tree::Config::open
//            (src/protocols/merkle_tree.rs:145-186), challenge_indices
//            (src/protocols/challenge_indices.rs:6-50)
//   sumcheck sumcheck::Config::prove without masks (src/protocols/sumcheck.rs:69-162)
//   pow      proof_of_work::Config::prove (src/protocols/proof_of_work.rs:68-166)
// for base-field vectors committed batch_size at a time (one or more commitments
// proven together) and MultilinearExtension linear forms.
//
// Every vector of the protocol (lifted vector, covector, codewords, Merkle trees)
// lives in HBM for the whole proof; the host only sees the transcript values:
// roots, OOD and sumcheck messages, opened rows and sibling hashes.
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <functional>
#include <memory>
#include <vector>

#include "../../include/whir_amd.h"
#include "algebra.h"
#include "common.h"
#include "config.h"
#include "gl.h"
#include "hash.h"
#include "ntt.h"
#include "transcript.h"
#include "handles.h"

using namespace whir;


namespace whir {

struct DevWitness {
  int Ds = 1;               // degree of the committed elements
  size_t N = 0, M = 0;      // codeword length, number of messages (columns)
  size_t num_vectors = 1;   // vectors in this commitment (M = num_vectors * interleaving depth)
  DevBuf cw;                // column-major codeword cw[j*N + r]
  DevBuf nodes;             // Merkle nodes [padded leaves][layer]...[root]
  // D words each; ood_values is point-major (out_domain_samples x num_vectors)
  std::vector<uint64_t> ood_points, ood_values;
};

}  // namespace whir

struct whir_witness {
  DevWitness w;
};

namespace whir {
namespace {

template <int D>
using E = gl::Ext<D>;

template <int D>
E<D> ld(const uint64_t* p) {
  E<D> r;
  for (int d = 0; d < D; d++) r.c[d] = p[d];
  return r;
}
template <int D>
void push(std::vector<uint64_t>& v, const E<D>& e) {
  for (int d = 0; d < D; d++) v.push_back(e.c[d]);
}

template <int D>
E<D> verifier_field(Transcript& t) {
  E<D> r;
  t.verifier_field(D, r.c);
  return r;
}

// geometric_challenge, src/protocols/geometric_challenge.rs:10-24
template <int D>
std::vector<E<D>> geometric_challenge(Transcript& t, size_t count) {
  std::vector<E<D>> out;
  if (count == 0) return out;
  if (count == 1) {
    out.push_back(gl::ext_base<D>(1));
    return out;
  }
  E<D> x = verifier_field<D>(t);
  E<D> cur = gl::ext_base<D>(1);
  for (size_t i = 0; i < count; i++) {
    out.push_back(cur);
    cur = gl::emul<D>(cur, x);
  }
  return out;
}

// challenge_indices, src/protocols/challenge_indices.rs:6-50
std::vector<size_t> challenge_indices(Transcript& t, size_t num_leaves, size_t count, bool dedup) {
  std::vector<size_t> idx;
  if (count == 0) return idx;
  WHIR_REQUIRE(num_leaves && (num_leaves & (num_leaves - 1)) == 0,
               "Number of leaves must be a power of two for unbiased results.");
  if (num_leaves == 1) return dedup ? std::vector<size_t>{0} : std::vector<size_t>(count, 0);
  size_t lg = 0;
  while (((size_t)1 << lg) < num_leaves) lg++;
  size_t size_bytes = (lg + 7) / 8;
  std::vector<uint8_t> entropy(count * size_bytes);
  for (auto& b : entropy) t.verifier_bytes(&b, 1);
  for (size_t i = 0; i < count; i++) {
    size_t acc = 0;
    for (size_t k = 0; k < size_bytes; k++) acc = (acc << 8) | entropy[i * size_bytes + k];
    idx.push_back(acc % num_leaves);
  }
  if (dedup) {
    std::sort(idx.begin(), idx.end());
    idx.erase(std::unique(idx.begin(), idx.end()), idx.end());
  }
  return idx;
}

// proof_of_work::Config::prove, src/protocols/proof_of_work.rs:68-166
void pow_prove(Transcript& t, int engine, uint64_t threshold, hipStream_t s) {
  if (threshold == UINT64_MAX) return;
  uint8_t ch[32];
  t.verifier_bytes(ch, 32);
  const uint64_t window = 1ull << 32;
  uint64_t nonce = 0;
  for (uint64_t first = 0;; first += window) {
    if (pow_grind_dev(engine, ch, threshold, first, window, &nonce, s)) break;
    WHIR_REQUIRE(first <= UINT64_MAX - 2 * window, "Proof of Work failed to solve.");
  }
  t.prover_u64(nonce);
}

// Covector views (see CovView) for initial covectors of at most this many terms;
// WHIR_COV_VIEW=0 always materializes the covector (for comparison).
constexpr size_t COV_VIEW_MAX_TERMS = 4;
static bool cov_view_enabled() {
  static const bool on = [] {
    const char* e = getenv("WHIR_COV_VIEW");
    return !e || atoi(e) != 0;
  }();
  return on;
}

int log2_exact(uint64_t n) {
  int k = 0;
  while ((1ull << k) < n) k++;
  return (1ull << k) == n ? k : -1;
}

uint64_t coset_size(uint64_t masked_len, uint64_t N) {
  // cooley_tukey.rs:436-439 restricted to the power-of-two codewords WHIR uses
  uint64_t c = 1;
  while (c < masked_len) c <<= 1;
  WHIR_REQUIRE(N % c == 0, "coset size must divide the codeword length");
  return c;
}

// irs_commit::Config::commit, src/protocols/irs_commit.rs:292-350 (no masks):
// vecs[0..nvec) each hold vector_size elements of Ds words on the device; vector
// b fills messages (columns) [b*depth, (b+1)*depth) of the committed matrix.
// The host side of a commit whose vector arrives from host memory: the coefficients go
// up a chunk of messages at a time on a copy stream, and each chunk's encode starts as
// soon as its copy has landed (the encode of message j needs message j only), so the
// upload overlaps the encodes of the earlier chunks.
struct HostUpload {
  const uint64_t* host;  // the vector in host memory (pinned memory overlaps best)
  uint64_t* dev;         // its device copy, filled by the upload
};

// A second stream per thread for host->device copies (default priority).
thread_local struct CopyStream {
  hipStream_t s = nullptr;
  std::vector<hipEvent_t> ev;
  ~CopyStream() { release(); }
  void release() {
    for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    ev.clear();
    if (s) (void)hipStreamDestroy(s);
    s = nullptr;
  }
  hipStream_t get() {
    if (!s) WHIR_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    return s;
  }
  hipEvent_t event(size_t i) {
    while (ev.size() <= i) {
      hipEvent_t e = nullptr;
      WHIR_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
      ev.push_back(e);
    }
    return ev[i];
  }
} t_copy;

// Messages per uploaded chunk of a host-memory commit (4 chunks of 4 messages at the
// usual interleaving depth 16).
constexpr size_t UPLOAD_CHUNK_MESSAGES = 4;

// Out-of-domain evaluations of at most this many coefficient-point products run on the
// host (Horner's rule over coefficients copied to pinned host memory).
constexpr size_t HOST_OOD_TERMS = 512;

// Pinned host memory for short vectors that the host works on (out-of-domain samples and
// short sumchecks read them here, and folded vectors go back from here), one buffer per
// host thread (each worker thread drives its own stream).  It is not the result mailbox:
// results queued ahead (an opening's rows) may still be there.  Copies into it are
// followed by a mailbox wait; a later use on the thread rewrites it only after such a
// wait, by which time the stream has run the copies that read it.
constexpr size_t HOST_STAGE_WORDS = 2048;
struct HostStage {
  uint64_t* p = nullptr;
  ~HostStage() {
    if (p) (void)hipHostFree(p);
  }
};
static uint64_t* host_stage() {
  thread_local HostStage st;
  if (!st.p) WHIR_HIP(hipHostMalloc((void**)&st.p, HOST_STAGE_WORDS * 8, hipHostMallocDefault));
  return st.p;
}

template <int D>
void irs_commit(Transcript& t, const IrsCfg& c, int engine, const uint64_t* const* vecs, size_t nvec, int Ds,
                DevWitness& w, hipStream_t s, const HostUpload* up = nullptr) {
  WHIR_REQUIRE(nvec == c.num_vectors, "irs commit: the configuration commits to a different number of vectors");
  const size_t depth = c.interleaving_depth, M = c.num_messages(), L = c.message_length(), N = c.codeword_length;
  uint64_t C = coset_size(L, N);
  w.Ds = Ds;
  w.N = N;
  w.M = M;
  w.num_vectors = nvec;
  w.cw = DevBuf(M * N * Ds * 8, s);
  if (up) {
    WHIR_REQUIRE(nvec == 1 && Ds == 1, "irs commit: host upload of one base-field vector only");
    const size_t G = std::min(UPLOAD_CHUNK_MESSAGES, depth);
    DevBuf scratch(rs_scratch_words(1, G, N, C) * 8, s);
    const hipStream_t cs = t_copy.get();
    // the copy stream must not overwrite the device vector before earlier work on s is done
    WHIR_HIP(hipEventRecord(t_copy.event(0), s));
    WHIR_HIP(hipStreamWaitEvent(cs, t_copy.event(0), 0));
    for (size_t j0 = 0, k = 1; j0 < depth; j0 += G, k++) {
      const size_t g = std::min(G, depth - j0);
      WHIR_HIP(hipMemcpyAsync(up->dev + j0 * L, up->host + j0 * L, g * L * 8, hipMemcpyHostToDevice, cs));
      WHIR_HIP(hipEventRecord(t_copy.event(k), cs));
      WHIR_HIP(hipStreamWaitEvent(s, t_copy.event(k), 0));
      RsArgs a{up->dev + j0 * L, L, L, nullptr, 0, w.cw.as<uint64_t>() + j0 * N, N, N / C, log2_exact(C), g};
      rs_encode_cols(1, a, scratch.as<uint64_t>(), s);
    }
  } else {
    DevBuf scratch(rs_scratch_words(Ds, depth, N, C) * 8, s);
    for (size_t b = 0; b < nvec; b++) {
      RsArgs a{vecs[b], L, L, nullptr, 0, w.cw.as<uint64_t>() + b * depth * N * Ds, N, N / C, log2_exact(C), depth};
      rs_encode_cols(Ds, a, scratch.as<uint64_t>(), s);
    }
  }
  size_t nn = merkle_num_nodes(N);
  w.nodes = DevBuf(32 * nn, s);
  uint64_t* mbox = result_mailbox(s);
  uint8_t root[32];
  const bool const_rows = L == 1 && (N & (N - 1)) == 0;  // every column is a constant: all rows are equal
  if (const_rows && const_rows_on_host(engine, 8 * M * Ds, N)) {
    commit_const_rows_host(Ds, w.cw.as<uint64_t>(), N, M, w.nodes.as<uint8_t>(), s, root);
  } else if (!const_rows && host_top_ok(engine)) {
    commit_cols_host_top(engine, Ds, w.cw.as<uint64_t>(), N, M, w.nodes.as<uint8_t>(), s, root);
  } else {
    // The root lands in the stream's pinned mailbox from the last tree launch.
    mailbox_arm(s);
    if (const_rows)
      commit_const_rows_dev(engine, Ds, w.cw.as<uint64_t>(), N, M, w.nodes.as<uint8_t>(), s, (uint8_t*)mbox);
    else
      commit_cols_dev(engine, Ds, w.cw.as<uint64_t>(), N, M, w.nodes.as<uint8_t>(), s, (uint8_t*)mbox);
    mailbox_wait(s);
    memcpy(root, mbox, 32);
  }
  t.prover_hash(root);
  w.ood_points.clear();
  w.ood_values.clear();
  for (size_t i = 0; i < c.out_domain_samples; i++) push<D>(w.ood_points, verifier_field<D>(t));
  const size_t nood = c.out_domain_samples;
  if (nood && nvec * nood * c.vector_size <= HOST_OOD_TERMS && nvec * c.vector_size * Ds <= HOST_STAGE_WORDS) {
    // Short vectors: their coefficients come to the host (one wait, which the values
    // needed anyway) and Horner's rule runs there, instead of an evaluation launch that
    // one block would spend ~15 us on.
    const size_t n = c.vector_size;
    uint64_t* stage = host_stage();
    mailbox_arm(s);
    for (size_t b = 0; b < nvec; b++)
      WHIR_HIP(hipMemcpyAsync(stage + b * n * Ds, vecs[b], n * Ds * 8, hipMemcpyDefault, s));
    mailbox_wait(s);
    w.ood_values.resize(nood * nvec * D);
    for (size_t i = 0; i < nood; i++) {
      gl::Ext<D> x;
      for (int d = 0; d < D; d++) x.c[d] = w.ood_points[i * D + d];
      for (size_t b = 0; b < nvec; b++) {
        const uint64_t* cf = stage + b * n * Ds;
        gl::Ext<D> v = gl::ext_zero<D>();
        for (size_t k = n; k-- > 0;) {
          gl::Ext<D> ck;
          if (Ds == 1) ck = gl::ext_base<D>(cf[k]);
          else for (int d = 0; d < D; d++) ck.c[d] = cf[k * Ds + d];
          v = gl::eadd<D>(gl::emul<D>(v, x), ck);
        }
        std::copy(v.c, v.c + D, &w.ood_values[(i * nvec + b) * D]);
        t.prover_field(D, &w.ood_values[(i * nvec + b) * D], 1);
      }
    }
  } else if (nood) {
    // All samples of a vector in shared passes over it, written to the mailbox
    // (vector b at words [b*pad*D, (b+1)*pad*D)); sent point-major.
    const size_t pad = (nood + 3) / 4 * 4;
    WHIR_REQUIRE(nvec * pad * D <= RESULT_MAILBOX_WORDS, "too many out-of-domain samples");
    for (size_t b = 0; b < nvec; b++) {
      if (b + 1 == nvec) mailbox_arm(s);
      univariate_eval_many_dev(Ds, D, vecs[b], c.vector_size, w.ood_points.data(), nood, mbox + b * pad * D, s,
                               true);
    }
    mailbox_wait(s);
    w.ood_values.resize(nood * nvec * D);
    for (size_t i = 0; i < nood; i++)
      for (size_t b = 0; b < nvec; b++) {
        std::copy(mbox + (b * pad + i) * D, mbox + (b * pad + i + 1) * D, &w.ood_values[(i * nvec + b) * D]);
        t.prover_field(D, &w.ood_values[(i * nvec + b) * D], 1);
      }
  }
}

// in_domain_challenges (irs_commit.rs:486-500): indices and their evaluation points.
// exps (optional) receives each point's exponent e (the point is g^e, g of order N).
void in_domain(Transcript& t, const IrsCfg& c, std::vector<size_t>& idx, std::vector<uint64_t>& pts,
               std::vector<uint64_t>* exps = nullptr) {
  idx = challenge_indices(t, c.codeword_length, c.in_domain_samples, c.deduplicate_in_domain);
  const size_t N = c.codeword_length;
  uint64_t C = coset_size(c.masked_message_length(), N), Q = N / C;
  uint64_t g = gl::root_of_unity(N);
  // g^e for every index (e < N, transpose_permute order) from a table of g^(d 16^w),
  // d < 16, one row per 4-bit window of e: 15 multiplies per window to build and one
  // per window per point, against a square-and-multiply per point (these run on the
  // host between two GPU steps of the proof).
  int bits = 0;
  while (((uint64_t)1 << bits) < N) bits++;
  const int windows = std::max(1, (bits + 3) / 4);
  std::vector<uint64_t> tab((size_t)windows * 16);
  uint64_t base = g;
  for (int w = 0; w < windows; w++) {
    uint64_t* row = &tab[(size_t)w * 16];
    row[0] = 1;
    row[1] = base;
    for (int d = 2; d < 16; d++) row[d] = gl::mul(row[d - 1], base);
    base = gl::mul(row[15], base);
  }
  pts.clear();
  if (exps) exps->clear();
  for (size_t i : idx) {
    const uint64_t e = (i / C) + (i % C) * Q;
    if (exps) exps->push_back(e);
    uint64_t v = 1;
    for (int w = 0; w < windows; w++) {
      const unsigned d = (unsigned)(e >> (4 * w)) & 15;
      if (d) v = v == 1 ? tab[(size_t)w * 16 + d] : gl::mul(v, tab[(size_t)w * 16 + d]);
    }
    pts.push_back(v);
  }
}

// The per-witness part of irs_commit::Config::open (irs_commit.rs:410-428): the rows at
// idx and their Merkle path (merkle_tree.rs:145-186), gathered by one launch, and with
// D > 0 each row's sum against the M column weights wts (D words each).  The launch
// and the read-back are split so that a caller can queue more work (the covector and
// the next sumcheck pass) behind the gather and synchronize once: open_rows_launch
// queues the gather with its results in the stream's mailbox from word `off` (in a
// device buffer when they do not fit), open_rows_finish, after the caller has
// synchronized the stream, copies them out and hints rows (k x M elements of Ds words)
// and hashes.  Hints do not enter the sponge, so hinting after later challenges were
// drawn leaves the transcript as it is.
struct OpenLaunch {
  const DevWitness* w = nullptr;
  size_t k = 0, M = 0, ns = 0, row_words = 0, hash_off = 0, val_off = 0, out_words = 0, off = 0;
  int D = 0;
  bool in_mailbox = true;
  std::shared_ptr<DevBuf> dout;
};

OpenLaunch open_rows_launch(const DevWitness& w, const std::vector<size_t>& idx, hipStream_t s, int D,
                            const uint64_t* wts, size_t off) {
  OpenLaunch L;
  const size_t N = w.N, M = w.M, k = idx.size();
  // Sibling node indices in hint order (they depend on the indices alone, so rows
  // and siblings are gathered in one launch).
  std::vector<size_t> cur(idx);
  std::sort(cur.begin(), cur.end());
  cur.erase(std::unique(cur.begin(), cur.end()), cur.end());
  std::vector<uint64_t> idx_sib(idx.begin(), idx.end());
  size_t layers = merkle_layers(N), noff = 0, width = (size_t)1 << layers;
  while (width > 1) {
    std::vector<size_t> next;
    for (size_t q = 0; q < cur.size();) {
      size_t a = cur[q];
      if (q + 1 < cur.size() && cur[q + 1] == (a ^ 1)) {
        q += 2;
      } else {
        idx_sib.push_back(noff + (a ^ 1));
        q += 1;
      }
      next.push_back(a >> 1);
    }
    cur.swap(next);
    noff += width;
    width >>= 1;
  }
  L.w = &w;
  L.k = k;
  L.M = M;
  L.D = D;
  L.ns = idx_sib.size() - k;
  L.row_words = k * M * w.Ds;
  L.hash_off = (L.row_words + 1) & ~(size_t)1;  // hashes 16-byte aligned
  L.val_off = L.hash_off + 4 * L.ns;
  L.out_words = L.val_off + (D > 0 ? k * D : 0);
  L.off = (off + 1) & ~(size_t)1;
  if (!(k + L.ns)) return L;
  // one upload: the row and node indices, then the column weights
  const size_t nidx = idx_sib.size();
  if (D > 0) idx_sib.insert(idx_sib.end(), wts, wts + M * D);
  DevBuf dis(idx_sib.size() * 8, s);
  upload_async(dis.ptr, idx_sib.data(), idx_sib.size() * 8, s);
  const uint64_t* dw = D > 0 ? dis.as<uint64_t>() + nidx : nullptr;
  uint64_t* out;
  // WHIR_OPEN_MAILBOX_WORDS lowers the mailbox room openings may use (a test knob that
  // sends them through the device-buffer path).
  size_t room = RESULT_MAILBOX_WORDS;
  if (const char* e = getenv("WHIR_OPEN_MAILBOX_WORDS")) room = std::min<size_t>(room, (size_t)atoll(e));
  if (L.off + L.out_words <= room) {
    out = result_mailbox(s) + L.off;
  } else {
    L.in_mailbox = false;
    L.dout = std::make_shared<DevBuf>(L.out_words * 8, s);
    out = L.dout->as<uint64_t>();
  }
  gather_open_dev(w.Ds, w.cw.as<uint64_t>(), N, M, dis.as<uint64_t>(), k, w.nodes.as<uint8_t>(), L.ns, out,
                  (uint8_t*)(out + L.hash_off), s, D, dw, out + L.val_off);
  return L;
}

void open_rows_finish(Transcript& t, OpenLaunch& L, std::vector<uint64_t>& rows, uint64_t* vals, hipStream_t s) {
  rows.assign(L.row_words, 0);
  std::vector<uint8_t> hashes(32 * L.ns);
  if (L.D > 0) std::fill(vals, vals + L.k * L.D, 0);
  if (L.k + L.ns) {
    const uint64_t* res;
    std::vector<uint64_t> h;
    if (L.in_mailbox) {
      res = result_mailbox(s) + L.off;
    } else {
      h.resize(L.out_words);
      download_sync(h.data(), L.dout->ptr, h.size() * 8, s);
      res = h.data();
      L.dout.reset();
    }
    std::copy(res, res + L.row_words, rows.begin());
    memcpy(hashes.data(), res + L.hash_off, hashes.size());
    if (L.D > 0) std::copy(res + L.val_off, res + L.val_off + L.k * L.D, vals);
  }
  t.hint_field_vec(L.w->Ds, rows.data(), L.k * L.M);
  t.hint_hashes(hashes.data(), L.ns);
}

// Mailbox words below this offset are left to the sumcheck passes.
constexpr size_t OPEN_MAILBOX_OFF = 64;

// irs_commit::Config::open (irs_commit.rs:385-432) over one or more witnesses:
// one set of in-domain indices, then each witness's rows and Merkle path.  Launch:
// draws the indices, computes their points and queues every witness's gather.
std::vector<OpenLaunch> irs_open_launch(Transcript& t, const IrsCfg& c, const std::vector<const DevWitness*>& ws,
                                        std::vector<uint64_t>& pts, hipStream_t s, int D = 0,
                                        const std::vector<std::vector<uint64_t>>* wts = nullptr,
                                        std::vector<uint64_t>* exps = nullptr) {
  std::vector<size_t> idx;
  in_domain(t, c, idx, pts, exps);
  std::vector<OpenLaunch> ls;
  size_t off = OPEN_MAILBOX_OFF;
  for (size_t w = 0; w < ws.size(); w++) {
    ls.push_back(open_rows_launch(*ws[w], idx, s, D, D > 0 ? (*wts)[w].data() : nullptr, off));
    if (ls.back().in_mailbox) off = ls.back().off + ls.back().out_words;
  }
  return ls;
}

// Finish (after the stream was synchronized): rows[w] holds witness w's opened rows;
// with D > 0, vals[w] (k x D words) their sums against the column weights.
void irs_open_finish(Transcript& t, std::vector<OpenLaunch>& ls, std::vector<std::vector<uint64_t>>& rows,
                     std::vector<std::vector<uint64_t>>& vals, hipStream_t s) {
  rows.resize(ls.size());
  vals.assign(ls.size(), std::vector<uint64_t>());
  for (size_t w = 0; w < ls.size(); w++) {
    if (ls[w].D > 0) vals[w].resize(ls[w].k * ls[w].D);
    open_rows_finish(t, ls[w], rows[w], ls[w].D > 0 ? vals[w].data() : nullptr, s);
  }
}

// sumcheck::Config::prove (no masks), staggered fold + round polynomial.  The
// first operand starts at a_src with a_Ds words per element (1 for the initial
// round over the committed base-field vector, which is folded into `a` without
// being lifted first); from the first fold on it lives in `a`.  Each round's
// (c0, c2) lands in the stream's pinned mailbox from the reducing launch itself.
//
// Rounds go two per pass over the vectors while they have at least 16 elements (a
// power of two): the pass sums round k's polynomial and round k+1's coefficients in
// round k's challenge (sumcheck_pair_dev), so round k+1's message needs no pass and no
// host round trip of its own; the two challenges are folded into the vectors by the
// next pass.  The messages, challenges and folded vectors are the ones the round-by-
// round loop below makes.
//
// view, when given, is b as a covector view (b's contents are then unused): pairs passes
// read it (sumcheck_pair_view_dev), each double fold folds its high tables, and b gets
// the view's values when the view ends (the last double fold writes them, any other
// use of b writes them first).
template <int D>
std::vector<E<D>> sumcheck_prove_pairs(Transcript& t, const SumcheckCfg& sc, int engine, const uint64_t* a_src,
                                       int a_Ds, uint64_t* a, uint64_t* b, size_t& n, E<D>& sum, hipStream_t s,
                                       const std::function<void()>& after_first, const CovView* view) {
  CovView cur = view ? *view : CovView{};  // cur.nt == 0: b holds the covector
  DevBuf folded_hi[2];                      // high tables of the folded views, in turn
  int fh = 0;
  // The view folded by the pending pair (the pass that folds the vectors uses it).
  auto fold_view = [&](const E<D>& r1, const E<D>& r2) {
    folded_hi[fh] = DevBuf((size_t)cur.nt * (cur.nhi >> 2) * D * 8, s);
    const CovView f = covview_fold2_dev(D, cur, r1.c, r2.c, folded_hi[fh].as<uint64_t>(), s);
    fh ^= 1;
    return f;
  };
  // A view pass needs four high-table blocks per quarter (two more bits to fold).
  auto view_ok = [&](size_t folds) { return cur.nt && (cur.nhi >> (2 * folds)) >= 4; };
  auto materialize = [&] {
    if (cur.nt) covview_write_dev(D, cur, b, (size_t)cur.nhi << cur.sh, s);  // still unfolded
    cur = CovView{};
  };
  bool first = true;
  auto synced = [&] {
    mailbox_wait(s);
    if (first && after_first) after_first();
    first = false;
  };
  std::vector<E<D>> chals, pend;  // pend: challenges not yet folded into (a_src, b)
  uint64_t* mbox = result_mailbox(s);
  size_t len = n;                 // length after the pending folds
  size_t round = 0;
  auto send = [&](const E<D>& c0, const E<D>& c2) {
    const E<D> c1 = gl::esub<D>(gl::esub<D>(gl::esub<D>(sum, c0), c0), c2);
    t.prover_field(D, c0.c, 1);
    t.prover_field(D, c2.c, 1);
    pow_prove(t, engine, sc.round_pow, s);
    const E<D> r = verifier_field<D>(t);
    chals.push_back(r);
    sum = gl::eadd<D>(gl::emul<D>(gl::eadd<D>(gl::emul<D>(c2, r), c1), r), c0);
    return r;
  };
  while (round < sc.num_rounds) {
    mailbox_arm(s);  // the pass below ends with its sums in the mailbox
    if (sc.num_rounds - round >= 2 && len >= 16 && pend.size() != 1) {
      // round k's polynomial and round k+1's coefficients, folding the pending pair first
      if (cur.nt && !view_ok(pend.empty() ? 0 : 1)) materialize();
      if (pend.empty()) {
        if (cur.nt) sumcheck_pair_view_dev(0, a_Ds, D, a_src, a, b, len, nullptr, nullptr, cur, mbox, s);
        else sumcheck_pair_dev(0, a_Ds, D, a_src, a, b, len, nullptr, nullptr, mbox, s);
      } else {
        if (cur.nt) {
          cur = fold_view(pend[0], pend[1]);
          sumcheck_pair_view_dev(1, a_Ds, D, a_src, a, b, len, pend[0].c, pend[1].c, cur, mbox, s);
        } else {
          sumcheck_pair_dev(1, a_Ds, D, a_src, a, b, len, pend[0].c, pend[1].c, mbox, s);
        }
        a_src = a;
        a_Ds = D;
      }
      synced();
      uint64_t h[8 * D];
      memcpy(h, mbox, sizeof h);
      const E<D> r1 = send(ld<D>(h), ld<D>(h + D));
      const E<D> r1sq = gl::emul<D>(r1, r1);
      const E<D> c0 = gl::eadd<D>(gl::eadd<D>(ld<D>(h + 2 * D), gl::emul<D>(r1, ld<D>(h + 3 * D))),
                                  gl::emul<D>(r1sq, ld<D>(h + 4 * D)));
      const E<D> c2 = gl::eadd<D>(gl::eadd<D>(ld<D>(h + 5 * D), gl::emul<D>(r1, ld<D>(h + 6 * D))),
                                  gl::emul<D>(r1sq, ld<D>(h + 7 * D)));
      const E<D> r2 = send(c0, c2);
      pend = {r1, r2};
      len >>= 2;
      round += 2;
    } else {
      // one round: its polynomial after folding whatever is pending
      if (pend.size() == 2) {
        if (cur.nt && !view_ok(1)) materialize();
        if (cur.nt) {
          cur = fold_view(pend[0], pend[1]);
          sumcheck_pair_view_dev(2, a_Ds, D, a_src, a, b, len, pend[0].c, pend[1].c, cur, mbox, s);
        } else {
          sumcheck_pair_dev(2, a_Ds, D, a_src, a, b, len, pend[0].c, pend[1].c, mbox, s);
        }
        a_src = a;
        a_Ds = D;
      } else if (pend.size() == 1) {
        materialize();
        fold_and_poly_mixed_dev(a_Ds, D, a_src, a, b, 2 * len, pend[0].c, mbox, s);
        a_src = a;
        a_Ds = D;
      } else {
        materialize();
        sumcheck_poly_mixed_dev(a_Ds, D, a_src, b, len, mbox, s);
      }
      synced();
      uint64_t h[2 * D];
      memcpy(h, mbox, sizeof h);
      pend = {send(ld<D>(h), ld<D>(h + D))};
      len >>= 1;
      round += 1;
    }
  }
  if (pend.size() == 2) {
    if (cur.nt && cur.nhi < 4) materialize();
    if (cur.nt) {
      const CovView f = fold_view(pend[0], pend[1]);
      cur = CovView{};
      sumcheck_pair_view_dev(3, a_Ds, D, a_src, a, b, len, pend[0].c, pend[1].c, f, nullptr, s);  // writes b
    } else {
      sumcheck_pair_dev(3, a_Ds, D, a_src, a, b, len, pend[0].c, pend[1].c, nullptr, s);
    }
  } else if (pend.size() == 1) {
    materialize();
    fold_both_dev(a_Ds, D, a_src, a, b, 2 * len, pend[0].c, s);
  } else {
    materialize();
    if (a_src != a) lift_dev(D, a_src, len, a, s);
  }
  n = len;
  return chals;
}

//
// after_first, when given, runs once after the first synchronization (before the first
// message is sent), or at the end when there are no rounds: a caller that queued
// work ahead of the sumcheck (an opening's gather) reads its results there.
//
// view: b given as a covector view of length n (see sumcheck_prove_pairs).
// Sumchecks over at most this many elements run on the host: both vectors come over in
// one wait, every round's sums, challenge and fold are host arithmetic (a few hundred
// extension products at most), and the folded vectors go back with one copy.  On the
// GPU such a sumcheck is four or five one-block launches of 4-16 us each with a host
// round trip per one or two rounds.
constexpr size_t HOST_SUMCHECK_MAX = 64;

// sumcheck::Config::prove on the host for short vectors (a and b of length n in device
// memory, extension elements): the same messages, challenges and folds as the device
// rounds (x'_i = x_i + r (x_(h+i) - x_i), c0 = sum x_i y_i, c2 = sum dx_i dy_i over
// i < h = n/2).
template <int D>
std::vector<E<D>> sumcheck_prove_host(Transcript& t, const SumcheckCfg& sc, int engine, uint64_t* a, uint64_t* b,
                                      size_t& n, E<D>& sum, hipStream_t s, const std::function<void()>& after_first) {
  const size_t words = n * D;
  WHIR_REQUIRE(n <= HOST_SUMCHECK_MAX && 2 * words <= HOST_STAGE_WORDS, "sumcheck: vectors too long for the host path");
  uint64_t* stage = host_stage();
  mailbox_arm(s);
  WHIR_HIP(hipMemcpyAsync(stage, a, words * 8, hipMemcpyDefault, s));
  WHIR_HIP(hipMemcpyAsync(stage + words, b, words * 8, hipMemcpyDefault, s));
  mailbox_wait(s);  // (results queued ahead, e.g. an opening's rows, are in the mailbox)
  if (after_first) after_first();
  std::vector<E<D>> x(n), y(n);
  for (size_t i = 0; i < n; i++) {
    x[i] = ld<D>(stage + i * D);
    y[i] = ld<D>(stage + words + i * D);
  }
  std::vector<E<D>> chals;
  size_t len = n;
  for (size_t round = 0; round < sc.num_rounds; round++) {
    const size_t h = len / 2;
    E<D> c0 = gl::ext_zero<D>(), c2 = gl::ext_zero<D>();
    for (size_t i = 0; i < h; i++) {
      c0 = gl::eadd<D>(c0, gl::emul<D>(x[i], y[i]));
      c2 = gl::eadd<D>(c2, gl::emul<D>(gl::esub<D>(x[h + i], x[i]), gl::esub<D>(y[h + i], y[i])));
    }
    const E<D> c1 = gl::esub<D>(gl::esub<D>(gl::esub<D>(sum, c0), c0), c2);
    t.prover_field(D, c0.c, 1);
    t.prover_field(D, c2.c, 1);
    pow_prove(t, engine, sc.round_pow, s);
    const E<D> r = verifier_field<D>(t);
    chals.push_back(r);
    sum = gl::eadd<D>(gl::emul<D>(gl::eadd<D>(gl::emul<D>(c2, r), c1), r), c0);
    for (size_t i = 0; i < h; i++) {
      x[i] = gl::eadd<D>(x[i], gl::emul<D>(r, gl::esub<D>(x[h + i], x[i])));
      y[i] = gl::eadd<D>(y[i], gl::emul<D>(r, gl::esub<D>(y[h + i], y[i])));
    }
    len = h;
  }
  // The folded vectors back to the device.
  for (size_t i = 0; i < len; i++) {
    std::copy(x[i].c, x[i].c + D, stage + i * D);
    std::copy(y[i].c, y[i].c + D, stage + (len + i) * D);
  }
  WHIR_HIP(hipMemcpyAsync(a, stage, len * D * 8, hipMemcpyDefault, s));
  WHIR_HIP(hipMemcpyAsync(b, stage + len * D, len * D * 8, hipMemcpyDefault, s));
  n = len;
  return chals;
}

template <int D>
std::vector<E<D>> sumcheck_prove(Transcript& t, const SumcheckCfg& sc, int engine, const uint64_t* a_src, int a_Ds,
                                 uint64_t* a, uint64_t* b, size_t& n, E<D>& sum, hipStream_t s,
                                 const std::function<void()>& after_first = {}, const CovView* view = nullptr) {
  WHIR_REQUIRE(n == sc.initial_size, "sumcheck: vector length does not match the configuration");
  if (n <= HOST_SUMCHECK_MAX && n >= 2 && (n & (n - 1)) == 0 && sc.num_rounds >= 1 && !view && a_src == a &&
      a_Ds == D && ((size_t)1 << sc.num_rounds) <= n)
    return sumcheck_prove_host<D>(t, sc, engine, a, b, n, sum, s, after_first);
  if (n >= 16 && (n & (n - 1)) == 0 && sc.num_rounds >= 2)
    return sumcheck_prove_pairs<D>(t, sc, engine, a_src, a_Ds, a, b, n, sum, s, after_first, view);
  if (view) covview_write_dev(D, *view, b, n, s);
  bool first = true;
  std::vector<E<D>> chals;
  uint64_t* mbox = result_mailbox(s);
  bool have_prev = false;
  E<D> prev;
  for (size_t round = 0; round < sc.num_rounds; round++) {
    mailbox_arm(s);
    if (have_prev) {
      n = fold_and_poly_mixed_dev(a_Ds, D, a_src, a, b, n, prev.c, mbox, s);
      a_src = a;
      a_Ds = D;
    } else {
      sumcheck_poly_mixed_dev(a_Ds, D, a_src, b, n, mbox, s);
    }
    mailbox_wait(s);
    if (first && after_first) after_first();
    first = false;
    uint64_t h[2 * D];
    memcpy(h, mbox, sizeof h);
    E<D> c0 = ld<D>(h), c2 = ld<D>(h + D);
    E<D> c1 = gl::esub<D>(gl::esub<D>(gl::esub<D>(sum, c0), c0), c2);
    t.prover_field(D, c0.c, 1);
    t.prover_field(D, c2.c, 1);
    pow_prove(t, engine, sc.round_pow, s);
    E<D> r = verifier_field<D>(t);
    chals.push_back(r);
    sum = gl::eadd<D>(gl::emul<D>(gl::eadd<D>(gl::emul<D>(c2, r), c1), r), c0);
    prev = r;
    have_prev = true;
  }
  if (have_prev) {
    n = fold_both_dev(a_Ds, D, a_src, a, b, n, prev.c, s);
  } else if (a_src != a) {
    lift_dev(D, a_src, n, a, s);
  }
  if (first && after_first) {
    mailbox_wait(s);
    after_first();
  }
  return chals;
}

// eq_weights (src/algebra/multilinear.rs:154-158) on the host: 2^k weights, x_0 = MSB.
template <int D>
std::vector<E<D>> eq_weights(const std::vector<E<D>>& z) {
  std::vector<E<D>> cur{gl::ext_base<D>(1)};
  for (const auto& zi : z) {
    E<D> om = gl::esub<D>(gl::ext_base<D>(1), zi);
    std::vector<E<D>> nxt(cur.size() * 2);
    for (size_t x = 0; x < cur.size(); x++) {
      nxt[2 * x] = gl::emul<D>(cur[x], om);
      nxt[2 * x + 1] = gl::emul<D>(cur[x], zi);
    }
    cur.swap(nxt);
  }
  return cur;
}

template <int D>
void commit_impl(const WhirCfg& cfg, Transcript& t, const uint64_t* const* vecs, size_t nvec, DevWitness& w,
                 hipStream_t s) {
  irs_commit<D>(t, cfg.initial, cfg.engine, vecs, nvec, 1, w, s);
}

// ------------------------------------------------------------------ batched commits
// whir::Config::commit (irs_commit.rs:292-350) of many independent instances, each with
// its own configuration, transcript and vector, as shared launches: the instances are
// grouped by shape (message length, codeword length, messages, hash engine), and each
// group's encodes, leaf hashes and Merkle levels run as one launch per step over
// device tables of the instances' vectors, codewords and trees.  That launch sequence
// depends only on the batch's layout (the list of shapes), so it is recorded once as a
// HIP graph (stream capture) and replayed for every later batch of the same layout,
// with the tables rewritten before each replay.  Then, on the host, each root goes to
// its transcript and the OOD points come out; one batched launch evaluates every
// instance's vector at its points, and the values go to the transcripts.  Every
// transcript sees the calls of a plain commit, in the same order, so the witnesses
// and transcripts are those of commit_impl.
struct CommitShape {
  int engine;
  size_t L, N, M;
  bool operator==(const CommitShape& o) const {
    return engine == o.engine && L == o.L && N == o.N && M == o.M;
  }
};

struct CommitGroup {
  CommitShape shape;
  std::vector<size_t> members;  // positions in the batch
  DevBuf tabs;                  // device pointer tables: [messages][codewords][nodes]
  DevBuf scratch;               // the group's NTT scratch (groups run concurrently)
  size_t root0 = 0;             // first root slot of the group
  hipEvent_t done = nullptr;
};

struct CommitPlan {
  std::vector<CommitShape> layout;
  std::vector<CommitGroup> groups;
  DevBuf roots;  // 32 bytes per batched instance
  hipEvent_t fork = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  uint64_t uses = 0;
  ~CommitPlan() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (fork) (void)hipEventDestroy(fork);
    for (CommitGroup& g : groups)
      if (g.done) (void)hipEventDestroy(g.done);
  }
};

// Side streams of the calling thread (default priority), one per concurrent group of
// a batched commit; they fork from and join back into the thread's stream.
struct SideStreams {
  std::vector<hipStream_t> s;
  ~SideStreams() { release(); }
  void release() {
    for (hipStream_t x : s) (void)hipStreamDestroy(x);
    s.clear();
  }
  hipStream_t get(size_t i) {
    while (s.size() <= i) {
      hipStream_t x = nullptr;
      WHIR_HIP(hipStreamCreateWithFlags(&x, hipStreamNonBlocking));
      s.push_back(x);
    }
    return s[i];
  }
};
thread_local SideStreams t_side;
constexpr size_t MAX_SIDE_STREAMS = 8;

// Plans of the calling thread's stream, most recently used first (a few layouts).
thread_local std::vector<std::unique_ptr<CommitPlan>> t_plans;
constexpr size_t MAX_PLANS = 4;

}  // namespace

// Frees the calling thread's plans (their buffers go back to the thread's stream, their
// graphs are destroyed); called before the thread's stream is released.
void release_commit_plans() {
  t_plans.clear();
  t_side.release();
  t_copy.release();
}

namespace {

bool batchable(const WhirCfg& c) {
  const IrsCfg& irs = c.initial;
  const size_t L = irs.message_length(), N = irs.codeword_length;
  return irs.num_vectors == 1 && irs.mask_length == 0 && L > 0 && (L & (L - 1)) == 0 && N >= 8 &&
         (N & (N - 1)) == 0 && N % L == 0;
}

// The groups' launches, each group on a side stream forked from s and joined back
// (in a capture these become parallel branches of the graph).
void run_commit_plan(CommitPlan& P, hipStream_t s) {
  WHIR_HIP(hipEventRecord(P.fork, s));
  for (size_t gi = 0; gi < P.groups.size(); gi++) {
    CommitGroup& g = P.groups[gi];
    const hipStream_t gs = t_side.get(gi % MAX_SIDE_STREAMS);
    WHIR_HIP(hipStreamWaitEvent(gs, P.fork, 0));
    const CommitShape& sh = g.shape;
    const size_t cnt = g.members.size();
    const uint64_t* const* msg_tab = (const uint64_t* const*)g.tabs.ptr;
    uint64_t* const* cw_tab = (uint64_t* const*)(g.tabs.as<uint64_t*>() + cnt);
    uint8_t* const* nodes_tab = (uint8_t* const*)(g.tabs.as<uint64_t*>() + 2 * cnt);
    RsArgs a{nullptr, sh.L, sh.L, nullptr, 0, nullptr, sh.N, sh.N / sh.L, log2_exact(sh.L), sh.M * cnt};
    a.msg_tab = msg_tab;
    a.out_tab = cw_tab;
    a.mpi = sh.M;
    rs_encode_cols(1, a, g.scratch.as<uint64_t>(), gs);
    commit_cols_many_dev(sh.engine, 1, (const uint64_t* const*)cw_tab, nodes_tab, cnt, sh.N, sh.M,
                         P.roots.as<uint8_t>() + 32 * g.root0, gs);
    WHIR_HIP(hipEventRecord(g.done, gs));
  }
  for (CommitGroup& g : P.groups) WHIR_HIP(hipStreamWaitEvent(s, g.done, 0));
}

CommitPlan& commit_plan(const std::vector<CommitShape>& layout, hipStream_t s) {
  for (size_t i = 0; i < t_plans.size(); i++)
    if (t_plans[i]->layout == layout) {
      if (i) std::rotate(t_plans.begin(), t_plans.begin() + i, t_plans.begin() + i + 1);
      return *t_plans[0];
    }
  auto P = std::make_unique<CommitPlan>();
  P->layout = layout;
  for (size_t k = 0; k < layout.size(); k++) {
    auto it = std::find_if(P->groups.begin(), P->groups.end(), [&](const CommitGroup& g) { return g.shape == layout[k]; });
    if (it == P->groups.end()) {
      P->groups.emplace_back();
      P->groups.back().shape = layout[k];
      it = P->groups.end() - 1;
    }
    it->members.push_back(k);
  }
  size_t slot = 0;
  for (CommitGroup& g : P->groups) {
    const size_t cnt = g.members.size();
    g.root0 = slot;
    slot += cnt;
    g.tabs = DevBuf(3 * cnt * 8, s);
    g.scratch = DevBuf(rs_scratch_words(1, g.shape.M * cnt, g.shape.N, g.shape.L) * 8, s);
    WHIR_HIP(hipEventCreateWithFlags(&g.done, hipEventDisableTiming));
  }
  P->roots = DevBuf(32 * layout.size(), s);
  WHIR_HIP(hipEventCreateWithFlags(&P->fork, hipEventDisableTiming));
  if (t_plans.size() >= MAX_PLANS) t_plans.pop_back();
  t_plans.insert(t_plans.begin(), std::move(P));
  return *t_plans[0];
}

template <int D>
void finish_ood(const std::vector<const WhirCfg*>& cfgs, std::vector<Transcript*>& ts,
                const std::vector<const uint64_t*>& vecs, std::vector<DevWitness*>& ws, const std::vector<size_t>& idx,
                hipStream_t s) {
  std::vector<UniJob> jobs;
  std::vector<std::pair<size_t, size_t>> where;  // (instance, sample)
  for (size_t k : idx) {
    const IrsCfg& c = cfgs[k]->initial;
    ws[k]->ood_points.clear();
    for (size_t i = 0; i < c.out_domain_samples; i++) {
      const E<D> x = verifier_field<D>(*ts[k]);
      push<D>(ws[k]->ood_points, x);
      UniJob j{vecs[k], c.vector_size, {0, 0, 0}};
      for (int d = 0; d < D; d++) j.x[d] = x.c[d];
      jobs.push_back(j);
      where.emplace_back(k, i);
    }
  }
  std::vector<uint64_t> vals(jobs.size() * D);
  if (!jobs.empty()) {
    DevBuf out(jobs.size() * D * 8, s);
    univariate_eval_batch_dev(D, jobs.data(), jobs.size(), out.as<uint64_t>(), s);
    download_sync(vals.data(), out.ptr, vals.size() * 8, s);
  }
  for (size_t q = 0; q < jobs.size(); q++) {
    DevWitness& w = *ws[where[q].first];
    w.ood_values.insert(w.ood_values.end(), &vals[q * D], &vals[(q + 1) * D]);
    ts[where[q].first]->prover_field(D, &vals[q * D], 1);
  }
}

void commit_many(const std::vector<const WhirCfg*>& cfgs, std::vector<Transcript*>& ts,
                 const std::vector<const uint64_t*>& vecs, std::vector<DevWitness*>& ws, hipStream_t s) {
  const size_t n = cfgs.size();
  // Instances outside the batched path commit one by one.
  std::vector<size_t> batched;
  for (size_t k = 0; k < n; k++) {
    if (batchable(*cfgs[k])) {
      batched.push_back(k);
      continue;
    }
    switch (cfgs[k]->D) {
      case 1: commit_impl<1>(*cfgs[k], *ts[k], &vecs[k], 1, *ws[k], s); break;
      case 2: commit_impl<2>(*cfgs[k], *ts[k], &vecs[k], 1, *ws[k], s); break;
      case 3: commit_impl<3>(*cfgs[k], *ts[k], &vecs[k], 1, *ws[k], s); break;
    }
  }
  if (batched.empty()) return;
  std::vector<CommitShape> layout;
  for (size_t k : batched) {
    const IrsCfg& c = cfgs[k]->initial;
    layout.push_back({cfgs[k]->engine, c.message_length(), c.codeword_length, c.num_messages()});
  }
  CommitPlan& P = commit_plan(layout, s);
  // This batch's witnesses and the tables that point the plan's launches at them.
  for (CommitGroup& g : P.groups) {
    const size_t cnt = g.members.size();
    std::vector<uint64_t> tab(3 * cnt);
    for (size_t i = 0; i < cnt; i++) {
      const size_t k = batched[g.members[i]];
      DevWitness& w = *ws[k];
      w.Ds = 1;
      w.N = g.shape.N;
      w.M = g.shape.M;
      w.num_vectors = 1;
      w.cw = DevBuf(g.shape.M * g.shape.N * 8, s);
      w.nodes = DevBuf(32 * merkle_num_nodes(g.shape.N), s);
      tab[i] = (uint64_t)(uintptr_t)vecs[k];
      tab[cnt + i] = (uint64_t)(uintptr_t)w.cw.ptr;
      tab[2 * cnt + i] = (uint64_t)(uintptr_t)w.nodes.ptr;
    }
    upload_async(g.tabs.ptr, tab.data(), tab.size() * 8, s);
  }
  if (P.exec) {
    WHIR_HIP(hipGraphLaunch(P.exec, s));
  } else {
    // First batch of this layout: run the launches (which also builds the twiddle
    // tables and tunes new NTT shapes), then record them for the later batches.
    run_commit_plan(P, s);
    WHIR_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    set_capture_active(true);
    try {
      run_commit_plan(P, s);
    } catch (...) {
      set_capture_active(false);
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(s, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    set_capture_active(false);
    WHIR_HIP(hipStreamEndCapture(s, &P.graph));
    WHIR_HIP(hipGraphInstantiate(&P.exec, P.graph, nullptr, nullptr, 0));
  }
  P.uses++;
  std::vector<uint8_t> roots(32 * batched.size());
  download_sync(roots.data(), P.roots.ptr, roots.size(), s);
  for (const CommitGroup& g : P.groups)
    for (size_t i = 0; i < g.members.size(); i++) {
      const size_t k = batched[g.members[i]];
      ts[k]->prover_hash(&roots[32 * (g.root0 + i)]);
    }
  // OOD samples of every batched instance, grouped by extension degree.
  for (int D = 1; D <= 3; D++) {
    std::vector<size_t> idx;
    for (size_t k : batched)
      if (cfgs[k]->D == D) idx.push_back(k);
    if (idx.empty()) continue;
    switch (D) {
      case 1: finish_ood<1>(cfgs, ts, vecs, ws, idx, s); break;
      case 2: finish_ood<2>(cfgs, ts, vecs, ws, idx, s); break;
      case 3: finish_ood<3>(cfgs, ts, vecs, ws, idx, s); break;
    }
  }
}

// whir::Config::prove, src/protocols/whir/prover.rs:50-316: vecs[0..V) are the
// committed base-field vectors, witnesses ws[0..W) hold them batch_size at a time
// in order (V = W * batch_size).  Linear form k is a MultilinearExtension (kinds
// null or kinds[k] == WHIR_FORM_MLE; it takes the next num_variables elements of
// `points`), a UnivariateEvaluation (WHIR_FORM_UNIVARIATE; the next element of
// `points`) or a Covector (the next device vector in `covs`, vector_size elements
// of D words); evals[(k*V + j)*D] is vector j under form k.
template <int D>
void prove_impl(const WhirCfg& cfg, Transcript& t, const uint64_t* const* vecs, size_t V,
                const DevWitness* const* ws, size_t W, const int* kinds, size_t np, const uint64_t* points,
                const uint64_t* const* covs, const uint64_t* evals, uint64_t* eval_point_out, hipStream_t s) {
  const int eng = cfg.engine;
  const size_t B = cfg.initial.num_vectors;
  WHIR_REQUIRE(V > 0 && W > 0, "prove: no vectors");
  WHIR_REQUIRE(V == W * B, "prove: the number of vectors must be the number of witnesses times batch_size");
  for (size_t w = 0; w < W; w++)
    WHIR_REQUIRE(ws[w] && ws[w]->num_vectors == B && ws[w]->N == cfg.initial.codeword_length &&
                     ws[w]->M == cfg.initial.num_messages(),
                 "prove: witness does not match the configuration");
  size_t n = cfg.initial.vector_size;
  size_t nv = 0;
  while (((size_t)1 << nv) < n) nv++;
  uint64_t* mbox = result_mailbox(s);

  // Complete the OOD matrix with its cross terms (prover.rs:98-125): the rows of
  // witness w hold its own vectors; every other vector is evaluated at w's OOD
  // points here and sent.
  std::vector<uint64_t> opts;  // all OOD points, witness by witness
  for (size_t w = 0; w < W; w++) opts.insert(opts.end(), ws[w]->ood_points.begin(), ws[w]->ood_points.end());
  const size_t nood = opts.size() / D;
  std::vector<uint64_t> cross;  // cross[(j*nood + o)*D]: vector j at OOD point o
  if (W > 1 && nood) {
    const size_t pad = (nood + 3) / 4 * 4;
    WHIR_REQUIRE(V * pad * D <= RESULT_MAILBOX_WORDS, "too many out-of-domain cross terms");
    for (size_t j = 0; j < V; j++) {
      if (j + 1 == V) mailbox_arm(s);
      univariate_eval_many_dev(1, D, vecs[j], n, opts.data(), nood, mbox + j * pad * D, s, true);
    }
    mailbox_wait(s);
    cross.resize(V * nood * D);
    for (size_t j = 0; j < V; j++) std::copy(mbox + j * pad * D, mbox + (j * pad + nood) * D, &cross[j * nood * D]);
  }
  std::vector<E<D>> omat;  // nood x V
  for (size_t w = 0, o = 0; w < W; w++) {
    const size_t wn = ws[w]->ood_points.size() / D;
    for (size_t i = 0; i < wn; i++, o++)
      for (size_t j = 0; j < V; j++) {
        if (j / B == w) {
          omat.push_back(ld<D>(&ws[w]->ood_values[(i * B + j % B) * D]));
        } else {
          omat.push_back(ld<D>(&cross[(j * nood + o) * D]));
          t.prover_field(D, omat.back().c, 1);
        }
      }
  }

  // Random linear combination of the vectors ([1] for one vector, no transcript).
  std::vector<E<D>> vector_rlc = geometric_challenge<D>(t, V);
  DevBuf vec(n * D * 8, s), cov(n * D * 8, s);
  // The sumcheck's first operand: the committed vector itself while there is one
  // (read from the base field without lifting it), else the combination.
  const uint64_t* a0 = vecs[0];
  int a0_Ds = 1;
  if (V > 1) {
    lift_dev(D, vecs[0], n, vec.as<uint64_t>(), s);
    for (size_t j = 1; j < V; j++) mixed_scalar_mul_add_dev(D, vec.as<uint64_t>(), vector_rlc[j].c, vecs[j], n, s);
    a0 = vec.as<uint64_t>();
    a0_Ds = D;
  }
  auto rlc_dot = [&](const E<D>* row) {
    E<D> acc = gl::ext_zero<D>();
    for (size_t j = 0; j < V; j++) acc = gl::eadd<D>(acc, gl::emul<D>(vector_rlc[j], row[j]));
    return acc;
  };
  std::vector<E<D>> crlc = geometric_challenge<D>(t, np + nood);
  const bool has_constraints = !crlc.empty();
  // The initial covector is kept as a view (a few product terms in small tables) when it
  // has no Covector forms: the first sumcheck pass reads it from the tables instead of
  // from 2^nv values in HBM, and the first fold writes the folded values.
  bool use_view = false;
  DevBuf view_buf;
  CovView view;
  {
    // cov = sum_{MLE forms k} crlc_k eq(point_k, .) + sum_o crlc_(np+o) (1, x_o, x_o^2, ...)
    // in one pass, then + crlc_k covector_k for the Covector forms (Covector::accumulate,
    // linear_form/covector.rs:21-23); the order of the sums does not change the field values.
    // UnivariateEvaluation forms join the OOD terms (UnivariateEvaluation::accumulate,
    // linear_form/univariate_evaluation.rs:60-65, is the same geometric sum).
    std::vector<uint64_t> erlc, srlc, mle_pts, geo_pts;
    std::vector<size_t> cov_forms;
    size_t nmle = 0;
    const uint64_t* pp = points;
    for (size_t k = 0; k < np; k++) {
      const int kind = kinds ? kinds[k] : WHIR_FORM_MLE;
      if (kind == WHIR_FORM_COVECTOR) {
        cov_forms.push_back(k);
      } else if (kind == WHIR_FORM_UNIVARIATE) {
        push<D>(srlc, crlc[k]);
        geo_pts.insert(geo_pts.end(), pp, pp + D);
        pp += D;
      } else {
        push<D>(erlc, crlc[k]);
        mle_pts.insert(mle_pts.end(), pp, pp + nv * D);
        pp += nv * D;
        nmle++;
      }
    }
    for (size_t k = 0; k < nood; k++) push<D>(srlc, crlc[np + k]);
    geo_pts.insert(geo_pts.end(), opts.begin(), opts.end());
    const size_t ngeo = srlc.size() / D;
    use_view = cov_view_enabled() && cov_forms.empty() && nmle + ngeo > 0 && nmle + ngeo <= COV_VIEW_MAX_TERMS &&
               n == ((size_t)1 << nv) && nv >= 4;
    if (use_view) {
      view_buf = DevBuf(covview_words(D, nmle, ngeo, nv) * 8, s);
      view = covview_build_dev(D, view_buf.as<uint64_t>(), nv, mle_pts.data(), erlc.data(), nmle, srlc.data(),
                               geo_pts.data(), ngeo, s);
    } else {
      covector_dev(D, cov.as<uint64_t>(), n, false, mle_pts.data(), erlc.data(), nmle, nv, srlc.data(),
                   geo_pts.data(), ngeo, s);
      for (size_t q = 0; q < cov_forms.size(); q++)
        scalar_mul_add_dev(D, cov.as<uint64_t>(), crlc[cov_forms[q]].c, covs[q], n, s);
    }
  }
  E<D> the_sum = gl::ext_zero<D>();
  {
    std::vector<E<D>> row(V);
    for (size_t k = 0; k < np; k++) {
      for (size_t j = 0; j < V; j++) row[j] = ld<D>(evals + (k * V + j) * D);
      the_sum = gl::eadd<D>(the_sum, gl::emul<D>(crlc[k], rlc_dot(row.data())));
    }
    for (size_t o = 0; o < nood; o++) the_sum = gl::eadd<D>(the_sum, gl::emul<D>(crlc[np + o], rlc_dot(&omat[o * V])));
  }
  std::vector<E<D>> fr;
  if (has_constraints) {
    fr = sumcheck_prove<D>(t, cfg.initial_sumcheck, eng, a0, a0_Ds, vec.as<uint64_t>(), cov.as<uint64_t>(), n,
                           the_sum, s, {}, use_view ? &view : nullptr);
  } else {
    for (size_t i = 0; i < cfg.initial_sumcheck.num_rounds; i++) fr.push_back(verifier_field<D>(t));
    pow_prove(t, eng, cfg.initial_skip_pow, s);
    const uint64_t* src = a0;
    int src_Ds = a0_Ds;
    for (const auto& f : fr) {
      n = fold_into_dev(src_Ds, D, src, vec.as<uint64_t>(), n, f.c, s);
      src = vec.as<uint64_t>();
      src_Ds = D;
    }
    if (fr.empty() && a0 != vec.as<uint64_t>()) lift_dev(D, a0, n, vec.as<uint64_t>(), s);
    fill_zero(cov.ptr, n * D * 8, s);  // the covector is all zeros
  }
  std::vector<E<D>> eval_point(fr);

  // In-domain value of opened row q: dot(tensor_product(poly_rlc, eq_weights(r)),
  // row) over the rows of every opened witness side by side (irs_commit.rs:535-537).
  // The weight of column col of witness w (column w*M + col of the concatenated rows,
  // i.e. vector g / depth, chunk g % depth); the gather launch of each opening sums
  // its rows against them, and the witnesses' sums are added here.
  auto column_weights = [&](const std::vector<const DevWitness*>& pw, const std::vector<E<D>>& poly_rlc,
                            const std::vector<E<D>>& w_eq) {
    std::vector<std::vector<uint64_t>> wts(pw.size());
    const size_t depth = w_eq.size();
    for (size_t w = 0; w < pw.size(); w++)
      for (size_t col = 0; col < pw[w]->M; col++) {
        const size_t g = w * pw[w]->M + col;
        push<D>(wts[w], gl::emul<D>(poly_rlc[g / depth], w_eq[g % depth]));
      }
    return wts;
  };

  const IrsCfg* prev_cfg = &cfg.initial;
  std::vector<const DevWitness*> prev_ws(ws, ws + W);
  std::unique_ptr<DevWitness> held;
  for (const RoundCfg& rc : cfg.rounds) {
    auto nw = std::make_unique<DevWitness>();
    const uint64_t* vp = vec.as<uint64_t>();
    irs_commit<D>(t, rc.irs, eng, &vp, 1, D, *nw, s);
    pow_prove(t, eng, rc.pow, s);
    // The opening's gather, the covector terms and the sumcheck's first pass are queued
    // back to back with one synchronization: the STIR challenge depends on the
    // transcript alone (hints do not enter it), and the in-domain values, which only
    // the claimed sum needs, come back with the first pass's sums.
    std::vector<uint64_t> pts, exps;
    const std::vector<std::vector<uint64_t>> wts = column_weights(prev_ws, vector_rlc, eq_weights<D>(fr));
    std::vector<OpenLaunch> opens = irs_open_launch(t, *prev_cfg, prev_ws, pts, s, D, &wts, &exps);
    // STIR constraints: OOD points of the new commitment, then the opened rows.
    const size_t nood = nw->ood_values.size() / D, kq = pts.size();
    std::vector<E<D>> srlc = geometric_challenge<D>(t, nood + kq);
    if (!srlc.empty()) {
      std::vector<uint64_t> flat;
      for (auto& e : srlc) push<D>(flat, e);
      covector_stir_dev(D, cov.as<uint64_t>(), n, true, flat.data(), nw->ood_points.data(), nood, exps.data(),
                        pts.data(), kq, log2_exact(prev_cfg->codeword_length), s);
    }
    for (size_t k = 0; k < nood; k++)
      the_sum = gl::eadd<D>(the_sum, gl::emul<D>(srlc[k], ld<D>(&nw->ood_values[k * D])));
    auto finish_open = [&] {
      std::vector<std::vector<uint64_t>> rows, vals;
      irs_open_finish(t, opens, rows, vals, s);
      for (size_t q = 0; q < kq; q++) {
        E<D> v = gl::ext_zero<D>();
        for (const auto& vw : vals) v = gl::eadd<D>(v, ld<D>(&vw[q * D]));
        the_sum = gl::eadd<D>(the_sum, gl::emul<D>(srlc[nood + q], v));
      }
    };
    fr = sumcheck_prove<D>(t, rc.sumcheck, eng, vec.as<uint64_t>(), D, vec.as<uint64_t>(), cov.as<uint64_t>(), n,
                           the_sum, s, finish_open);
    eval_point.insert(eval_point.end(), fr.begin(), fr.end());
    held = std::move(nw);
    prev_cfg = &rc.irs;
    prev_ws.assign(1, held.get());
    vector_rlc.assign(1, gl::ext_base<D>(1));
  }
  // Send the final vector directly.
  WHIR_REQUIRE(n == cfg.final_sumcheck.initial_size, "final vector length mismatch");
  std::vector<uint64_t> fin(n * D);
  download_sync(fin.data(), vec.ptr, n * D * 8, s);
  for (size_t i = 0; i < n; i++) t.prover_field(D, &fin[i * D], 1);
  pow_prove(t, eng, cfg.final_pow, s);
  // The final opening is queued ahead of the final sumcheck's first pass and read back
  // with it.
  std::vector<uint64_t> fpts;
  std::vector<OpenLaunch> fopens = irs_open_launch(t, *prev_cfg, prev_ws, fpts, s);
  auto finish_final = [&] {
    std::vector<std::vector<uint64_t>> rows, vals;
    irs_open_finish(t, fopens, rows, vals, s);
  };
  fr = sumcheck_prove<D>(t, cfg.final_sumcheck, eng, vec.as<uint64_t>(), D, vec.as<uint64_t>(), cov.as<uint64_t>(),
                         n, the_sum, s, finish_final);
  eval_point.insert(eval_point.end(), fr.begin(), fr.end());
  if (eval_point_out)
    for (size_t i = 0; i < eval_point.size(); i++)
      for (int d = 0; d < D; d++) eval_point_out[i * D + d] = eval_point[i].c[d];
  WHIR_HIP(hipStreamSynchronize(s));
}

}  // namespace
}  // namespace whir

#define API_BEGIN try {
#define API_END                                \
  }                                            \
  catch (const std::exception& e) {            \
    set_last_error(e.what());                  \
    return 1;                                  \
  }                                            \
  catch (...) {                                \
    set_last_error("unknown error");           \
    return 1;                                  \
  }                                            \
  set_last_error("");                          \
  return 0;

extern "C" {

int whir_config_new(const whir_params* p, size_t size, whir_config** out) {
  API_BEGIN
  WhirParams wp;
  wp.degree = p->degree;
  wp.security_level = p->security_level;
  wp.pow_bits = p->pow_bits;
  wp.initial_folding_factor = p->initial_folding_factor;
  wp.folding_factor = p->folding_factor;
  wp.unique_decoding = p->unique_decoding != 0;
  wp.starting_log_inv_rate = p->starting_log_inv_rate;
  wp.batch_size = p->batch_size;
  wp.hash_engine = p->hash_engine;
  WHIR_REQUIRE(engine_supports_size(wp.hash_engine, 64), "hash engine cannot hash Merkle nodes");
  auto* c = new whir_config{WhirCfg::make(size, wp)};
  *out = c;
  API_END
}

int whir_config_disable_pow(whir_config* c) {
  API_BEGIN
  c->cfg.disable_pow();
  API_END
}

int whir_config_describe(const whir_config* c, char* buf, size_t cap, size_t* len) {
  API_BEGIN
  std::string s = c->cfg.describe();
  *len = s.size();
  if (buf && cap) {
    size_t n = std::min(cap - 1, s.size());
    memcpy(buf, s.data(), n);
    buf[n] = 0;
  }
  API_END
}

void whir_config_free(whir_config* c) { delete c; }

int whir_transcript_new(const uint8_t protocol_id[64], const uint8_t session_id[32], const uint8_t* instance,
                        size_t len, whir_transcript** out) {
  API_BEGIN
  auto* t = new whir_transcript;
  auto d = std::make_unique<Duplex>(protocol_id, session_id, instance, len);
  t->duplex = d.get();
  t->t = std::move(d);
  *out = t;
  API_END
}

int whir_transcript_from_vtable(const whir_transcript_vtable* vt, whir_transcript** out) {
  API_BEGIN
  WHIR_REQUIRE(vt && vt->prover_field && vt->verifier_field && vt->verifier_bytes && vt->prover_hash &&
                   vt->prover_u64 && vt->hint_hash && vt->hint_field_vec,
               "transcript vtable: every callback must be set");
  auto* t = new whir_transcript;
  t->t = std::make_unique<VtableTranscript>(*vt);
  *out = t;
  API_END
}

int whir_transcript_proof(const whir_transcript* t, uint8_t* narg, size_t narg_cap, size_t* narg_len,
                          uint8_t* hints, size_t hints_cap, size_t* hints_len) {
  API_BEGIN
  WHIR_REQUIRE(t->duplex, "transcript_proof: only the built-in transcript holds proof bytes");
  const auto& a = t->duplex->narg();
  const auto& h = t->duplex->hints();
  *narg_len = a.size();
  *hints_len = h.size();
  if (narg) {
    WHIR_REQUIRE(narg_cap >= a.size(), "transcript_proof: narg buffer too small");
    memcpy(narg, a.data(), a.size());
  }
  if (hints) {
    WHIR_REQUIRE(hints_cap >= h.size(), "transcript_proof: hints buffer too small");
    memcpy(hints, h.data(), h.size());
  }
  API_END
}

void whir_transcript_free(whir_transcript* t) { delete t; }

int whir_commit_batch_dev(const whir_config* c, whir_transcript* t, const uint64_t* const* vectors, size_t nvec,
                          whir_witness** out) {
  API_BEGIN
  WHIR_REQUIRE(c && t && vectors && out, "commit: null argument");
  for (size_t b = 0; b < nvec; b++) WHIR_REQUIRE(vectors[b], "commit: null vector");
  hipStream_t s = thread_stream();
  auto w = std::make_unique<whir_witness>();
  switch (c->cfg.D) {
    case 1: commit_impl<1>(c->cfg, *t->t, vectors, nvec, w->w, s); break;
    case 2: commit_impl<2>(c->cfg, *t->t, vectors, nvec, w->w, s); break;
    case 3: commit_impl<3>(c->cfg, *t->t, vectors, nvec, w->w, s); break;
  }
  *out = w.release();
  API_END
}

int whir_commit_dev(const whir_config* c, whir_transcript* t, const uint64_t* vector, whir_witness** out) {
  return whir_commit_batch_dev(c, t, &vector, 1, out);
}

int whir_commit_host(const whir_config* c, whir_transcript* t, const uint64_t* host_vector, uint64_t* dev_vector,
                     whir_witness** out) {
  API_BEGIN
  WHIR_REQUIRE(c && t && host_vector && dev_vector && out, "commit_host: null argument");
  WHIR_REQUIRE(c->cfg.initial.num_vectors == 1, "commit_host: configurations of batch_size 1 only");
  hipStream_t s = thread_stream();
  auto w = std::make_unique<whir_witness>();
  HostUpload up{host_vector, dev_vector};
  const uint64_t* v = dev_vector;
  switch (c->cfg.D) {
    case 1: irs_commit<1>(*t->t, c->cfg.initial, c->cfg.engine, &v, 1, 1, w->w, s, &up); break;
    case 2: irs_commit<2>(*t->t, c->cfg.initial, c->cfg.engine, &v, 1, 1, w->w, s, &up); break;
    case 3: irs_commit<3>(*t->t, c->cfg.initial, c->cfg.engine, &v, 1, 1, w->w, s, &up); break;
  }
  *out = w.release();
  API_END
}

int whir_commit_many_dev(const whir_config* const* cfgs, whir_transcript* const* ts, const uint64_t* const* vectors,
                         size_t count, whir_witness** out) {
  API_BEGIN
  WHIR_REQUIRE(count == 0 || (cfgs && ts && vectors && out), "commit_many: null argument");
  std::vector<const WhirCfg*> cs(count);
  std::vector<Transcript*> tr(count);
  std::vector<const uint64_t*> vs(count);
  for (size_t k = 0; k < count; k++) {
    WHIR_REQUIRE(cfgs[k] && ts[k] && vectors[k], "commit_many: null configuration, transcript or vector");
    WHIR_REQUIRE(cfgs[k]->cfg.initial.num_vectors == 1, "commit_many: configurations of batch_size 1 only");
    for (size_t j = 0; j < k; j++)
      WHIR_REQUIRE(ts[j] != ts[k], "commit_many: each instance needs its own transcript");
    cs[k] = &cfgs[k]->cfg;
    tr[k] = ts[k]->t.get();
    vs[k] = vectors[k];
    out[k] = nullptr;
  }
  std::vector<std::unique_ptr<whir_witness>> ws(count);
  std::vector<DevWitness*> wp(count);
  for (size_t k = 0; k < count; k++) {
    ws[k] = std::make_unique<whir_witness>();
    wp[k] = &ws[k]->w;
  }
  commit_many(cs, tr, vs, wp, thread_stream());
  for (size_t k = 0; k < count; k++) out[k] = ws[k].release();
  API_END
}

int whir_prove_forms_dev(const whir_config* c, whir_transcript* t, const uint64_t* const* vectors, size_t nvec,
                         const whir_witness* const* witnesses, size_t nw, const int* kinds, size_t nforms,
                         const uint64_t* points, const uint64_t* const* covectors, const uint64_t* evals,
                         uint64_t* eval_point) {
  API_BEGIN
  WHIR_REQUIRE(c && t && vectors && witnesses, "prove: null argument");
  WHIR_REQUIRE(nforms == 0 || evals, "prove: null evaluations");
  size_t nmle = 0, ncov = 0;
  for (size_t k = 0; k < nforms; k++) {
    const int kind = kinds ? kinds[k] : WHIR_FORM_MLE;
    WHIR_REQUIRE(kind == WHIR_FORM_MLE || kind == WHIR_FORM_COVECTOR || kind == WHIR_FORM_UNIVARIATE,
                 "prove: unknown linear form kind");
    if (kind == WHIR_FORM_COVECTOR) ncov++;
    else nmle++;  // takes words from `points`
  }
  WHIR_REQUIRE(nmle == 0 || points, "prove: null points");
  WHIR_REQUIRE(ncov == 0 || covectors, "prove: null covectors");
  for (size_t q = 0; q < ncov; q++) WHIR_REQUIRE(covectors[q], "prove: null covector");
  std::vector<const DevWitness*> ws(nw);
  for (size_t i = 0; i < nw; i++) {
    WHIR_REQUIRE(witnesses[i], "prove: null witness");
    ws[i] = &witnesses[i]->w;
  }
  for (size_t j = 0; j < nvec; j++) WHIR_REQUIRE(vectors[j], "prove: null vector");
  hipStream_t s = thread_stream();
  switch (c->cfg.D) {
    case 1:
      prove_impl<1>(c->cfg, *t->t, vectors, nvec, ws.data(), nw, kinds, nforms, points, covectors, evals, eval_point, s);
      break;
    case 2:
      prove_impl<2>(c->cfg, *t->t, vectors, nvec, ws.data(), nw, kinds, nforms, points, covectors, evals, eval_point, s);
      break;
    case 3:
      prove_impl<3>(c->cfg, *t->t, vectors, nvec, ws.data(), nw, kinds, nforms, points, covectors, evals, eval_point, s);
      break;
  }
  API_END
}

int whir_prove_batch_dev(const whir_config* c, whir_transcript* t, const uint64_t* const* vectors, size_t nvec,
                         const whir_witness* const* witnesses, size_t nw, const uint64_t* points, size_t np,
                         const uint64_t* evals, uint64_t* eval_point) {
  return whir_prove_forms_dev(c, t, vectors, nvec, witnesses, nw, nullptr, np, points, nullptr, evals, eval_point);
}

int whir_prove_dev(const whir_config* c, whir_transcript* t, const uint64_t* vector, const whir_witness* w,
                   const uint64_t* points, size_t np, const uint64_t* evals, uint64_t* eval_point) {
  return whir_prove_batch_dev(c, t, &vector, 1, &w, 1, points, np, evals, eval_point);
}

void whir_witness_free(whir_witness* w) { delete w; }

int whir_mle_evaluate_dev(int D, const uint64_t* vector, size_t nv, const uint64_t* point, uint64_t* out) {
  API_BEGIN
  WHIR_REQUIRE(D >= 1 && D <= 3, "degree must be 1, 2 or 3");
  hipStream_t s = thread_stream();
  size_t n = (size_t)1 << nv;
  DevBuf eq(n * D * 8, s), lv(n * D * 8, s), o(D * 8, s);
  fill_zero(eq.ptr, n * D * 8, s);
  uint64_t one[3] = {1, 0, 0};
  eval_eq_dev(D, eq.as<uint64_t>(), point, nv, one, s);
  lift_dev(D, vector, n, lv.as<uint64_t>(), s);
  dot_dev(D, lv.as<uint64_t>(), eq.as<uint64_t>(), n, o.as<uint64_t>(), s);
  WHIR_HIP(hipMemcpyAsync(out, o.ptr, D * 8, hipMemcpyDeviceToHost, s));
  WHIR_HIP(hipStreamSynchronize(s));
  API_END
}

}  // extern "C"
