// Batched and served proving: a pool of worker threads, each proving on its own
// HIP stream, that admits jobs by an HBM budget.
//
// The reference proves one instance per call (src/bin/main.rs:152-190 times one
// commit + prove); this layer has no counterpart there.  Each WHIR proof is a
// chain of ~50 Fiat-Shamir round trips between host and device, so a single
// proof leaves the GPU idle between its rounds and small instances are bound by
// launch and round-trip latency, not by the device.  Running many proofs at
// once, each on its own stream, overlaps one proof's host work with the others'
// kernels; the device interleaves their launches.  Every job uses the same
// prover entry points as a plain call (whir_commit_dev + whir_prove_dev), so a
// proof made in the mix is byte-identical to the same proof made alone.
//
// Admission: a job's peak device memory is estimated from its configuration
// (job_memory); a worker starts the first queued job that fits in what the
// budget has left, so small requests pass a large one that is waiting for
// memory, but only a bounded number of times: once the job at the head of the
// queue has been passed max_pass times, no other job starts until it has.  What
// the budget has left counts the running jobs' estimates and the buffers the
// streams keep cached for reuse; when a job fits only without the idle workers'
// caches, those are handed back to the device first.  A job larger than the whole
// budget runs only when nothing else does.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/whir_amd.h"
#include "common.h"
#include "config.h"
#include "handles.h"

using namespace whir;

namespace whir {
namespace {

// Peak device bytes of one commit + prove: the covector and folded vector, the
// initial codeword (plus its NTT scratch) and Merkle nodes, which the caller keeps
// for the whole proof, and the largest pair of consecutive round witnesses (a
// round's witness is committed while the previous round's, which that round opens,
// is still alive).  With several commitments (multi-vector jobs) every initial
// witness stays alive until the first round has opened it.
size_t job_memory(const WhirCfg& c, size_t commitments = 1) {
  auto wit = [](const IrsCfg& irs, int Ds) {
    const size_t N = irs.codeword_length, M = irs.num_messages();
    return 2 * M * N * (size_t)Ds * 8 + 64 * N;
  };
  const size_t n = c.initial.vector_size;
  size_t mx = 0;
  for (size_t k = 0; k < c.rounds.size(); k++) {
    size_t w = wit(c.rounds[k].irs, c.D);
    if (k) w += wit(c.rounds[k - 1].irs, c.D);
    mx = std::max(mx, w);
  }
  return 2 * n * (size_t)c.D * 8 + commitments * wit(c.initial, 1) + mx;
}

double now_s(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

struct Job {
  uint64_t id = 0;
  int kind = 0;  // 0: commit + prove, 1: commit only, 2: batch_size-vector commitments + one proof,
                 // 3: prove only (the witness is given)
  const whir_config* cfg = nullptr;
  whir_transcript* t = nullptr;
  const uint64_t* vector = nullptr;
  std::vector<const uint64_t*> vectors;  // kind 2
  const whir_witness* witness = nullptr;  // kind 3
  std::vector<uint64_t> points, evals;
  size_t np = 0;
  size_t mem = 0;
  // results
  bool done = false;
  int status = 0;
  std::string error;
  std::vector<uint64_t> eval_point;
  double t_submit = 0, t_start = 0, t_end = 0;
  int passed = 0;  // times a later job started while this one waited at the head of the queue
};

}  // namespace
}  // namespace whir

struct whir_executor {
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<Job*> queue;
  std::unordered_map<uint64_t, std::unique_ptr<Job>> jobs;
  std::vector<std::thread> workers;
  size_t budget = 0, in_use = 0;
  int running = 0;
  int running_large = 0;  // running jobs of vectors of at least 2^large_log elements
  // While a large job runs, at most large_cap jobs run: the large job keeps more of the
  // device, which cuts the tail latency of a mixed request stream (serving at 750
  // requests/s, d log-uniform [14, 24], 8 workers: p99 13.6 -> 12.6-13.1 ms, p50 +2-4%;
  // vectors of 2^23 and up: the 2^16..2^22 batch keeps its throughput).
  int large_log = 23, large_cap = 6;
  // The job at the head of the queue may be passed by later jobs this many times
  // (WHIR_EXEC_MAX_PASS; 0 = strict order).
  int max_pass = 16;
  uint64_t next_id = 1;
  bool stop = false;
  int device = 0;
  std::vector<hipStream_t> streams;  // per worker, set when the worker starts
  std::vector<char> busy;            // per worker: running a job
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();

  bool is_large(const Job* j) const { return j->cfg->cfg.initial.vector_size >= ((size_t)1 << large_log); }

  // Cached bytes of the idle workers' streams (buffers no running job can be using).
  size_t idle_cached() const {
    size_t b = 0;
    for (size_t i = 0; i < streams.size(); i++)
      if (!busy[i] && streams[i]) b += pool_cached_bytes(streams[i]);
    return b;
  }
  void trim_idle() {
    for (size_t i = 0; i < streams.size(); i++)
      if (!busy[i] && streams[i]) pool_trim(streams[i]);
  }

  // Whether j can start now (called with mu held): always when nothing runs, else
  // when the running estimates, the cached buffers and j fit the budget, handing the
  // idle workers' caches back to the device when only that makes it fit.
  bool admit(const Job* j) {
    const size_t cached = pool_cached_bytes(nullptr);
    if (in_use + j->mem + cached <= budget) return true;
    const size_t idle = idle_cached();
    if (running == 0 || in_use + j->mem + (cached - std::min(cached, idle)) <= budget) {
      if (idle) trim_idle();
      return true;
    }
    return false;
  }

  // First queued job that fits the remaining budget (or any job when idle).  With a
  // large-job cap, no job starts while a large one runs beside large_cap - 1 others.
  Job* pick() {
    if (large_cap > 0 && running_large > 0 && running >= large_cap) return nullptr;
    if (queue.empty()) return nullptr;
    Job* head = queue.front();
    if (admit(head)) {
      queue.pop_front();
      return head;
    }
    if (head->passed >= max_pass) return nullptr;  // the head has waited long enough: drain for it
    for (auto it = queue.begin() + 1; it != queue.end(); ++it) {
      Job* j = *it;
      if (admit(j)) {
        head->passed++;
        queue.erase(it);
        return j;
      }
    }
    return nullptr;
  }

  void run(Job* j) {
    int st = 0;
    std::string err;
    std::vector<uint64_t> ep;
    const WhirCfg& c = j->cfg->cfg;
    size_t rounds = c.initial_sumcheck.num_rounds + c.final_sumcheck.num_rounds;
    for (const RoundCfg& r : c.rounds) rounds += r.sumcheck.num_rounds;
    if (j->kind == 2) {
      // Commit each group of batch_size vectors in order, then prove them together.
      const size_t B = c.initial.num_vectors, nw = j->vectors.size() / B;
      std::vector<whir_witness*> ws;
      for (size_t g = 0; g < nw && st == 0; g++) {
        whir_witness* w = nullptr;
        st = whir_commit_batch_dev(j->cfg, j->t, j->vectors.data() + g * B, B, &w);
        if (w) ws.push_back(w);
      }
      if (st == 0) {
        ep.assign(rounds * (size_t)c.D, 0);
        std::vector<const whir_witness*> cws(ws.begin(), ws.end());
        st = whir_prove_batch_dev(j->cfg, j->t, j->vectors.data(), j->vectors.size(), cws.data(), cws.size(),
                                  j->points.data(), j->np, j->evals.data(), ep.data());
      }
      if (st) err = whir_last_error();
      for (whir_witness* w : ws) whir_witness_free(w);
    } else if (j->kind == 3) {
      ep.assign(rounds * (size_t)c.D, 0);
      st = whir_prove_dev(j->cfg, j->t, j->vector, j->witness, j->points.data(), j->np, j->evals.data(), ep.data());
      if (st) err = whir_last_error();
    } else {
      whir_witness* w = nullptr;
      st = whir_commit_dev(j->cfg, j->t, j->vector, &w);
      if (st == 0 && j->kind == 0) {
        ep.assign(rounds * (size_t)c.D, 0);  // one folding challenge per variable
        st = whir_prove_dev(j->cfg, j->t, j->vector, w, j->points.data(), j->np, j->evals.data(), ep.data());
      } else if (st == 0) {
        st = hipStreamSynchronize(thread_stream()) == hipSuccess ? 0 : 1;
      }
      if (st) err = whir_last_error();
      if (w) whir_witness_free(w);
    }
    std::lock_guard<std::mutex> g(mu);
    j->status = st;
    j->error = err;
    j->eval_point.swap(ep);
  }

  void loop(size_t wi) {
    (void)hipSetDevice(device);
    {
      hipStream_t s = nullptr;
      try {
        s = thread_stream();
      } catch (...) {
      }
      std::lock_guard<std::mutex> g(mu);
      streams[wi] = s;
    }
    for (;;) {
      Job* j = nullptr;
      {
        std::unique_lock<std::mutex> lk(mu);
        cv_work.wait(lk, [&] { return stop || (j = pick()) != nullptr; });
        if (!j) break;  // stop requested and nothing admissible left
        in_use += j->mem;
        running++;
        busy[wi] = 1;
        if (is_large(j)) running_large++;
        j->t_start = now_s(t0);
      }
      run(j);
      {
        std::lock_guard<std::mutex> g(mu);
        j->t_end = now_s(t0);
        j->done = true;
        in_use -= j->mem;
        running--;
        busy[wi] = 0;
        if (is_large(j)) running_large--;
      }
      cv_done.notify_all();
      cv_work.notify_all();
    }
    {
      std::lock_guard<std::mutex> g(mu);
      streams[wi] = nullptr;  // its cache goes with it
    }
    release_thread_resources();
  }
};

extern "C" {

#define EXEC_API_BEGIN try {
#define EXEC_API_END                           \
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

int whir_config_memory(const whir_config* cfg, size_t* bytes) {
  EXEC_API_BEGIN
  WHIR_REQUIRE(cfg && bytes, "config_memory: null argument");
  *bytes = job_memory(cfg->cfg);
  EXEC_API_END
}

int whir_executor_new(int num_workers, size_t hbm_budget, whir_executor** out) {
  EXEC_API_BEGIN
  WHIR_REQUIRE(num_workers >= 1 && num_workers <= 64, "executor: 1..64 workers");
  WHIR_REQUIRE(out, "executor: null output");
  auto e = std::make_unique<whir_executor>();
  WHIR_HIP(hipGetDevice(&e->device));
  if (hbm_budget == 0) {
    size_t free_b = 0, total_b = 0;
    WHIR_HIP(hipMemGetInfo(&free_b, &total_b));
    hbm_budget = free_b / 10 * 8;  // leave a fifth of what is free for tables and other users
  }
  e->budget = hbm_budget;
  // Large-job cap: while a job of at least 2^WHIR_EXEC_LARGE_LOG elements runs, no job
  // starts while WHIR_EXEC_LARGE_CAP jobs run (0 turns the cap off).
  if (const char* v = getenv("WHIR_EXEC_LARGE_CAP")) e->large_cap = atoi(v);
  if (const char* v = getenv("WHIR_EXEC_LARGE_LOG")) e->large_log = atoi(v);
  if (const char* v = getenv("WHIR_EXEC_MAX_PASS")) e->max_pass = atoi(v);
  e->streams.assign(num_workers, nullptr);
  e->busy.assign(num_workers, 0);
  whir_executor* p = e.get();
  for (int i = 0; i < num_workers; i++) e->workers.emplace_back([p, i] { p->loop((size_t)i); });
  *out = e.release();
  EXEC_API_END
}

void whir_executor_free(whir_executor* e) {
  if (!e) return;
  {
    std::unique_lock<std::mutex> lk(e->mu);
    e->cv_done.wait(lk, [&] { return e->queue.empty() && e->running == 0; });
    e->stop = true;
  }
  e->cv_work.notify_all();
  for (auto& t : e->workers) t.join();
  delete e;
}

int whir_executor_submit(whir_executor* e, int kind, const whir_config* cfg, whir_transcript* t,
                         const uint64_t* vector, const uint64_t* points, size_t num_points,
                         const uint64_t* evaluations, uint64_t* job_id) {
  EXEC_API_BEGIN
  WHIR_REQUIRE(e && cfg && t && vector && job_id, "executor_submit: null argument");
  WHIR_REQUIRE(kind == 0 || kind == 1, "executor_submit: kind is 0 (commit + prove) or 1 (commit)");
  auto j = std::make_unique<Job>();
  j->kind = kind;
  j->cfg = cfg;
  j->t = t;
  j->vector = vector;
  j->np = kind == 0 ? num_points : 0;
  if (kind == 0) {
    size_t nv = 0;
    while (((size_t)1 << nv) < cfg->cfg.initial.vector_size) nv++;
    const size_t D = (size_t)cfg->cfg.D;
    if (num_points) WHIR_REQUIRE(points && evaluations, "executor_submit: null points or evaluations");
    j->points.assign(points, points + num_points * nv * D);
    j->evals.assign(evaluations, evaluations + num_points * D);
  }
  j->mem = job_memory(cfg->cfg);
  {
    std::lock_guard<std::mutex> g(e->mu);
    WHIR_REQUIRE(!e->stop, "executor_submit: executor is shutting down");
    j->id = e->next_id++;
    j->t_submit = now_s(e->t0);
    *job_id = j->id;
    e->queue.push_back(j.get());
    e->jobs.emplace(j->id, std::move(j));
  }
  e->cv_work.notify_one();
  EXEC_API_END
}

int whir_executor_submit_prove(whir_executor* e, const whir_config* cfg, whir_transcript* t, const uint64_t* vector,
                               const whir_witness* witness, const uint64_t* points, size_t num_points,
                               const uint64_t* evaluations, uint64_t* job_id) {
  EXEC_API_BEGIN
  WHIR_REQUIRE(e && cfg && t && vector && witness && job_id, "executor_submit_prove: null argument");
  if (num_points) WHIR_REQUIRE(points && evaluations, "executor_submit_prove: null points or evaluations");
  auto j = std::make_unique<Job>();
  j->kind = 3;
  j->cfg = cfg;
  j->t = t;
  j->vector = vector;
  j->witness = witness;
  j->np = num_points;
  size_t nv = 0;
  while (((size_t)1 << nv) < cfg->cfg.initial.vector_size) nv++;
  const size_t D = (size_t)cfg->cfg.D;
  j->points.assign(points, points + num_points * nv * D);
  j->evals.assign(evaluations, evaluations + num_points * D);
  j->mem = job_memory(cfg->cfg);
  {
    std::lock_guard<std::mutex> g(e->mu);
    WHIR_REQUIRE(!e->stop, "executor_submit_prove: executor is shutting down");
    j->id = e->next_id++;
    j->t_submit = now_s(e->t0);
    *job_id = j->id;
    e->queue.push_back(j.get());
    e->jobs.emplace(j->id, std::move(j));
  }
  e->cv_work.notify_one();
  EXEC_API_END
}

int whir_executor_submit_batch(whir_executor* e, const whir_config* cfg, whir_transcript* t,
                               const uint64_t* const* vectors, size_t num_vectors, const uint64_t* points,
                               size_t num_points, const uint64_t* evaluations, uint64_t* job_id) {
  EXEC_API_BEGIN
  WHIR_REQUIRE(e && cfg && t && vectors && job_id, "executor_submit_batch: null argument");
  const size_t B = cfg->cfg.initial.num_vectors;
  WHIR_REQUIRE(num_vectors > 0 && num_vectors % B == 0,
               "executor_submit_batch: the number of vectors must be a multiple of batch_size");
  for (size_t i = 0; i < num_vectors; i++) WHIR_REQUIRE(vectors[i], "executor_submit_batch: null vector");
  if (num_points) WHIR_REQUIRE(points && evaluations, "executor_submit_batch: null points or evaluations");
  auto j = std::make_unique<Job>();
  j->kind = 2;
  j->cfg = cfg;
  j->t = t;
  j->vectors.assign(vectors, vectors + num_vectors);
  j->np = num_points;
  size_t nv = 0;
  while (((size_t)1 << nv) < cfg->cfg.initial.vector_size) nv++;
  const size_t D = (size_t)cfg->cfg.D;
  j->points.assign(points, points + num_points * nv * D);
  j->evals.assign(evaluations, evaluations + num_points * num_vectors * D);
  j->mem = job_memory(cfg->cfg, num_vectors / B);
  {
    std::lock_guard<std::mutex> g(e->mu);
    WHIR_REQUIRE(!e->stop, "executor_submit_batch: executor is shutting down");
    j->id = e->next_id++;
    j->t_submit = now_s(e->t0);
    *job_id = j->id;
    e->queue.push_back(j.get());
    e->jobs.emplace(j->id, std::move(j));
  }
  e->cv_work.notify_one();
  EXEC_API_END
}

int whir_executor_wait(whir_executor* e, uint64_t job_id, double times[3], uint64_t* eval_point) {
  EXEC_API_BEGIN
  WHIR_REQUIRE(e, "executor_wait: null executor");
  std::unique_ptr<Job> j;
  {
    std::unique_lock<std::mutex> lk(e->mu);
    auto it = e->jobs.find(job_id);
    WHIR_REQUIRE(it != e->jobs.end(), "executor_wait: unknown job id");
    Job* p = it->second.get();
    e->cv_done.wait(lk, [&] { return p->done; });
    j = std::move(it->second);
    e->jobs.erase(it);
  }
  if (times) {
    times[0] = j->t_submit;
    times[1] = j->t_start;
    times[2] = j->t_end;
  }
  if (eval_point && !j->eval_point.empty()) std::copy(j->eval_point.begin(), j->eval_point.end(), eval_point);
  if (j->status) throw Error("job " + std::to_string(job_id) + ": " + j->error);
  EXEC_API_END
}

int whir_executor_clock(whir_executor* e, double* seconds) {
  EXEC_API_BEGIN
  WHIR_REQUIRE(e && seconds, "executor_clock: null argument");
  *seconds = now_s(e->t0);
  EXEC_API_END
}

}  // extern "C"
