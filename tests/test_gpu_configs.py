"""The benchmark configurations end to end on the GPU, each proof compared byte for
byte with the oracle prover's (the threaded C oracle under oracle/whir_protocol.py),
not only accepted by the verifier: a wrong codeword row that no query opens changes
the Merkle root and so every later transcript byte, which acceptance cannot see.

  C1  d = 20, Goldilocks2, rate 1/2, folding 4, Blake3, 100-bit list decoding, 1 point
  C2  d = 24, Goldilocks2, rate 1/4, folding 4, Blake3, 100-bit list decoding, 2 points
  C3  d = 26, Goldilocks3, rate 1/2, folding 4, SHA3, 100-bit, the configuration's own
      PoW (query-phase grinding), PCS mode with 1 point and LDT mode with none
  C4  64 independent instances, Goldilocks2, Blake3, rate 1/2, 100-bit,
      num_variables ~ U{16..22}, 1-4 points, through the executor: every proof equals
      the same proof made alone, and the smallest and largest equal the oracle's
  C5  Poisson arrivals (num_variables log-uniform in [14, 24], 1-4 points, rate 1/2
      or 1/4) served by the executor: every served proof equals its plain proof

C1, C2 and C4 disable PoW as the benchmark does (at 100 bits over Goldilocks2 the
fold PoW asks for 31-55 bits per sumcheck round; src/protocols/whir/mod.rs:107-118
is the reference's own switch).  The reference CLI's shapes: src/bin/main.rs:152-190.
"""
import hashlib
import random
import threading
import time

import numpy as np
import pytest

import whir_amd as W
from oracle import oracle as O
from oracle import whir_protocol as WP
from whir_amd import prover as GP

pytestmark = pytest.mark.gpu


def params(D, engine, lir, sec=100, pow_bits=20):
    return dict(security_level=sec, pow_bits=pow_bits, initial_folding_factor=4, folding_factor=4,
                unique_decoding=False, starting_log_inv_rate=lir, hash_engine=engine)


def uniform(rng, n):
    return rng.integers(0, O.P, size=n, dtype=np.uint64)


def oracle_prove(nv, D, prm, vec, points, pid, sid, disable_pow):
    cfg = WP.WhirConfig.new(1 << nv, WP.ProtocolParameters(**prm), D)
    if disable_pow:
        cfg.disable_pow()
    evals = [WP.mle_evaluate_base(D, vec, pt) for pt in points]
    t = WP.Transcript(pid, sid)
    w = WP.commit(t, cfg, vec)
    ep, _ = WP.prove(t, cfg, vec, w, points, evals)
    return cfg, evals, (bytes(t.narg), bytes(t.hints)), np.concatenate(ep) if ep else np.zeros(0, np.uint64)


def gpu_prove(cfg, nv, D, dv, points, pid, sid):
    flat = [np.concatenate(pt) for pt in points]
    evals = [GP.mle_evaluate(D, dv, nv, p) for p in flat]
    t = GP.Transcript(pid, sid)
    w = GP.commit(cfg, t, dv)
    ep = GP.prove(cfg, t, dv, w, np.concatenate(flat) if flat else np.zeros(0, np.uint64),
                  np.concatenate(evals) if evals else np.zeros(0, np.uint64))
    return evals, t.proof(), ep


SINGLE = [
    # name, nv, D, engine, log_inv_rate, points, disable_pow
    ("C1", 20, 2, W.HASH_BLAKE3, 1, 1, True),
    ("C2", 24, 2, W.HASH_BLAKE3, 2, 2, True),
    ("C3-PCS", 26, 3, W.HASH_SHA3, 1, 1, False),
    ("C3-LDT", 26, 3, W.HASH_SHA3, 1, 0, False),
]


@pytest.mark.timeout(900)
@pytest.mark.parametrize("name,nv,D,engine,lir,npts,dis", SINGLE, ids=[c[0] for c in SINGLE])
def test_config_proof_equals_oracle(name, nv, D, engine, lir, npts, dis):
    print(f"[{name}] d={nv} D={D} engine={engine} rate=1/{1 << lir} points={npts} pow={'off' if dis else 'on'}",
          flush=True)
    rng = np.random.default_rng(nv * 10 + D)
    vec = uniform(rng, 1 << nv)
    r = random.Random(nv)
    points = [[np.array([r.randrange(O.P) for _ in range(D)], dtype=np.uint64) for _ in range(nv)]
              for _ in range(npts)]
    prm = params(D, engine, lir)
    pid, sid = WP.protocol_ids(f"config {name}")
    gcfg = GP.Config(1 << nv, GP.ProtocolParameters(degree=D, **prm))
    if dis:
        gcfg.disable_pow()
    dv = W.DeviceBuffer.from_array(vec)
    gevals, gproof, gep = gpu_prove(gcfg, nv, D, dv, points, pid, sid)
    dv.close()
    cfg, oevals, oproof, oep = oracle_prove(nv, D, prm, vec, points, pid, sid, dis)
    for a, b in zip(gevals, oevals):
        assert np.array_equal(a, b), f"{name}: MLE evaluation differs"
    assert gproof[0] == oproof[0], f"{name}: narg differs from the oracle prover's"
    assert gproof[1] == oproof[1], f"{name}: hints differ from the oracle prover's"
    assert np.array_equal(gep, oep), f"{name}: folding randomness differs"
    WP.verify(WP.TranscriptReader(pid, sid, *gproof), cfg, oevals, points)


def _ids(tag, k):
    return (hashlib.sha3_512(f"{tag} pid {k}".encode()).digest(), hashlib.sha3_256(f"{tag} sid {k}".encode()).digest())


class Inst:
    def __init__(self, tag, k, nv, lir, npts, rng, cache, vec=None, dv=None):
        key = (nv, lir)
        if key not in cache:
            cache[key] = GP.Config(1 << nv, GP.ProtocolParameters(degree=2, **params(2, W.HASH_BLAKE3, lir)))
            cache[key].disable_pow()
        self.cfg, self.nv, self.lir = cache[key], nv, lir
        self.vec = uniform(rng, 1 << nv) if vec is None else vec
        self.dv = W.DeviceBuffer.from_array(self.vec) if dv is None else dv
        self.pts = uniform(rng, npts * nv * 2)
        self.evals = np.concatenate([GP.mle_evaluate(2, self.dv, nv, self.pts[k2 * nv * 2:(k2 + 1) * nv * 2])
                                     for k2 in range(npts)])
        self.npts = npts
        self.pid, self.sid = _ids(tag, k)

    def plain(self):
        t = GP.Transcript(self.pid, self.sid)
        w = GP.commit(self.cfg, t, self.dv)
        GP.prove(self.cfg, t, self.dv, w, self.pts, self.evals)
        return t.proof()

    def oracle(self):
        points = [[self.pts[(k * self.nv + i) * 2:(k * self.nv + i + 1) * 2] for i in range(self.nv)]
                  for k in range(self.npts)]
        return oracle_prove(self.nv, 2, params(2, W.HASH_BLAKE3, self.lir), self.vec, points, self.pid, self.sid,
                            True)[2]


@pytest.mark.timeout(600)
def test_c4_64_instances_through_the_executor():
    rng = np.random.default_rng(64)
    cache = {}
    insts = [Inst("C4", k, int(rng.integers(16, 23)), 1, int(rng.integers(1, 5)), rng, cache) for k in range(64)]
    print(f"[C4] 64 instances, num_variables {sorted(i.nv for i in insts)}", flush=True)
    ex = GP.Executor(workers=12)
    try:
        jobs = []
        for i in insts:
            t = GP.Transcript(i.pid, i.sid)
            jobs.append((ex.submit(i.cfg, t, i.dv, i.pts, i.evals), t))
        got = []
        for jid, t in jobs:
            ex.wait(jid)
            got.append(t.proof())
    finally:
        ex.close()
    for k, (i, g) in enumerate(zip(insts, got)):
        assert g == i.plain(), f"C4 instance {k} (d={i.nv}) differs from its plain proof"
    small = min(range(64), key=lambda k: insts[k].nv)
    large = max(range(64), key=lambda k: insts[k].nv)
    for k in (small, large):
        assert got[k] == insts[k].oracle(), f"C4 instance {k} (d={insts[k].nv}) differs from the oracle's proof"


@pytest.mark.timeout(600)
def test_c5_poisson_serving_every_proof_equals_plain():
    rng = np.random.default_rng(5)
    r = np.random.default_rng(6)
    cache = {}
    vecs = {d: uniform(rng, 1 << d) for d in range(14, 25)}
    dvs = {d: W.DeviceBuffer.from_array(v) for d, v in vecs.items()}
    seconds, qps = 3.0, 300.0
    arrivals = np.cumsum(r.exponential(1.0 / qps, size=int(qps * seconds * 1.5) + 16))
    arrivals = arrivals[arrivals < seconds]
    reqs = []
    for k in range(len(arrivals)):
        d = int(np.clip(np.floor(np.exp(r.uniform(np.log(14), np.log(25)))), 14, 24))
        reqs.append(Inst("C5", k, d, int(r.integers(1, 3)), int(r.integers(1, 5)), r, cache, vecs[d], dvs[d]))
    print(f"[C5] {len(reqs)} requests over {seconds} s", flush=True)
    ex = GP.Executor(workers=8)
    results = {}
    jobs = []
    cv = threading.Condition()

    def waiter():
        k = 0
        while True:
            with cv:
                while k >= len(jobs):
                    cv.wait()
                item = jobs[k]
            if item is None:
                return
            jid, t, idx = item
            ex.wait(jid)
            results[idx] = t.proof()
            k += 1

    th = threading.Thread(target=waiter)
    th.start()
    try:
        t0 = time.perf_counter()
        for idx, (inst, at) in enumerate(zip(reqs, arrivals)):
            delay = at - (time.perf_counter() - t0)
            if delay > 0:
                time.sleep(delay)
            t = GP.Transcript(inst.pid, inst.sid)
            jid = ex.submit(inst.cfg, t, inst.dv, inst.pts, inst.evals)
            with cv:
                jobs.append((jid, t, idx))
                cv.notify()
        with cv:
            jobs.append(None)
            cv.notify()
        th.join()
    finally:
        ex.close()
    assert len(results) == len(reqs)
    for idx, inst in enumerate(reqs):
        assert results[idx] == inst.plain(), f"C5 request {idx} (d={inst.nv}) differs from its plain proof"
