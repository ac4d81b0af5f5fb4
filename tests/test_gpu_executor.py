"""Batched / served proving (whir_executor_*): every proof made in the mix equals
the same proof made alone, byte for byte; admission by HBM budget; failures stay
with their own job.

The plain path here (whir_commit_dev + whir_prove_dev on the calling thread) is
itself pinned to the oracle prover by tests/test_gpu_prover.py; one case of the
mix is also checked against the oracle directly.
"""
import hashlib

import numpy as np
import pytest

import whir_amd as W
from oracle import oracle as O
from oracle import whir_protocol as WP
from whir_amd import prover as GP

pytestmark = pytest.mark.gpu


def instance(seed, nv, D=2, lir=1, engine=W.HASH_BLAKE3, num_points=2, pow_bits=0):
    rng = np.random.default_rng(seed)
    vec = rng.integers(0, O.P, size=1 << nv, dtype=np.uint64)
    pts = rng.integers(0, O.P, size=num_points * nv * D, dtype=np.uint64)
    params = GP.ProtocolParameters(degree=D, security_level=32, pow_bits=pow_bits, initial_folding_factor=4,
                                   folding_factor=4, starting_log_inv_rate=lir, hash_engine=engine)
    cfg = GP.Config(1 << nv, params)
    if pow_bits == 0:
        cfg.disable_pow()
    dv = W.DeviceBuffer.from_array(vec)
    evals = np.concatenate([GP.mle_evaluate(D, dv, nv, pts[k * nv * D:(k + 1) * nv * D])
                            for k in range(num_points)]) if num_points else np.zeros(0, np.uint64)
    pid = hashlib.sha3_512(f"executor test {seed} {nv} {D} {lir}".encode()).digest()
    sid = hashlib.sha3_256(f"session {seed}".encode()).digest()
    return dict(seed=seed, nv=nv, D=D, lir=lir, engine=engine, vec=vec, cfg=cfg, dv=dv, pts=pts, evals=evals,
                pid=pid, sid=sid, params=params)


def plain(inst):
    t = GP.Transcript(inst["pid"], inst["sid"])
    w = GP.commit(inst["cfg"], t, inst["dv"])
    ep = GP.prove(inst["cfg"], t, inst["dv"], w, inst["pts"], inst["evals"])
    return t.proof(), ep


MIX = [(11, 12), (12, 10), (13, 16), (14, 11), (15, 14), (16, 13), (17, 10), (18, 15), (19, 12), (20, 16),
       (21, 10), (22, 14)]


@pytest.fixture(scope="module")
def mix():
    insts = [instance(seed, nv, num_points=1 + seed % 3, lir=1 + seed % 2) for seed, nv in MIX]
    return insts, [plain(i) for i in insts]


@pytest.mark.parametrize("workers", [1, 4, 8])
def test_executor_proofs_equal_plain(mix, workers):
    insts, expected = mix
    ex = GP.Executor(workers=workers)
    try:
        jobs = []
        for inst in insts:
            t = GP.Transcript(inst["pid"], inst["sid"])
            jobs.append((ex.submit(inst["cfg"], t, inst["dv"], inst["pts"], inst["evals"]), t, inst))
        for (jid, t, inst), (exp_proof, exp_ep) in zip(jobs, expected):
            times, ep = ex.wait(jid)
            assert times[0] <= times[1] <= times[2]
            assert t.proof() == exp_proof, f"proof differs for instance seed={inst['seed']} nv={inst['nv']}"
            assert np.array_equal(ep, exp_ep)
    finally:
        ex.close()


def test_executor_large_job_cap(mix, monkeypatch):
    """A job of 2^23 elements among small ones: while it runs the executor admits at
    most the capped number of jobs (here 2, so the cap is what orders the mix); every
    proof still equals its plain one and the executor drains."""
    monkeypatch.setenv("WHIR_EXEC_LARGE_CAP", "2")
    monkeypatch.setenv("WHIR_EXEC_LARGE_LOG", "23")
    insts, expected = mix
    big = instance(99, 23, num_points=1)
    big_exp = plain(big)
    ex = GP.Executor(workers=8)
    try:
        jobs = []
        for k, (inst, exp) in enumerate(zip([big] + insts, [big_exp] + expected)):
            t = GP.Transcript(inst["pid"], inst["sid"])
            jobs.append((ex.submit(inst["cfg"], t, inst["dv"], inst["pts"], inst["evals"]), t, inst, exp))
        spans = []
        for jid, t, inst, (exp_proof, exp_ep) in jobs:
            times, ep = ex.wait(jid)
            spans.append((times[1], times[2]))
            assert t.proof() == exp_proof, f"proof differs for instance seed={inst['seed']} nv={inst['nv']}"
            assert np.array_equal(ep, exp_ep)
    finally:
        ex.close()
    # a small job that started while the large one ran found nothing else running
    (s_big, e_big), small = spans[0], spans[1:]
    for j, (s_j, _) in enumerate(small):
        if s_big < s_j < e_big:
            others = sum(1 for i, (s_i, e_i) in enumerate(small) if i != j and s_i <= s_j < e_i)
            assert others == 0, "more jobs ran beside the large one than the cap allows"


def test_executor_mix_matches_oracle_and_verifies(mix):
    """One instance of the mix against the oracle prover and verifier."""
    insts, expected = mix
    inst = insts[0]
    (narg, hints), _ = expected[0]
    ocfg = WP.WhirConfig.new(1 << inst["nv"], WP.ProtocolParameters(
        security_level=32, pow_bits=0, initial_folding_factor=4, folding_factor=4, unique_decoding=False,
        starting_log_inv_rate=inst["lir"], hash_engine=inst["engine"]), inst["D"])
    ocfg.disable_pow()
    nv, D = inst["nv"], inst["D"]
    pts = [[inst["pts"][(k * nv + i) * D:(k * nv + i + 1) * D] for i in range(nv)]
           for k in range(inst["evals"].size // D)]
    evs = [inst["evals"][k * D:(k + 1) * D] for k in range(len(pts))]
    t = WP.Transcript(inst["pid"], inst["sid"])
    w = WP.commit(t, ocfg, inst["vec"])
    WP.prove(t, ocfg, inst["vec"], w, pts, evs)
    assert (bytes(t.narg), bytes(t.hints)) == (narg, hints)
    WP.verify(WP.TranscriptReader(inst["pid"], inst["sid"], narg, hints), ocfg, evs, pts)


def test_executor_budget_serializes(mix):
    """A 1-byte budget admits one job at a time (a job larger than the budget runs
    only when nothing else does): runs never overlap, results unchanged."""
    insts, expected = mix
    ex = GP.Executor(workers=4, hbm_budget=1)
    try:
        jobs = []
        for inst in insts[:6]:
            t = GP.Transcript(inst["pid"], inst["sid"])
            jobs.append((ex.submit(inst["cfg"], t, inst["dv"], inst["pts"], inst["evals"]), t))
        spans = []
        for (jid, t), (exp_proof, _) in zip(jobs, expected):
            times, _ = ex.wait(jid)
            spans.append((times[1], times[2]))
            assert t.proof() == exp_proof
        spans.sort()
        for (s0, e0), (s1, e1) in zip(spans, spans[1:]):
            assert s1 >= e0, "jobs overlapped under a budget that fits none"
    finally:
        ex.close()


def test_executor_budget_admits_by_memory(mix):
    """A budget that fits exactly two of the largest jobs never runs more than
    that much at once."""
    insts, expected = mix
    mems = [GP.config_memory(i["cfg"]) for i in insts]
    budget = 2 * max(mems)
    ex = GP.Executor(workers=8, hbm_budget=budget)
    try:
        jobs = []
        for inst in insts:
            t = GP.Transcript(inst["pid"], inst["sid"])
            jobs.append((ex.submit(inst["cfg"], t, inst["dv"], inst["pts"], inst["evals"]), t))
        spans = []
        for k, ((jid, t), (exp_proof, _)) in enumerate(zip(jobs, expected)):
            times, _ = ex.wait(jid)
            spans.append((times[1], times[2], mems[k]))
            assert t.proof() == exp_proof
        for s, _, _ in spans:
            live = sum(m for (s2, e2, m) in spans if s2 <= s < e2)
            assert live <= budget
    finally:
        ex.close()


def test_executor_commit_only_matches_plain_commit(mix):
    insts, _ = mix
    ex = GP.Executor(workers=4)
    try:
        for inst in insts[:4]:
            t0 = GP.Transcript(inst["pid"], inst["sid"])
            GP.commit(inst["cfg"], t0, inst["dv"])
            t1 = GP.Transcript(inst["pid"], inst["sid"])
            jid = ex.submit(inst["cfg"], t1, inst["dv"], kind=GP.Executor.COMMIT_ONLY)
            ex.wait(jid)
            assert t1.proof() == t0.proof()
    finally:
        ex.close()


class _FailingTranscript:
    """A transcript whose callbacks (whir_transcript_vtable) refuse the first
    prover message: the job fails on the host after its commit kernels ran."""

    def __init__(self):
        import ctypes
        c_int, vp, u64p, u8p, sz = (ctypes.c_int, ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint64),
                                    ctypes.POINTER(ctypes.c_uint8), ctypes.c_size_t)
        F = ctypes.CFUNCTYPE

        class VT(ctypes.Structure):
            _fields_ = [("ctx", vp),
                        ("prover_field", F(c_int, vp, c_int, u64p, sz)),
                        ("verifier_field", F(c_int, vp, c_int, u64p)),
                        ("verifier_bytes", F(c_int, vp, u8p, sz)),
                        ("prover_hash", F(c_int, vp, u8p)),
                        ("prover_u64", F(c_int, vp, ctypes.c_uint64)),
                        ("hint_hash", F(c_int, vp, u8p)),
                        ("hint_field_vec", F(c_int, vp, c_int, u64p, sz))]

        fail = lambda *a: 1  # noqa: E731
        self.vt = VT(None, *[f(fail) for (_, f) in VT._fields_[1:]])
        h = vp()
        W._call("whir_transcript_from_vtable", ctypes.byref(self.vt), ctypes.byref(h))
        self.h = h.value

    def __del__(self):
        try:
            W.lib().whir_transcript_free(self.h)
        except Exception:
            pass


def test_executor_failure_stays_with_its_job(mix):
    """A job whose transcript refuses its messages fails with that message; the
    jobs around it are unaffected."""
    insts, expected = mix
    ex = GP.Executor(workers=2)
    try:
        a, b = insts[1], insts[3]
        ta, tb = GP.Transcript(a["pid"], a["sid"]), GP.Transcript(b["pid"], b["sid"])
        tx = _FailingTranscript()
        ja = ex.submit(a["cfg"], ta, a["dv"], a["pts"], a["evals"])
        jx = ex.submit(a["cfg"], tx, a["dv"], a["pts"], a["evals"])
        jb = ex.submit(b["cfg"], tb, b["dv"], b["pts"], b["evals"])
        ex.wait(ja)
        with pytest.raises(W.WhirError, match="transcript callback failed: prover_hash"):
            ex.wait(jx)
        ex.wait(jb)
        assert ta.proof() == expected[1][0]
        assert tb.proof() == expected[3][0]
    finally:
        ex.close()


def batch_instance(seed, nv, B, NW, D=2, num_points=2):
    rng = np.random.default_rng(seed)
    vecs = [rng.integers(0, O.P, size=1 << nv, dtype=np.uint64) for _ in range(B * NW)]
    pts = rng.integers(0, O.P, size=num_points * nv * D, dtype=np.uint64)
    params = GP.ProtocolParameters(degree=D, security_level=32, pow_bits=0, initial_folding_factor=4,
                                   folding_factor=4, starting_log_inv_rate=1, batch_size=B)
    cfg = GP.Config(1 << nv, params).disable_pow()
    dvs = [W.DeviceBuffer.from_array(v) for v in vecs]
    evals = np.array([[GP.mle_evaluate(D, dv, nv, pts[k * nv * D:(k + 1) * nv * D]) for dv in dvs]
                      for k in range(num_points)], dtype=np.uint64)
    pid = hashlib.sha3_512(f"executor batch {seed} {nv} {B} {NW}".encode()).digest()
    sid = hashlib.sha3_256(f"batch session {seed}".encode()).digest()
    return dict(nv=nv, B=B, NW=NW, D=D, vecs=vecs, cfg=cfg, dvs=dvs, pts=pts, evals=evals, pid=pid, sid=sid)


def plain_batch(bi):
    t = GP.Transcript(bi["pid"], bi["sid"])
    B = bi["B"]
    ws = [GP.commit_batch(bi["cfg"], t, bi["dvs"][w * B:(w + 1) * B]) for w in range(bi["NW"])]
    ep = GP.prove_batch(bi["cfg"], t, bi["dvs"], ws, bi["pts"], bi["evals"])
    return t.proof(), ep


def test_executor_batch_jobs_equal_plain(mix):
    """Multi-vector jobs (several commitments of batch_size vectors in one proof)
    interleaved with single-vector jobs: every proof equals the one made alone."""
    insts, expected = mix
    bis = [batch_instance(30 + k, nv, B, NW) for k, (nv, B, NW) in enumerate([(12, 2, 1), (10, 1, 3), (13, 2, 2),
                                                                              (11, 3, 1)])]
    bexp = [plain_batch(bi) for bi in bis]
    ex = GP.Executor(workers=4)
    try:
        jobs = []
        for k in range(4):
            bi = bis[k]
            tb = GP.Transcript(bi["pid"], bi["sid"])
            jobs.append(("b", k, ex.submit_batch(bi["cfg"], tb, bi["dvs"], bi["pts"], bi["evals"]), tb))
            inst = insts[k]
            ts = GP.Transcript(inst["pid"], inst["sid"])
            jobs.append(("s", k, ex.submit(inst["cfg"], ts, inst["dv"], inst["pts"], inst["evals"]), ts))
        for kind, k, jid, t in jobs:
            _, ep = ex.wait(jid)
            if kind == "b":
                assert t.proof() == bexp[k][0], f"batch job {k} differs"
                assert np.array_equal(ep, bexp[k][1])
            else:
                assert t.proof() == expected[k][0]
    finally:
        ex.close()


def test_executor_batch_job_rejects_a_partial_batch():
    bi = batch_instance(40, 10, 2, 1)
    ex = GP.Executor(workers=1)
    try:
        with pytest.raises(W.WhirError, match="multiple of batch_size"):
            ex.submit_batch(bi["cfg"], GP.Transcript(bi["pid"], bi["sid"]), bi["dvs"][:1],
                            bi["pts"], bi["evals"][:, :1])
    finally:
        ex.close()


def _hip_used_bytes():
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    assert hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return total.value - free.value


def test_executor_device_memory_stays_within_budget(mix):
    """Mixed sizes under a budget of about two of the largest jobs: the device memory
    in use (hipMemGetInfo, sampled while the jobs run) stays within the budget plus a
    small allowance for pool granularity, because the buffers streams keep cached
    count against the budget and idle workers' caches go back to the device when a
    job needs the room.  Every proof is unchanged."""
    import threading
    import time
    insts, expected = mix
    mems = [GP.config_memory(i["cfg"]) for i in insts]
    budget = 2 * max(mems) + max(mems) // 2
    W.synchronize()
    base = _hip_used_bytes()
    peak = [base]
    stop = threading.Event()

    def sample():
        while not stop.is_set():
            peak[0] = max(peak[0], _hip_used_bytes())
            time.sleep(0.0005)

    th = threading.Thread(target=sample)
    th.start()
    ex = GP.Executor(workers=6, hbm_budget=budget)
    try:
        jobs = []
        for rep in range(3):
            for inst, exp in zip(insts[::-1], expected[::-1]):
                t = GP.Transcript(inst["pid"], inst["sid"])
                jobs.append((ex.submit(inst["cfg"], t, inst["dv"], inst["pts"], inst["evals"]), t, exp))
        for jid, t, (exp_proof, _) in jobs:
            ex.wait(jid)
            assert t.proof() == exp_proof
    finally:
        ex.close()
        stop.set()
        th.join()
    allowance = (64 << 20) + budget // 20
    assert peak[0] - base <= budget + allowance, \
        f"device memory grew by {(peak[0] - base) >> 20} MiB under a {budget >> 20} MiB budget"


def test_executor_head_of_queue_is_not_passed_forever(mix, monkeypatch):
    """A job that fits the budget only when nothing else runs, submitted ahead of a
    stream of small jobs: at most WHIR_EXEC_MAX_PASS of them start before it (without
    the bound every one of them would), and every proof is unchanged."""
    monkeypatch.setenv("WHIR_EXEC_MAX_PASS", "3")
    insts, expected = mix
    big, big_exp = insts[-1], expected[-1]
    small = [(insts[k % 4], expected[k % 4]) for k in range(40)]
    budget = GP.config_memory(big["cfg"])
    ex = GP.Executor(workers=4, hbm_budget=budget)
    try:
        first = small[0]
        t0 = GP.Transcript(first[0]["pid"], first[0]["sid"])
        jobs = [(ex.submit(first[0]["cfg"], t0, first[0]["dv"], first[0]["pts"], first[0]["evals"]), t0, first[1],
                 "s")]
        tb = GP.Transcript(big["pid"], big["sid"])
        jobs.append((ex.submit(big["cfg"], tb, big["dv"], big["pts"], big["evals"]), tb, big_exp, "big"))
        for inst, exp in small[1:]:
            t = GP.Transcript(inst["pid"], inst["sid"])
            jobs.append((ex.submit(inst["cfg"], t, inst["dv"], inst["pts"], inst["evals"]), t, exp, "s"))
        starts = []
        for jid, t, (exp_proof, _), tag in jobs:
            times, _ = ex.wait(jid)
            starts.append((times[1], tag))
            assert t.proof() == exp_proof
    finally:
        ex.close()
    big_start = [s for s, tag in starts if tag == "big"][0]
    passed = sum(1 for s, tag in starts[2:] if tag == "s" and s < big_start)
    assert passed <= 3, f"{passed} later jobs started before the waiting head of the queue"
