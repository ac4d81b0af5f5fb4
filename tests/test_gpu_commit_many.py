"""Batched commits (whir_commit_many_dev): many independent instances committed by
shared launches recorded as a HIP graph per batch layout and replayed.  Every
instance's transcript after the batched commit, and the proof made from its witness,
equal those of a plain whir_commit_dev + whir_prove_dev of the same instance; graph
replays with new vectors (and plain work on the same stream between them) each give
the plain results; witnesses from a batch feed prove-only executor jobs."""
import hashlib

import numpy as np
import pytest

import whir_amd as W
from oracle import oracle as O
from whir_amd import prover as GP

pytestmark = pytest.mark.gpu


class Inst:
    def __init__(self, tag, nv, D=2, lir=1, engine=W.HASH_BLAKE3, npts=1, seed=0, k0=4):
        rng = np.random.default_rng(seed)
        self.nv, self.D = nv, D
        self.cfg = GP.Config(1 << nv, GP.ProtocolParameters(degree=D, security_level=32, pow_bits=0,
                                                            initial_folding_factor=k0, folding_factor=4,
                                                            starting_log_inv_rate=lir, hash_engine=engine))
        self.cfg.disable_pow()
        self.vec = rng.integers(0, O.P, size=1 << nv, dtype=np.uint64)
        self.dv = W.DeviceBuffer.from_array(self.vec)
        self.pts = rng.integers(0, O.P, size=npts * nv * D, dtype=np.uint64)
        self.evals = np.concatenate([GP.mle_evaluate(D, self.dv, nv, self.pts[k * nv * D:(k + 1) * nv * D])
                                     for k in range(npts)])
        self.pid = hashlib.sha3_512(f"commit many {tag}".encode()).digest()
        self.sid = hashlib.sha3_256(f"commit many session {tag}".encode()).digest()

    def transcript(self):
        return GP.Transcript(self.pid, self.sid)

    def plain(self):
        t = self.transcript()
        w = GP.commit(self.cfg, t, self.dv)
        after_commit = t.proof()
        GP.prove(self.cfg, t, self.dv, w, self.pts, self.evals)
        return after_commit, t.proof()


def check_batch(insts, expected):
    ts = [i.transcript() for i in insts]
    ws = GP.commit_many([i.cfg for i in insts], ts, [i.dv for i in insts])
    for k, (i, t, w) in enumerate(zip(insts, ts, ws)):
        assert t.proof() == expected[k][0], f"instance {k} (d={i.nv}): transcript after the batched commit differs"
        GP.prove(i.cfg, t, i.dv, w, i.pts, i.evals)
        assert t.proof() == expected[k][1], f"instance {k} (d={i.nv}): proof from the batched witness differs"


MIX = [
    # nv, D, lir, engine, points
    (12, 2, 1, W.HASH_BLAKE3, 1),
    (14, 2, 1, W.HASH_BLAKE3, 2),
    (12, 2, 1, W.HASH_BLAKE3, 3),   # same shape as the first: one group
    (16, 2, 2, W.HASH_BLAKE3, 1),
    (13, 3, 1, W.HASH_SHA3, 1),     # SHA3 leaves (the unfused leaf path), Goldilocks3 OOD values
    (18, 2, 1, W.HASH_BLAKE3, 1),   # two-pass NTT shape
    (14, 1, 1, W.HASH_BLAKE3, 2),
    (14, 2, 1, W.HASH_BLAKE3, 1),   # same shape as the second
    (4, 2, 1, W.HASH_BLAKE3, 1),    # too small for the batched path: committed on its own
]


def test_commit_many_equals_plain_and_replays():
    rounds = []
    for rep in range(3):
        insts = [Inst(f"{rep}-{k}", nv, D, lir, eng, npts, seed=100 * rep + k)
                 for k, (nv, D, lir, eng, npts) in enumerate(MIX)]
        rounds.append((insts, [i.plain() for i in insts]))
    for rep, (insts, expected) in enumerate(rounds):
        print(f"[commit_many] replay {rep}", flush=True)
        check_batch(insts, expected)
        # plain work on the same stream between graph replays
        insts[0].plain()


def test_commit_many_c4_layout_and_executor_prove_only():
    """The 64-instance layout of the batched benchmark (num_variables ~ U{16..22},
    Goldilocks2, rate 1/2, Blake3): batched witnesses proved by prove-only executor
    jobs give the plain proofs, on two batches of the same layout (a replay)."""
    rng = np.random.default_rng(64)
    sizes = [int(rng.integers(16, 23)) for _ in range(64)]
    for rep in range(2):
        insts = [Inst(f"c4 {rep} {k}", nv, 2, 1, W.HASH_BLAKE3, 1 + k % 4, seed=1000 * rep + k)
                 for k, nv in enumerate(sizes)]
        expected = [i.plain() for i in insts]
        ts = [i.transcript() for i in insts]
        ws = GP.commit_many([i.cfg for i in insts], ts, [i.dv for i in insts])
        ex = GP.Executor(workers=8)
        try:
            jobs = [ex.submit_prove(i.cfg, t, i.dv, w, i.pts, i.evals) for i, t, w in zip(insts, ts, ws)]
            for jid in jobs:
                ex.wait(jid)
        finally:
            ex.close()
        for k, (t, e) in enumerate(zip(ts, expected)):
            assert t.proof() == e[1], f"batch {rep} instance {k} (d={insts[k].nv}) differs from its plain proof"


def test_commit_many_argument_errors():
    a = Inst("err", 10)
    t = a.transcript()
    with pytest.raises(W.WhirError, match="own transcript"):
        GP.commit_many([a.cfg, a.cfg], [t, t], [a.dv, a.dv])
    cfg2 = GP.Config(1 << 10, GP.ProtocolParameters(degree=2, security_level=32, pow_bits=0, batch_size=2))
    with pytest.raises(W.WhirError, match="batch_size 1"):
        GP.commit_many([cfg2], [a.transcript()], [a.dv])
    assert GP.commit_many([], [], []) == []
