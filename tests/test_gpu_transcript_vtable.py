"""The GPU prover driven through whir_transcript_from_vtable (the route a Rust
binding takes to run the prover over its own transcript, src/transcript/mod.rs:183-267):
callbacks wrapping the oracle's transcript receive every typed operation, in the
order the built-in transcript receives them, so the callback transcript's bytes
equal the built-in transcript's proof and the oracle prover's proof.  This covers
the delayed hint_field_vec / hint_hash calls of openings queued ahead of the next
sumcheck pass, several commitments in one proof and PoW nonces."""
import random

import numpy as np
import pytest

import whir_amd as W
from oracle import oracle as O
from oracle import whir_protocol as WP
from whir_amd import prover as GP

pytestmark = pytest.mark.gpu


class OracleOps:
    """The oracle transcript behind the callbacks; records the operation sequence."""

    def __init__(self, pid, sid):
        self.t = WP.Transcript(pid, sid)
        self.ops = []

    def prover_field(self, D, words):
        self.ops.append(("prover_field", D, len(words) // D))
        self.t.prover_field(D, words)

    def verifier_field(self, D):
        self.ops.append(("verifier_field", D))
        return self.t.verifier_field(D)

    def verifier_bytes(self, n):
        self.ops.append(("verifier_bytes", n))
        return self.t.verifier_bytes(n)

    def prover_hash(self, h):
        self.ops.append(("prover_hash",))
        self.t.prover_hash(h)

    def prover_u64(self, v):
        self.ops.append(("prover_u64",))
        self.t.prover_u64(v)

    def hint_hash(self, h):
        self.ops.append(("hint_hash",))
        self.t.hint_hash(h)

    def hint_field_vec(self, D, words):
        self.ops.append(("hint_field_vec", D, len(words) // D))
        self.t.hint_field_vec(D, words)


def make(nv, D, engine, B, NW, npts, sec, pow_bits, disable_pow, seed):
    g = np.random.default_rng(seed)
    vecs = [g.integers(0, O.P, size=1 << nv, dtype=np.uint64) for _ in range(B * NW)]
    rng = random.Random(seed)
    points = [[np.array([rng.randrange(O.P) for _ in range(D)], dtype=np.uint64) for _ in range(nv)]
              for _ in range(npts)]
    params = dict(security_level=sec, pow_bits=pow_bits, initial_folding_factor=4, folding_factor=4,
                  unique_decoding=False, starting_log_inv_rate=1, batch_size=B, hash_engine=engine)
    pid, sid = WP.protocol_ids(f"vtable nv={nv} D={D} B={B} W={NW} e={engine} pb={pow_bits} s={seed}")
    return vecs, points, params, pid, sid


def gpu_run(t, nv, D, B, NW, vecs, points, params, disable_pow):
    cfg = GP.Config(1 << nv, GP.ProtocolParameters(degree=D, **params))
    if disable_pow:
        cfg.disable_pow()
    dvs = [W.DeviceBuffer.from_array(v) for v in vecs]
    evals = np.array([[GP.mle_evaluate(D, dv, nv, np.concatenate(pt)) for dv in dvs] for pt in points],
                     dtype=np.uint64).reshape(len(points), len(vecs), D)
    ws = [GP.commit_batch(cfg, t, dvs[w * B:(w + 1) * B]) for w in range(NW)]
    flat = np.concatenate([np.concatenate(pt) for pt in points]) if points else np.zeros(0, np.uint64)
    ep = GP.prove_batch(cfg, t, dvs, ws, flat, evals)
    return evals, ep


def oracle_run(nv, D, B, NW, vecs, points, params, pid, sid, disable_pow):
    cfg = WP.WhirConfig.new(1 << nv, WP.ProtocolParameters(**params), D)
    if disable_pow:
        cfg.disable_pow()
    evals = [[WP.mle_evaluate_base(D, v, pt) for v in vecs] for pt in points]
    t = WP.Transcript(pid, sid)
    ws = [WP.commit_batch(t, cfg, vecs[w * B:(w + 1) * B]) for w in range(NW)]
    WP.prove_batch(t, cfg, vecs, ws, points, evals)
    return bytes(t.narg), bytes(t.hints)


CASES = [
    # nv, D, engine, batch_size, witnesses, points, security, pow bits, disable pow
    (12, 2, W.HASH_BLAKE3, 1, 1, 2, 32, 0, True),     # single vector, paired sumcheck rounds
    (16, 2, W.HASH_BLAKE3, 1, 1, 1, 64, 0, True),     # three WHIR rounds of queued openings
    (10, 3, W.HASH_SHA3, 1, 1, 1, 40, 12, False),     # PoW nonces (prover_u64) in every round
    (11, 2, W.HASH_BLAKE3, 2, 2, 2, 32, 0, True),     # two commitments of two vectors: cross terms
    (7, 3, W.HASH_BLAKE3, 2, 2, 1, 32, 0, True),      # no WHIR rounds: final opening of both witnesses
]


@pytest.mark.parametrize("nv,D,engine,B,NW,npts,sec,pb,dis", CASES)
def test_vtable_transcript_proof_equals_builtin(nv, D, engine, B, NW, npts, sec, pb, dis):
    vecs, points, params, pid, sid = make(nv, D, engine, B, NW, npts, sec, pb, dis, seed=nv + 7 * B)
    # built-in transcript
    tb = GP.Transcript(pid, sid)
    ev_b, ep_b = gpu_run(tb, nv, D, B, NW, vecs, points, params, dis)
    narg_b, hints_b = tb.proof()
    # the same proof through the vtable, callbacks over the oracle's transcript
    ops = OracleOps(pid, sid)
    tv = GP.CallbackTranscript(ops)
    ev_v, ep_v = gpu_run(tv, nv, D, B, NW, vecs, points, params, dis)
    assert tv.error is None
    assert np.array_equal(ev_v, ev_b) and np.array_equal(ep_v, ep_b)
    assert bytes(ops.t.narg) == narg_b, "narg through the vtable differs from the built-in transcript's"
    assert bytes(ops.t.hints) == hints_b, "hints through the vtable differ from the built-in transcript's"
    # and both equal the oracle prover's proof
    assert (narg_b, hints_b) == oracle_run(nv, D, B, NW, vecs, points, params, pid, sid, dis)
    # every opening hints its rows once per witness, then the Merkle siblings
    kinds = [o[0] for o in ops.ops]
    assert kinds[0] == "prover_hash" and kinds.count("prover_hash") >= NW
    assert kinds.count("hint_field_vec") >= NW
    if not dis:
        assert "prover_u64" in kinds


def test_vtable_callback_error_names_the_operation():
    """A callback that raises fails the call with the operation's name; the
    exception is kept on the transcript object."""
    vecs, points, params, pid, sid = make(10, 2, W.HASH_BLAKE3, 1, 1, 1, 32, 0, True, seed=3)

    class Broken(OracleOps):
        def verifier_field(self, D):
            raise ValueError("no challenge today")

    tv = GP.CallbackTranscript(Broken(pid, sid))
    with pytest.raises(W.WhirError, match="transcript callback failed: verifier_field"):
        gpu_run(tv, 10, 2, 1, 1, vecs, points, params, True)
    assert isinstance(tv.error, ValueError)
