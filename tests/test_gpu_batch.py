"""Several vectors per commitment (batch_size > 1) and several commitments proven
together (whir/prover.rs:50-316 with vectors.len() = witnesses * batch_size): the
GPU proof equals the oracle's byte for byte and the oracle verifier accepts it."""
import random

import numpy as np
import pytest

import whir_amd as W
from oracle import oracle as O
from oracle import whir_protocol as WP
from whir_amd import prover as GP

pytestmark = pytest.mark.gpu


def case(nv, D, engine, k0, k, lir, B, NW, npts, ud=False, seed=3):
    V = B * NW
    g = np.random.default_rng(seed)
    vecs = [g.integers(0, O.P, size=1 << nv, dtype=np.uint64) for _ in range(V)]
    rng = random.Random(seed)
    points = [[np.array([rng.randrange(O.P) for _ in range(D)], dtype=np.uint64) for _ in range(nv)]
              for _ in range(npts)]
    params = dict(security_level=32, pow_bits=0, initial_folding_factor=k0, folding_factor=k, unique_decoding=ud,
                  starting_log_inv_rate=lir, batch_size=B, hash_engine=engine)
    pid, sid = WP.protocol_ids(f"gpu batch nv={nv} D={D} B={B} W={NW} e={engine} k0={k0} k={k} ud={ud}")
    return vecs, points, params, pid, sid


def oracle_batch(nv, D, B, NW, vecs, points, params, pid, sid):
    cfg = WP.WhirConfig.new(1 << nv, WP.ProtocolParameters(**params), D)
    cfg.disable_pow()
    evals = [[WP.mle_evaluate_base(D, v, pt) for v in vecs] for pt in points]
    t = WP.Transcript(pid, sid)
    ws = [WP.commit_batch(t, cfg, vecs[w * B:(w + 1) * B]) for w in range(NW)]
    ep, _ = WP.prove_batch(t, cfg, vecs, ws, points, evals)
    return cfg, evals, bytes(t.narg), bytes(t.hints), ep


def gpu_batch(nv, D, B, NW, vecs, points, params, pid, sid):
    cfg = GP.Config(1 << nv, GP.ProtocolParameters(degree=D, **params)).disable_pow()
    dvs = [W.DeviceBuffer.from_array(v) for v in vecs]
    evals = np.array([[GP.mle_evaluate(D, dv, nv, np.concatenate(pt)) for dv in dvs] for pt in points],
                     dtype=np.uint64).reshape(len(points), len(vecs), D)
    t = GP.Transcript(pid, sid)
    ws = [GP.commit_batch(cfg, t, dvs[w * B:(w + 1) * B]) for w in range(NW)]
    flat = np.concatenate([np.concatenate(pt) for pt in points]) if points else np.zeros(0, np.uint64)
    ep = GP.prove_batch(cfg, t, dvs, ws, flat, evals)
    narg, hints = t.proof()
    return evals, narg, hints, ep


CASES = [
    # nv, D, engine, k0, k, lir, batch_size, witnesses, points, unique decoding
    (10, 2, W.HASH_BLAKE3, 4, 4, 1, 2, 1, 2, False),
    (10, 2, W.HASH_BLAKE3, 4, 4, 1, 1, 2, 1, False),
    (11, 3, W.HASH_SHA3, 2, 3, 1, 2, 2, 2, False),
    (10, 1, W.HASH_BLAKE3, 4, 4, 1, 3, 1, 1, True),
    (12, 2, W.HASH_BLAKE3, 4, 4, 2, 4, 2, 0, False),
    (9, 2, W.HASH_SHA3, 3, 2, 1, 2, 2, 1, True),
    (10, 2, W.HASH_BLAKE3, 4, 4, 1, 8, 1, 1, False),   # 128-byte-word rows: the fused leaf + tree path
    (10, 2, W.HASH_BLAKE3, 4, 4, 1, 9, 1, 1, False),   # 1152-byte rows: multi-chunk BLAKE3 leaves
    (14, 2, W.HASH_BLAKE3, 4, 4, 2, 2, 3, 2, False),
    (7, 3, W.HASH_BLAKE3, 4, 4, 1, 2, 2, 1, False),    # no WHIR rounds: the final open takes the initial witnesses
]


@pytest.mark.parametrize("nv,D,engine,k0,k,lir,B,NW,npts,ud", CASES)
def test_gpu_batch_proof_equals_oracle(nv, D, engine, k0, k, lir, B, NW, npts, ud):
    vecs, points, params, pid, sid = case(nv, D, engine, k0, k, lir, B, NW, npts, ud)
    cfg, oevals, narg, hints, oep = oracle_batch(nv, D, B, NW, vecs, points, params, pid, sid)
    gevals, gnarg, ghints, gep = gpu_batch(nv, D, B, NW, vecs, points, params, pid, sid)
    assert np.array_equal(gevals.reshape(-1), np.concatenate([np.concatenate(r) for r in oevals])
                          if oevals else np.zeros(0, np.uint64))
    assert gnarg == narg, "narg strings differ"
    assert ghints == hints, "hints differ"
    assert np.array_equal(gep, np.concatenate(oep))
    WP.verify_batch(WP.TranscriptReader(pid, sid, gnarg, ghints), cfg, NW, oevals, points)


def test_gpu_batch_d20_two_commitments_of_two():
    nv, D, B, NW = 20, 2, 2, 2
    vecs, points, params, pid, sid = case(nv, D, W.HASH_BLAKE3, 4, 4, 1, B, NW, 2)
    params["security_level"] = 100
    params["pow_bits"] = 20
    cfg, oevals, narg, hints, oep = oracle_batch(nv, D, B, NW, vecs, points, params, pid, sid)
    gevals, gnarg, ghints, gep = gpu_batch(nv, D, B, NW, vecs, points, params, pid, sid)
    assert (gnarg, ghints) == (narg, hints)
    WP.verify_batch(WP.TranscriptReader(pid, sid, gnarg, ghints), cfg, NW, oevals, points)


def test_gpu_batch_argument_errors():
    nv, D = 10, 2
    cfg = GP.Config(1 << nv, GP.ProtocolParameters(degree=D, security_level=32, pow_bits=0, batch_size=2))
    dvs = [W.DeviceBuffer.from_array(np.arange(1 << nv, dtype=np.uint64)) for _ in range(3)]
    t = GP.Transcript(b"\0" * 64, b"\0" * 32)
    with pytest.raises(W.WhirError, match="different number of vectors"):
        GP.commit_batch(cfg, t, dvs[:3])
    t = GP.Transcript(b"\0" * 64, b"\0" * 32)
    w = GP.commit_batch(cfg, t, dvs[:2])
    with pytest.raises(W.WhirError, match="witnesses times batch_size"):
        GP.prove_batch(cfg, t, dvs[:3], [w], np.zeros(0, np.uint64), np.zeros(0, np.uint64))
