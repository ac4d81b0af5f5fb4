"""End-to-end parity of the GPU WHIR prover with the oracle prover.

The same transcript spec runs on both sides, so a GPU proof must equal the
oracle's proof byte for byte (narg string and hints), and the oracle's
restatement of the reference verifier (src/protocols/whir/verifier.rs:45-264)
must accept it.
"""
import random

import numpy as np
import pytest

import whir_amd as W
from oracle import oracle as O
from oracle import whir_protocol as WP
from whir_amd import prover as GP

pytestmark = pytest.mark.gpu


def rand_vec(seed, n):
    g = np.random.default_rng(seed)
    return g.integers(0, O.P, size=n, dtype=np.uint64)


def make_case(nv, D, engine, k0, k, lir, num_points, ud=False, sec=32, pow_bits=0, disable_pow=True, seed=1):
    desc = f"gpu-prover nv={nv} D={D} e={engine} k0={k0} k={k} lir={lir} ud={ud} sec={sec} pb={pow_bits}"
    pid, sid = WP.protocol_ids(desc)
    vec = rand_vec(seed, 1 << nv)
    rng = random.Random(seed)
    points = [[np.array([rng.randrange(O.P) for _ in range(D)], dtype=np.uint64) for _ in range(nv)]
              for _ in range(num_points)]
    params = dict(security_level=sec, pow_bits=pow_bits, initial_folding_factor=k0, folding_factor=k,
                  unique_decoding=ud, starting_log_inv_rate=lir, hash_engine=engine)
    return pid, sid, vec, points, params


def oracle_proof(nv, D, pid, sid, vec, points, params, disable_pow):
    cfg = WP.WhirConfig.new(1 << nv, WP.ProtocolParameters(**params), D)
    if disable_pow:
        cfg.disable_pow()
    evals = [WP.mle_evaluate_base(D, vec, pt) for pt in points]
    t = WP.Transcript(pid, sid)
    w = WP.commit(t, cfg, vec)
    ep, _ = WP.prove(t, cfg, vec, w, points, evals)
    return cfg, evals, bytes(t.narg), bytes(t.hints), ep


def gpu_proof(nv, D, pid, sid, vec, points, params, disable_pow):
    cfg = GP.Config(1 << nv, GP.ProtocolParameters(degree=D, **params))
    if disable_pow:
        cfg.disable_pow()
    dv = W.DeviceBuffer.from_array(vec)
    evals = [GP.mle_evaluate(D, dv, nv, np.concatenate(pt)) for pt in points]
    t = GP.Transcript(pid, sid)
    w = GP.commit(cfg, t, dv)
    flat_pts = np.concatenate([np.concatenate(pt) for pt in points]) if points else np.zeros(0, np.uint64)
    ep = GP.prove(cfg, t, dv, w, flat_pts, np.concatenate(evals) if evals else np.zeros(0, np.uint64))
    narg, hints = t.proof()
    return evals, narg, hints, ep


CASES = [
    # nv, D, engine, k0, k, lir, points
    (10, 2, W.HASH_BLAKE3, 4, 4, 1, 1),
    (10, 1, W.HASH_BLAKE3, 4, 4, 1, 2),
    (10, 3, W.HASH_SHA3, 4, 4, 1, 1),
    (12, 2, W.HASH_SHA3, 4, 4, 2, 2),
    (12, 3, W.HASH_BLAKE3, 4, 4, 1, 0),
    (11, 2, W.HASH_BLAKE3, 3, 2, 2, 1),
    (9, 2, W.HASH_BLAKE3, 1, 1, 1, 1),
    (14, 2, W.HASH_BLAKE3, 4, 4, 2, 2),
    (13, 3, W.HASH_SHA3, 2, 3, 1, 3),
    # five rounds per sumcheck (two pairs and a single round after a pending pair) and a
    # base-field run of paired rounds
    (16, 2, W.HASH_BLAKE3, 5, 5, 1, 1),
    (17, 1, W.HASH_BLAKE3, 4, 4, 1, 1),
    # initial covector views: too few high-table bits for a second double fold (the view
    # is written out mid-sumcheck), and five terms (four points and an OOD sample: the
    # covector is materialized from the start)
    (7, 2, W.HASH_BLAKE3, 4, 2, 1, 1),
    (12, 2, W.HASH_BLAKE3, 4, 4, 1, 4),
]


def test_gpu_proof_with_openings_through_device_buffers(monkeypatch):
    """Openings too large for the pinned mailbox are gathered into a device buffer and
    read back on their own (forced here by shrinking the mailbox room they may use)."""
    monkeypatch.setenv("WHIR_OPEN_MAILBOX_WORDS", "16")
    nv, D, engine, k0, k, lir, npts = 12, 2, W.HASH_BLAKE3, 4, 4, 1, 2
    pid, sid, vec, points, params = make_case(nv, D, engine, k0, k, lir, npts)
    _, _, narg, hints, _ = oracle_proof(nv, D, pid, sid, vec, points, params, True)
    _, gnarg, ghints, _ = gpu_proof(nv, D, pid, sid, vec, points, params, True)
    assert gnarg == narg and ghints == hints


@pytest.mark.parametrize("nv,D,engine,k0,k,lir,npts", CASES)
def test_gpu_proof_equals_oracle_proof(nv, D, engine, k0, k, lir, npts):
    pid, sid, vec, points, params = make_case(nv, D, engine, k0, k, lir, npts)
    cfg, evals, narg, hints, ep = oracle_proof(nv, D, pid, sid, vec, points, params, True)
    gevals, gnarg, ghints, gep = gpu_proof(nv, D, pid, sid, vec, points, params, True)
    for a, b in zip(gevals, evals):
        assert np.array_equal(a, b)
    assert gnarg == narg
    assert ghints == hints
    assert np.array_equal(gep, np.concatenate(ep))
    # the oracle verifier accepts the GPU proof
    WP.verify(WP.TranscriptReader(pid, sid, gnarg, ghints), cfg, evals, points)


@pytest.mark.parametrize("ud,sec,pb,engine", [(False, 40, 12, W.HASH_BLAKE3), (True, 30, 8, W.HASH_SHA3)])
def test_gpu_proof_with_pow(ud, sec, pb, engine):
    nv, D = 10, 3
    pid, sid, vec, points, params = make_case(nv, D, engine, 2, 2, 1, 1, ud=ud, sec=sec, pow_bits=pb)
    cfg, evals, narg, hints, ep = oracle_proof(nv, D, pid, sid, vec, points, params, False)
    gevals, gnarg, ghints, gep = gpu_proof(nv, D, pid, sid, vec, points, params, False)
    assert gnarg == narg and ghints == hints
    WP.verify(WP.TranscriptReader(pid, sid, gnarg, ghints), cfg, evals, points)


@pytest.mark.parametrize("kind", ["zeros", "max", "iota", "ones"])
@pytest.mark.parametrize("nv,D,k0,k,lir", [(12, 2, 4, 4, 1), (5, 3, 1, 1, 1), (4, 2, 4, 4, 2), (9, 1, 3, 2, 1)])
def test_gpu_proof_special_vectors(kind, nv, D, k0, k, lir):
    """Edge inputs: the all-zero vector (zero OOD values, zero sums), every
    coefficient p - 1 (carries at every reduction), v[i] = i (the reference CLI's
    input, src/bin/main.rs:150) and all ones (the reference tests' vector,
    whir/mod.rs:209); tiny sizes down to one folding round and no WHIR rounds."""
    n = 1 << nv
    vec = {"zeros": np.zeros(n, np.uint64), "max": np.full(n, O.P - 1, np.uint64),
           "iota": np.arange(n, dtype=np.uint64), "ones": np.ones(n, np.uint64)}[kind]
    pid, sid, _, points, params = make_case(nv, D, W.HASH_BLAKE3, k0, k, lir, 2, ud=(D == 1))
    cfg, evals, narg, hints, ep = oracle_proof(nv, D, pid, sid, vec, points, params, True)
    gevals, gnarg, ghints, gep = gpu_proof(nv, D, pid, sid, vec, points, params, True)
    assert (gnarg, ghints) == (narg, hints)
    assert np.array_equal(gep, np.concatenate(ep))
    if kind == "zeros":
        # The verifier divides by the final polynomial's value (verifier.rs:241-243),
        # zero here: the reference panics, the oracle rejects.
        with pytest.raises(WP.VerificationError, match="division by zero"):
            WP.verify(WP.TranscriptReader(pid, sid, gnarg, ghints), cfg, evals, points)
    else:
        WP.verify(WP.TranscriptReader(pid, sid, gnarg, ghints), cfg, evals, points)
