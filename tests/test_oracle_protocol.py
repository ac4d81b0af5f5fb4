"""Pins the oracle's protocol layer (oracle/whir_protocol.py) on the CPU.

- parameter selection against the round shapes of the three headline configs
  (an independent restatement of src/protocols/whir/config.rs:16-152 and
  src/protocols/irs_commit.rs:125-288 tabulated with the survey of the reference);
- prove -> verify round trips over fields, hashes, folding factors, rates and
  PoW settings (the reference's make_whir_things matrix, whir/mod.rs:176-894);
- negative tests: tampered proofs are rejected (whir/mod.rs:515-612 analogue).
"""
import math
import random

import numpy as np
import pytest

from oracle import oracle as O
from oracle import whir_protocol as WP


def shapes(cfg):
    out = [(cfg.initial.message_length(), cfg.initial.codeword_length, cfg.initial.out_domain_samples,
            cfg.initial.in_domain_samples)]
    for r in cfg.rounds:
        out.append((r.irs.message_length(), r.irs.codeword_length, r.irs.out_domain_samples,
                    r.irs.in_domain_samples))
    return out


def bits(thr):
    return round(WP.pow_difficulty(thr), 2)


def test_config_c1_shapes():
    p = WP.ProtocolParameters(security_level=100, pow_bits=20, starting_log_inv_rate=1)
    cfg = WP.WhirConfig.new(1 << 20, p, 2)
    assert shapes(cfg) == [(1 << 16, 1 << 17, 1, 187), (1 << 12, 1 << 16, 1, 42), (1 << 8, 1 << 15, 1, 24),
                           (1 << 4, 1 << 14, 1, 17), (1, 1 << 13, 1, 13)]
    assert bits(cfg.initial_sumcheck.round_pow) == 30.75
    assert [bits(r.pow) for r in cfg.rounds] == [19.66, 18.96, 17.69, 16.2]
    assert [bits(r.sumcheck.round_pow) for r in cfg.rounds] == [33.25, 35.75, 38.25, 40.75]
    assert bits(cfg.final_pow) == 16.42
    assert cfg.final_sumcheck.initial_size == 1 and cfg.final_sumcheck.num_rounds == 0


def test_config_c2_shapes():
    p = WP.ProtocolParameters(security_level=100, pow_bits=20, starting_log_inv_rate=2)
    cfg = WP.WhirConfig.new(1 << 24, p, 2)
    assert shapes(cfg) == [(1 << 20, 1 << 22, 1, 87), (1 << 16, 1 << 21, 1, 33), (1 << 12, 1 << 20, 1, 21),
                           (1 << 8, 1 << 19, 1, 15), (1 << 4, 1 << 18, 1, 12), (1, 1 << 17, 1, 10)]
    assert bits(cfg.initial_sumcheck.round_pow) == 42.25
    assert [bits(r.pow) for r in cfg.rounds] == [19.12, 19.82, 17.48, 18.56, 16.84]
    assert [bits(r.sumcheck.round_pow) for r in cfg.rounds] == [44.75, 47.25, 49.75, 52.25, 54.75]
    assert bits(cfg.final_pow) == 15.7


def test_config_c3_shapes():
    p = WP.ProtocolParameters(security_level=100, pow_bits=20, starting_log_inv_rate=1, hash_engine=O.ENGINE_SHA3)
    cfg = WP.WhirConfig.new(1 << 26, p, 3)
    assert shapes(cfg) == [(1 << 22, 1 << 23, 1, 187), (1 << 18, 1 << 22, 1, 42), (1 << 14, 1 << 21, 1, 24),
                           (1 << 10, 1 << 20, 1, 17), (1 << 6, 1 << 19, 1, 13), (1 << 2, 1 << 18, 1, 11)]
    assert cfg.initial_sumcheck.round_pow == 2**64 - 1
    assert [bits(r.pow) for r in cfg.rounds] == [19.66, 18.96, 17.69, 16.2, 16.42]
    assert bits(cfg.final_pow) == 12.77
    assert cfg.final_sumcheck.initial_size == 4


def test_config_unique_decoding_has_no_ood():
    p = WP.ProtocolParameters(security_level=32, pow_bits=0, unique_decoding=True)
    cfg = WP.WhirConfig.new(1 << 12, p, 3)
    assert cfg.initial.out_domain_samples == 0
    assert all(r.irs.out_domain_samples == 0 for r in cfg.rounds)


def test_transcript_is_deterministic_and_order_sensitive():
    a = WP.Transcript(b"p" * 64, b"s" * 32)
    b = WP.Transcript(b"p" * 64, b"s" * 32)
    a.prover_hash(b"\1" * 32)
    b.prover_hash(b"\1" * 32)
    assert a.verifier_bytes(40) == b.verifier_bytes(40)
    c = WP.Transcript(b"p" * 64, b"s" * 32)
    c.prover_hash(b"\2" * 32)
    d = WP.Transcript(b"p" * 64, b"s" * 32)
    d.prover_hash(b"\1" * 32)
    assert c.verifier_bytes(8) != d.verifier_bytes(8)
    # field challenges are canonical
    e = WP.Transcript(b"q" * 64, b"s" * 32)
    assert all(int(x) < O.P for _ in range(50) for x in e.verifier_field(3))


def rand_vec(seed, n):
    g = np.random.default_rng(seed)
    v = g.integers(0, O.P, size=n, dtype=np.uint64)
    return v


def run_roundtrip(nv, D, engine, k0, k, lir, num_points, ud=False, sec=32, pow_bits=0, disable_pow=True,
                  seed=1):
    p = WP.ProtocolParameters(security_level=sec, pow_bits=pow_bits, initial_folding_factor=k0,
                              folding_factor=k, unique_decoding=ud, starting_log_inv_rate=lir,
                              hash_engine=engine)
    cfg = WP.WhirConfig.new(1 << nv, p, D)
    if disable_pow:
        cfg.disable_pow()
    pid, sid = WP.protocol_ids(f"test nv={nv} D={D} e={engine} k0={k0} k={k} lir={lir}")
    vec = rand_vec(seed, 1 << nv)
    rng = random.Random(seed)
    points = [[np.array([rng.randrange(O.P) for _ in range(D)], dtype=np.uint64) for _ in range(nv)]
              for _ in range(num_points)]
    evals = [WP.mle_evaluate_base(D, vec, pt) for pt in points]
    t = WP.Transcript(pid, sid)
    w = WP.commit(t, cfg, vec)
    WP.prove(t, cfg, vec, w, points, evals)
    r = WP.TranscriptReader(pid, sid, bytes(t.narg), bytes(t.hints))
    WP.verify(r, cfg, evals, points)
    return cfg, pid, sid, bytes(t.narg), bytes(t.hints), evals, points


@pytest.mark.parametrize("D", [1, 2, 3])
@pytest.mark.parametrize("num_points", [0, 1, 2])
def test_prove_verify_roundtrip_fields(D, num_points):
    run_roundtrip(10, D, O.ENGINE_BLAKE3, 4, 4, 1, num_points)


@pytest.mark.parametrize("k0,k", [(1, 1), (2, 3), (3, 2), (4, 4)])
@pytest.mark.parametrize("engine", [O.ENGINE_SHA3, O.ENGINE_BLAKE3])
def test_prove_verify_roundtrip_folding(k0, k, engine):
    run_roundtrip(9, 2, engine, k0, k, 2, 1)


def test_prove_verify_unique_decoding_and_pow():
    run_roundtrip(8, 3, O.ENGINE_SHA3, 2, 2, 1, 1, ud=True, sec=24, pow_bits=6, disable_pow=False)
    run_roundtrip(8, 2, O.ENGINE_BLAKE3, 4, 4, 1, 2, ud=False, sec=40, pow_bits=10, disable_pow=False)


def test_tampered_proofs_are_rejected():
    cfg, pid, sid, narg, hints, evals, points = run_roundtrip(8, 2, O.ENGINE_BLAKE3, 2, 2, 1, 1)
    rng = random.Random(9)
    for _ in range(6):
        pos = rng.randrange(len(narg))
        bad = bytearray(narg)
        bad[pos] ^= 1 << rng.randrange(8)
        with pytest.raises((WP.VerificationError, AssertionError, ValueError)):
            WP.verify(WP.TranscriptReader(pid, sid, bytes(bad), hints), cfg, evals, points)
    for _ in range(6):
        pos = rng.randrange(len(hints))
        bad = bytearray(hints)
        bad[pos] ^= 1 << rng.randrange(8)
        with pytest.raises((WP.VerificationError, AssertionError, ValueError)):
            WP.verify(WP.TranscriptReader(pid, sid, narg, bytes(bad)), cfg, evals, points)
    # a wrong claimed evaluation
    wrong = [WP.e_add(2, evals[0], WP.e_one(2))]
    with pytest.raises(WP.VerificationError):
        WP.verify(WP.TranscriptReader(pid, sid, narg, hints), cfg, wrong, points)


# ---------------------------------------------------------------- several vectors / commitments
def run_batch(nv, D, B, W, num_points, engine=O.ENGINE_BLAKE3, k0=2, k=2, lir=1, ud=False, seed=7):
    """batch_size = B vectors per commitment, W commitments, all proven together
    (whir/prover.rs:50-316 with vectors.len() = W * B; the reference's tests of
    several witnesses, whir/mod.rs:346-894)."""
    p = WP.ProtocolParameters(security_level=32, pow_bits=0, initial_folding_factor=k0, folding_factor=k,
                              unique_decoding=ud, starting_log_inv_rate=lir, batch_size=B, hash_engine=engine)
    cfg = WP.WhirConfig.new(1 << nv, p, D)
    cfg.disable_pow()
    pid, sid = WP.protocol_ids(f"batch nv={nv} D={D} B={B} W={W} e={engine}")
    vecs = [rand_vec(seed + j, 1 << nv) for j in range(B * W)]
    rng = random.Random(seed)
    points = [[np.array([rng.randrange(O.P) for _ in range(D)], dtype=np.uint64) for _ in range(nv)]
              for _ in range(num_points)]
    evals = [[WP.mle_evaluate_base(D, v, pt) for v in vecs] for pt in points]
    t = WP.Transcript(pid, sid)
    ws = [WP.commit_batch(t, cfg, vecs[w * B:(w + 1) * B]) for w in range(W)]
    ep, _ = WP.prove_batch(t, cfg, vecs, ws, points, evals, check_invariants=True)
    r = WP.TranscriptReader(pid, sid, bytes(t.narg), bytes(t.hints))
    vp = WP.verify_batch(r, cfg, W, evals, points)
    assert len(vp) == len(ep) and all(np.array_equal(a, b) for a, b in zip(vp, ep))
    return cfg, pid, sid, bytes(t.narg), bytes(t.hints), vecs, ws, evals, points


@pytest.mark.parametrize("B,W", [(1, 2), (2, 1), (2, 2), (3, 1), (1, 3)])
@pytest.mark.parametrize("D", [2, 3])
@pytest.mark.parametrize("num_points", [0, 1, 2])
def test_batch_prove_verify_roundtrip(B, W, D, num_points):
    run_batch(10, D, B, W, num_points)


@pytest.mark.parametrize("B,W", [(2, 2), (1, 2)])
def test_batch_unique_decoding_and_sha3(B, W):
    run_batch(9, 2, B, W, 1, engine=O.ENGINE_SHA3, ud=True, k0=3, k=2)


def test_batch_single_vector_is_the_plain_proof():
    """One vector in one commitment: the batched prover is the plain prover."""
    cfg, pid, sid, narg, hints, vecs, ws, evals, points = run_batch(10, 2, 1, 1, 2)
    t = WP.Transcript(pid, sid)
    w = WP.commit(t, cfg, vecs[0])
    WP.prove(t, cfg, vecs[0], w, points, [row[0] for row in evals])
    assert (bytes(t.narg), bytes(t.hints)) == (narg, hints)


def test_batch_commit_columns_are_per_vector_blocks():
    """messages = vectors.flat_map(chunks) (irs_commit.rs:317-320): vector b owns
    columns [b*depth, (b+1)*depth) of the committed matrix, each block equal to
    that vector committed alone."""
    B, nv = 3, 10
    p = WP.ProtocolParameters(security_level=32, pow_bits=0, initial_folding_factor=2, folding_factor=2,
                              starting_log_inv_rate=1, batch_size=B)
    cfg = WP.WhirConfig.new(1 << nv, p, 2)
    p1 = WP.ProtocolParameters(security_level=32, pow_bits=0, initial_folding_factor=2, folding_factor=2,
                               starting_log_inv_rate=1)
    cfg1 = WP.WhirConfig.new(1 << nv, p1, 2)
    vecs = [rand_vec(40 + b, 1 << nv) for b in range(B)]
    wb = WP.commit_batch(WP.Transcript(b"a" * 64, b"b" * 32), cfg, vecs)
    depth = cfg.initial.interleaving_depth
    assert wb.matrix.shape == (cfg.initial.codeword_length, B * depth)
    for b in range(B):
        w1 = WP.commit(WP.Transcript(b"a" * 64, b"b" * 32), cfg1, vecs[b])
        assert np.array_equal(wb.matrix[:, b * depth:(b + 1) * depth], w1.matrix)
    # OOD values are point-major: value (i, b) = vector b at OOD point i
    for i, x in enumerate(wb.ood_points):
        for b in range(B):
            assert np.array_equal(wb.ood_values[i * B + b], O.univariate_evaluate(1, 2, vecs[b], x))


def test_batch_tampered_evaluation_is_rejected():
    cfg, pid, sid, narg, hints, vecs, ws, evals, points = run_batch(10, 2, 2, 2, 2)
    bad = [list(row) for row in evals]
    bad[1][3] = WP.e_add(2, bad[1][3], WP.e_one(2))
    with pytest.raises(WP.VerificationError):
        WP.verify_batch(WP.TranscriptReader(pid, sid, narg, hints), cfg, 2, bad, points)
    # swapping two commitments' vectors in the claim is caught too
    swapped = [[row[2], row[3], row[0], row[1]] for row in evals]
    with pytest.raises(WP.VerificationError):
        WP.verify_batch(WP.TranscriptReader(pid, sid, narg, hints), cfg, 2, swapped, points)


# ---------------------------------------------------------------- Covector linear forms
def covector_iota(n, D):
    """The reference tests' covector: element i is the field element i
    (whir/mod.rs:226-229, (0..1 << num_variables).map(EF::from))."""
    v = np.zeros(n * D, dtype=np.uint64)
    v[0::D] = np.arange(n, dtype=np.uint64)
    return WP.Covector(v)


def run_forms(nv, D, forms_spec, B=1, NW=1, seed=9):
    """forms_spec: string of 'm' (MLE point) and 'c' (covector) in transcript order."""
    p = WP.ProtocolParameters(security_level=32, pow_bits=0, initial_folding_factor=2, folding_factor=2,
                              starting_log_inv_rate=1, batch_size=B)
    cfg = WP.WhirConfig.new(1 << nv, p, D)
    cfg.disable_pow()
    pid, sid = WP.protocol_ids(f"forms {nv} {D} {forms_spec} {B} {NW}")
    vecs = [rand_vec(seed + j, 1 << nv) for j in range(B * NW)]
    rng = random.Random(seed)
    forms = []
    for k, f in enumerate(forms_spec):
        if f == "m":
            forms.append([np.array([rng.randrange(O.P) for _ in range(D)], dtype=np.uint64) for _ in range(nv)])
        elif f == "u":
            forms.append(WP.Univariate(np.array([rng.randrange(O.P) for _ in range(D)], dtype=np.uint64)))
        elif k == len(forms_spec) - 1:
            forms.append(covector_iota(1 << nv, D))
        else:
            g = np.random.default_rng(seed + 100 + k)
            forms.append(WP.Covector(g.integers(0, O.P, size=(1 << nv) * D, dtype=np.uint64)))
    evals = [[WP.form_evaluate(D, f, v) for v in vecs] for f in forms]
    t = WP.Transcript(pid, sid)
    ws = [WP.commit_batch(t, cfg, vecs[w * B:(w + 1) * B]) for w in range(NW)]
    WP.prove_batch(t, cfg, vecs, ws, forms, evals, check_invariants=True)
    WP.verify_batch(WP.TranscriptReader(pid, sid, bytes(t.narg), bytes(t.hints)), cfg, NW, evals, forms)
    return cfg, pid, sid, bytes(t.narg), bytes(t.hints), vecs, forms, evals


@pytest.mark.parametrize("spec", ["mmc", "c", "cm", "mcm", "cc", "u", "mu", "umc", "cuu"])
@pytest.mark.parametrize("D", [2, 3])
def test_covector_forms_roundtrip(spec, D):
    run_forms(10, D, spec)


def test_covector_forms_with_several_commitments():
    run_forms(9, 2, "mc", B=2, NW=2)


def test_covector_evaluate_is_the_mle_sum():
    """Covector evaluation = sum_i c_i v_i; with c = eq(z, .) it is the MLE at z."""
    nv, D = 8, 2
    rng = random.Random(3)
    v = rand_vec(4, 1 << nv)
    z = [np.array([rng.randrange(O.P) for _ in range(D)], dtype=np.uint64) for _ in range(nv)]
    eq = O.eval_eq(D, np.zeros((1 << nv) * D, dtype=np.uint64), np.concatenate(z), WP.e_one(D))
    assert np.array_equal(WP.covector_evaluate(D, WP.Covector(eq), v), WP.mle_evaluate_base(D, v, z))


def test_tampered_covector_claim_is_rejected():
    cfg, pid, sid, narg, hints, vecs, forms, evals = run_forms(10, 2, "mmc")
    bad = [list(r) for r in evals]
    bad[2][0] = WP.e_add(2, bad[2][0], WP.e_one(2))
    with pytest.raises(WP.VerificationError):
        WP.verify_batch(WP.TranscriptReader(pid, sid, narg, hints), cfg, 1, bad, forms)
    # a different covector with the claimed value unchanged is caught by the final claim
    other = list(forms)
    v = np.asarray(forms[2].vector, dtype=np.uint64).copy()
    v[0] = (int(v[0]) + 1) % O.P
    other[2] = WP.Covector(v)
    with pytest.raises(WP.VerificationError):
        WP.verify_batch(WP.TranscriptReader(pid, sid, narg, hints), cfg, 1, evals, other)


def test_univariate_form_mle_is_the_mle_of_its_powers():
    """UnivariateEvaluation::mle_evaluate (univariate_evaluation.rs:46-56) equals the
    multilinear extension of (1, x, x^2, ...) at the point, variable 0 the MSB."""
    nv, D = 7, 2
    rng = random.Random(12)
    x = np.array([rng.randrange(O.P) for _ in range(D)], dtype=np.uint64)
    r = [np.array([rng.randrange(O.P) for _ in range(D)], dtype=np.uint64) for _ in range(nv)]
    powers = np.concatenate(WP.geometric_sequence(D, x, 1 << nv))
    assert np.array_equal(WP.univariate_mle(D, x, r), O.multilinear_extend(D, powers, np.concatenate(r)))
    # and its evaluation on a vector is Horner's rule
    v = rand_vec(13, 1 << nv)
    assert np.array_equal(WP.form_evaluate(D, WP.Univariate(x), v), O.dot(D, powers, O.lift(D, v)))
