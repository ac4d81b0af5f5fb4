"""NTT schedules: every candidate schedule of the autotuner (four-step split,
per-pass LDS padding / radix 16 / tile size) computes the same Reed-Solomon
encoding as the oracle (cooley_tukey.rs:411-477 restated), and tuning leaves
encodings and whole proofs unchanged."""
import hashlib

import numpy as np
import pytest

import whir_amd as W
from oracle import oracle as O
from oracle import whir_protocol as WP
from whir_amd import prover as GP

pytestmark = pytest.mark.gpu


@pytest.fixture
def clean_schedules():
    W.ntt_autotune(False)
    W.ntt_clear_schedules()
    yield
    W.ntt_autotune(False)
    W.ntt_clear_schedules()


def tiles():
    # 8192-word tiles are offered where the device's LDS holds them (gfx950 does)
    return [4096, 8192]


def log2(n):
    return n.bit_length() - 1


# D, M, L, N: two-pass shapes with an even and an odd coset exponent, one single-pass shape
SHAPES = [(1, 4, 1 << 12, 1 << 14), (2, 3, 1 << 13, 1 << 14), (3, 2, 1 << 13, 1 << 15), (2, 5, 1 << 10, 1 << 12)]


@pytest.mark.parametrize("D,M,L,N", SHAPES)
def test_every_candidate_schedule_matches_oracle(D, M, L, N, clean_schedules):
    rng = np.random.default_rng(D * 100 + M)
    msgs = rng.integers(0, O.P, size=(M, L * D), dtype=np.uint64)
    exp = O.interleaved_encode(D, msgs, N)
    logC, logN = log2(L), log2(N)
    scheds = []
    if logC <= 11:
        for v in range(4):
            for tw in tiles():
                scheds.append((0, v, tw, 0, 4096))
    else:
        # the tuner's splits: the balanced one(s) and one level either side
        splits = sorted({(logC + 1) // 2, logC // 2, (logC + 1) // 2 + 1, logC // 2 - 1})
        for c1 in splits:
            for v in range(4):
                for tw in tiles():
                    scheds.append((c1, v, tw, 0, 4096))
                    scheds.append((c1, 0, 4096, v, tw))
    for s in scheds:
        W.ntt_set_schedule(D, logN, logC, M, *s)
        got = W.interleaved_rs_encode(list(msgs), None, N, degree=D)
        assert np.array_equal(got, exp), f"schedule {s} differs from the oracle"
    assert len(W.ntt_schedules()) == 1


def test_autotuned_encodes_match_oracle_and_are_recorded(clean_schedules):
    W.ntt_autotune(True)
    rng = np.random.default_rng(7)
    for D, M, L, N in SHAPES:
        msgs = rng.integers(0, O.P, size=(M, L * D), dtype=np.uint64)
        exp = O.interleaved_encode(D, msgs, N)
        for _ in range(2):  # the first call tunes, the second uses the table
            assert np.array_equal(W.interleaved_rs_encode(list(msgs), None, N, degree=D), exp)
    table = W.ntt_schedules()
    keys = {(e["degree"], e["log_n"], e["log_c"], e["messages"]) for e in table}
    assert keys == {(D, log2(N), log2(L), M) for D, M, L, N in SHAPES}
    for e in table:
        assert e["candidates"] >= 7 and e["tuned_ms"] > 0
        assert e["pass1"]["tile_words"] in tiles() and e["pass2"]["tile_words"] in tiles()


def test_pinned_schedule_is_validated(clean_schedules):
    with pytest.raises(W.WhirError, match="variant"):
        W.ntt_set_schedule(1, 14, 12, 4, 6, 4, 4096, 0, 4096)
    with pytest.raises(W.WhirError, match="split"):
        W.ntt_set_schedule(1, 20, 18, 4, 5, 0, 4096, 0, 4096)
    with pytest.raises(W.WhirError, match="tile"):
        W.ntt_set_schedule(1, 14, 12, 4, 6, 0, 1000, 0, 4096)
    with pytest.raises(W.WhirError, match="single-pass"):
        W.ntt_set_schedule(1, 12, 10, 4, 5, 0, 4096, 0, 4096)


def test_autotuned_proof_is_the_oracle_proof(clean_schedules):
    W.ntt_autotune(True)
    nv, D = 16, 2
    rng = np.random.default_rng(11)
    vec = rng.integers(0, O.P, size=1 << nv, dtype=np.uint64)
    pt = [rng.integers(0, O.P, size=D, dtype=np.uint64) for _ in range(nv)]
    params = dict(security_level=64, pow_bits=0, initial_folding_factor=4, folding_factor=4, starting_log_inv_rate=2)
    pid, sid = hashlib.sha3_512(b"autotune").digest(), hashlib.sha3_256(b"autotune").digest()
    ocfg = WP.WhirConfig.new(1 << nv, WP.ProtocolParameters(**params), D)
    ocfg.disable_pow()
    ev = WP.mle_evaluate_base(D, vec, pt)
    t = WP.Transcript(pid, sid)
    WP.prove(t, ocfg, vec, WP.commit(t, ocfg, vec), [pt], [ev])
    cfg = GP.Config(1 << nv, GP.ProtocolParameters(degree=D, **params)).disable_pow()
    dv = W.DeviceBuffer.from_array(vec)
    for _ in range(2):  # tuning run, then tuned run
        gt = GP.Transcript(pid, sid)
        GP.prove(cfg, gt, dv, GP.commit(cfg, gt, dv), np.concatenate(pt), ev)
        assert gt.proof() == (bytes(t.narg), bytes(t.hints))
    assert len(W.ntt_schedules()) >= 2  # the initial and the round codewords
