"""GPU parity of every hot-path kernel against the CPU oracle, through the C ABI.

Bit-exact comparisons on seeded inputs at sizes the oracle finishes in seconds,
plus size-independent properties at full sizes.
"""
import hashlib
import random

import numpy as np
import pytest

import whir_amd as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu
P = W.P


def rnd(rng, n):
    return np.array([rng.randrange(P) for _ in range(n)], dtype=np.uint64)


def rnd_np(seed, n):
    # uniform canonical values quickly: rejection of the (rare) >= p values
    g = np.random.default_rng(seed)
    v = g.integers(0, 2**64, size=n, dtype=np.uint64)
    bad = v >= np.uint64(P)
    while bad.any():
        v[bad] = g.integers(0, 2**64, size=int(bad.sum()), dtype=np.uint64)
        bad = v >= np.uint64(P)
    return v


# ------------------------------------------------------------------ Reed-Solomon
RS_CASES = [
    # (D, num_messages, L, N, mask_len)
    (1, 1, 1, 1, 0), (1, 3, 4, 8, 0), (1, 16, 16, 64, 0), (1, 2, 5, 16, 0), (1, 5, 32, 32, 0),
    (1, 2, 0, 4, 0), (1, 3, 7, 16, 1), (1, 16, 1 << 10, 1 << 12, 0), (1, 4, 1 << 11, 1 << 11, 0),
    (1, 3, 1 << 12, 1 << 13, 0), (1, 2, 3000, 1 << 13, 0), (1, 16, 1 << 13, 1 << 15, 0),
    (2, 3, 8, 32, 0), (2, 16, 1 << 9, 1 << 14, 0), (2, 2, 1 << 12, 1 << 14, 0), (2, 4, 100, 256, 3),
    (3, 3, 8, 32, 0), (3, 16, 1 << 8, 1 << 12, 0), (3, 2, 1 << 12, 1 << 13, 0),
    # one-coefficient messages (a broadcast, no transform), and one with a mask (transform)
    (1, 3, 1, 64, 0), (2, 16, 1, 1 << 14, 0), (3, 4, 1, 256, 0), (2, 4, 1, 64, 2),
]


@pytest.mark.parametrize("D,M,L,N,ml", RS_CASES)
def test_interleaved_encode_matches_oracle(D, M, L, N, ml):
    seed = hash((D, M, L, N, ml)) & 0xFFFF
    msgs = rnd_np(seed, M * L * D).reshape(M, L * D)
    masks = rnd_np(seed + 1, M * ml * D) if ml else None
    got = W.interleaved_rs_encode(list(msgs), masks, N, degree=D)
    exp = O.interleaved_encode(D, msgs, N, masks=masks.reshape(M, ml * D) if ml else None)
    assert got.shape == exp.shape
    assert np.array_equal(got, exp)


def test_interleaved_encode_rejects_bad_lengths():
    with pytest.raises(W.WhirError):
        W.interleaved_rs_encode([np.zeros(10, np.uint64)], None, 8)
    with pytest.raises(W.WhirError):
        W.interleaved_rs_encode([np.zeros(4, np.uint64)], None, 5)


def test_interleaved_encode_full_size_property():
    # The headline shape: 16 messages of 2^20, N = 2^22 (rate 1/4), device resident.
    D, M, L, N = 1, 16, 1 << 20, 1 << 22
    v = rnd_np(7, M * L)
    dm = W.DeviceBuffer.from_array(v)
    dout = W.DeviceBuffer(M * N * 8)
    W.rs_encode_dev(D, dm, M, L, N, dout, row_major=False)
    cw = dout.to_array()
    rng = random.Random(3)
    rows = [0, 1, N - 1] + [rng.randrange(N) for _ in range(5)]
    pts = O.evaluation_points(L, N, rows)
    for r, x in zip(rows, pts):
        for j in (0, 7, 15):
            exp = O.univariate_evaluate(1, 1, v[j * L:(j + 1) * L], [int(x)])
            assert int(cw[j * N + r]) == int(exp[0])
    # linearity: encode(a + b) == encode(a) + encode(b) on one message
    a, b = rnd_np(8, L), rnd_np(9, L)
    s = O.add(1, a, b)
    enc = []
    for vec in (a, b, s):
        d1 = W.DeviceBuffer.from_array(vec)
        do = W.DeviceBuffer(N * 8)
        W.rs_encode_dev(1, d1, 1, L, N, do)
        enc.append(do.to_array())
    assert np.array_equal(O.add(1, enc[0], enc[1]), enc[2])


# ------------------------------------------------------------------ hashes
HASH_SIZES = [1, 8, 31, 32, 63, 64, 65, 128, 135, 136, 137, 192, 256, 384, 512, 1023, 1024, 1025, 2100, 5000]


@pytest.mark.parametrize("engine", [W.HASH_BLAKE3, W.HASH_SHA3, W.HASH_SHA2, W.HASH_KECCAK])
def test_hash_many_matches_oracle(engine):
    rng = random.Random(engine)
    for size in HASH_SIZES:
        count = rng.choice([1, 3, 17, 130])
        data = bytes(rng.randrange(256) for _ in range(size * count))
        got = W.Engine(engine).hash_many(size, data)
        assert got == O.hash_many(engine, size, data, count), (engine, size)


def test_preferred_batch_size():
    """HashEngine::preferred_batch_size (src/hash/mod.rs:67-77): one message per lane,
    1024 per CU of the current device; an unknown engine is an error."""
    import torch
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    for e in (W.HASH_COPY, W.HASH_SHA2, W.HASH_KECCAK, W.HASH_SHA3, W.HASH_BLAKE3):
        assert W.Engine(e).preferred_batch_size() == 1024 * cus
    with pytest.raises(W.WhirError):
        W.Engine(99).preferred_batch_size()


def test_hash_many_copy_engine_and_empty():
    data = bytes(range(24)) * 5
    assert W.Engine(W.HASH_COPY).hash_many(24, data) == O.hash_many(0, 24, data, 5)
    # size 0: EMPTY_HASH per message (blake3_engine.rs:102-105)
    assert W.Engine(W.HASH_BLAKE3).hash_many_count(0, b"", 4) == O.hash_many(4, 0, b"", 4)
    assert W.Engine(W.HASH_SHA3).hash_many_count(0, b"", 2) == hashlib.sha3_256(b"").digest() * 2


@pytest.mark.parametrize("engine", [W.HASH_BLAKE3, W.HASH_SHA3])
@pytest.mark.parametrize("n", [1, 2, 3, 5, 64, 1000, 1 << 12, (1 << 13) + 7])
def test_merkle_commit_matches_oracle(engine, n):
    rng = random.Random(n * 10 + engine)
    leaves = bytes(rng.randrange(256) for _ in range(32 * n))
    assert W.merkle_commit(engine, leaves) == O.merkle_commit(engine, leaves)


@pytest.mark.parametrize("engine,n", [(W.HASH_BLAKE3, 1 << 19), (W.HASH_BLAKE3, 1 << 20),
                                      (W.HASH_BLAKE3, (1 << 20) + 3), (W.HASH_BLAKE3, (1 << 21) + 9),
                                      (W.HASH_SHA3, (1 << 22) + 5)])
def test_merkle_commit_wide_levels_match_oracle(engine, n):
    """Trees with levels of 2^19 nodes and more: three-level launches of eight nodes
    per lane (2^20 -> 2^17, and for the 2^23-wide padded tree 2^23 -> 2^20 -> 2^17),
    then the block-wide subtree launches above them (for BLAKE3, one quad of lanes per
    node once a launch has at most 2^16 lanes: 2^22 -> 2^19 -> 2^16 and then quads for
    the 2^21 + 9 tree; the 2^20 + 3 tree's launch from 2^21 -> 2^18 still takes one
    lane per node)."""
    rng = np.random.default_rng(n)
    leaves = rng.integers(0, 256, size=32 * n, dtype=np.uint8).tobytes()
    assert W.merkle_commit(engine, leaves) == O.merkle_commit(engine, leaves)


@pytest.mark.parametrize("D", [1, 2, 3])
@pytest.mark.parametrize("engine", [W.HASH_BLAKE3, W.HASH_SHA3])
def test_matrix_commit_and_column_leaves(D, engine):
    rows, cols = 1000, 16
    m = rnd_np(D * 100 + engine, rows * cols * D)
    exp_leaves = O.hash_rows(engine, D, m, rows, cols)
    exp_nodes = O.merkle_commit(engine, exp_leaves)
    assert W.matrix_commit(engine, D, m, rows, cols) == exp_nodes
    # column-major codeword -> leaves (the layout the prover keeps on device)
    cm = np.ascontiguousarray(m.reshape(rows, cols, D).transpose(1, 0, 2)).reshape(-1)
    dcw = W.DeviceBuffer.from_array(cm)
    dl = W.DeviceBuffer(32 * rows)
    W.hash_rows_cols_dev(engine, D, dcw, rows, cols, dl)
    assert dl.to_array(np.uint8).tobytes() == exp_leaves


@pytest.mark.parametrize("D,rows,cols", [(1, 4, 8), (1, 1 << 12, 16), (2, 1 << 11, 16), (3, 1 << 10, 8),
                                         (2, 8, 4), (1, 1000, 16), (2, 1 << 12, 12), (1, 2, 8), (3, 64, 16)])
@pytest.mark.parametrize("engine", [W.HASH_BLAKE3, W.HASH_SHA3])
def test_matrix_commit_cols_dev(D, rows, cols, engine):
    """Leaves + tree of a column-major codeword (fused BLAKE3 leaf/two-level kernel for
    power-of-two rows and 64-byte-multiple rows; the staged path otherwise)."""
    m = rnd_np(D * 1000 + rows + cols + engine, rows * cols * D)
    exp = O.merkle_commit(engine, O.hash_rows(engine, D, m, rows, cols))
    cm = np.ascontiguousarray(m.reshape(rows, cols, D).transpose(1, 0, 2)).reshape(-1)
    dcw = W.DeviceBuffer.from_array(cm)
    dn = W.DeviceBuffer(len(exp))
    W.matrix_commit_cols_dev(engine, D, dcw, rows, cols, dn)
    assert dn.to_array(np.uint8).tobytes() == exp


# ------------------------------------------------------------------ sumcheck and covectors
@pytest.mark.parametrize("D", [1, 2, 3])
@pytest.mark.parametrize("n", [0, 1, 2, 3, 5, 8, 100, 1 << 12, 5000])
def test_sumcheck_polynomial_matches_oracle(D, n):
    a, b = rnd_np(n + D, n * D), rnd_np(n + D + 99, n * D)
    c0, c2 = W.compute_sumcheck_polynomial(D, a, b)
    e0, e2 = O.sumcheck_poly(D, a, b)
    assert np.array_equal(c0, e0) and np.array_equal(c2, e2)


@pytest.mark.parametrize("D", [1, 2, 3])
@pytest.mark.parametrize("n", [0, 1, 2, 3, 7, 16, 1000, 1 << 12])
def test_fold_and_fused_fold_poly(D, n):
    a, b = rnd_np(n * 7 + D, n * D), rnd_np(n * 7 + D + 1, n * D)
    w = rnd_np(n + 5, D)
    assert np.array_equal(W.fold(D, a, w), O.fold(D, a, w))
    fa, fb, c0, c2 = W.fold_and_compute_polynomial(D, a, b, w)
    ea, eb = O.fold(D, a, w), O.fold(D, b, w)
    assert np.array_equal(fa, ea) and np.array_equal(fb, eb)
    e0, e2 = O.sumcheck_poly(D, ea, eb)
    assert np.array_equal(c0, e0) and np.array_equal(c2, e2)


@pytest.mark.parametrize("D", [1, 2, 3])
@pytest.mark.parametrize("k", [0, 1, 2, 5, 11, 14])
def test_eval_eq_matches_oracle(D, k):
    acc = rnd_np(k + 1, (1 << k) * D)
    z = rnd_np(k + 2, k * D)
    s = rnd_np(k + 3, D)
    assert np.array_equal(W.eval_eq(D, acc, z, s), O.eval_eq(D, acc, z, s))


@pytest.mark.parametrize("D", [1, 2, 3])
@pytest.mark.parametrize("n,m", [(1, 1), (5, 3), (4096, 2), (5000, 17), (70000, 4)])
def test_geometric_accumulate_matches_oracle(D, n, m):
    acc = rnd_np(n + m, n * D)
    s = rnd_np(n + m + 1, m * D)
    x = rnd_np(n + m + 2, m * D)
    assert np.array_equal(W.geometric_accumulate(D, acc, s, x), O.geometric_accumulate(D, acc, s, x))


@pytest.mark.parametrize("D", [2, 3])
def test_geometric_accumulate_mixed_base_and_extension_points(D):
    # STIR constraints mix extension (OOD) points with base-field (in-domain) points.
    n, m = 70000, 9
    acc = rnd_np(11, n * D)
    s = rnd_np(12, m * D)
    x = rnd_np(13, m * D).reshape(m, D)
    x[1:, 1:] = 0  # points 1..8 lie in the base field
    x[5] = 0       # and one is zero
    x[6, 0] = 1    # and one is one
    x = x.reshape(-1)
    assert np.array_equal(W.geometric_accumulate(D, acc, s, x), O.geometric_accumulate(D, acc, s, x))


@pytest.mark.parametrize("D", [1, 2, 3])
@pytest.mark.parametrize("n", [0, 1, 2, 100, 4096, 70001])
def test_univariate_evaluate_matches_oracle(D, n):
    c = rnd_np(n + 11, n)
    x = rnd_np(n + 12, D)
    assert np.array_equal(W.mixed_univariate_evaluate(1, D, c, x), O.univariate_evaluate(1, D, c, x))
    if D > 1:
        ce = rnd_np(n + 13, n * D)
        assert np.array_equal(W.mixed_univariate_evaluate(D, D, ce, x), O.univariate_evaluate(D, D, ce, x))


@pytest.mark.parametrize("D", [1, 2, 3])
def test_lift_axpy_dot(D):
    n = 3001
    v = rnd_np(1, n)
    assert np.array_equal(W.lift(D, v), O.lift(D, v))
    acc = rnd_np(2, n * D)
    w = rnd_np(3, D)
    assert np.array_equal(W.mixed_scalar_mul_add(D, acc, w, v), O.mixed_scalar_mul_add(D, acc, w, v))
    a, b = rnd_np(4, n * D), rnd_np(5, n * D)
    assert np.array_equal(W.dot(D, a, b), O.dot(D, a, b))


# ------------------------------------------------------------------ proof of work
@pytest.mark.parametrize("engine", [W.HASH_BLAKE3, W.HASH_SHA3])
@pytest.mark.parametrize("bits", [1.0, 6.0, 12.5, 16.0])
def test_pow_minimal_nonce_matches_oracle(engine, bits):
    ch = hashlib.sha3_256(f"pow-{engine}-{bits}".encode()).digest()
    thr = O.pow_threshold(bits)
    got = W.pow_grind_bounded(engine, ch, thr, 1 << 24)
    exp = O.pow_search(engine, ch, thr, 0, 1 << 24)
    assert got == exp
    assert W.pow_grind(engine, ch, thr) == exp


def test_pow_noop_threshold():
    assert W.pow_grind(W.HASH_BLAKE3, b"\0" * 32, 2**64 - 1) is None


def test_pow_unreachable_threshold_in_bounded_window():
    # threshold 0 is practically unreachable: the bounded search must return None.
    assert W.pow_grind_bounded(W.HASH_BLAKE3, b"\1" * 32, 0, 1 << 16) is None
