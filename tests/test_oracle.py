"""Pins the CPU oracle (oracle/whir_oracle.c) before it is trusted as the checker.

Each test checks an oracle function against something independent of it:
Python big-int arithmetic, hashlib, LLVM's BLAKE3 vectors, closed forms, brute
force, or a known-answer test taken from the reference crate's own tests.
"""
import hashlib
import json
import os
import random

import numpy as np
import pytest

from oracle import oracle as O

P = O.P
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def rnd(rng, n):
    return [rng.randrange(P) for _ in range(n)]


# ------------------------------------------------------------------ field
EDGE = [0, 1, 2, P - 1, P - 2, 2**32, 2**32 - 1, 2**32 + 1, 2**63, 2**64 - 2**32, P // 2]


def test_field_base_ops_vs_bigint():
    rng = random.Random(1)
    xs = EDGE + rnd(rng, 300)
    ys = EDGE[::-1] + rnd(rng, 300)
    a = np.array(xs, dtype=np.uint64)
    b = np.array(ys, dtype=np.uint64)
    assert [int(v) for v in O.add(1, a, b)] == [(x + y) % P for x, y in zip(xs, ys)]
    assert [int(v) for v in O.sub(1, a, b)] == [(x - y) % P for x, y in zip(xs, ys)]
    assert [int(v) for v in O.mul(1, a, b)] == [(x * y) % P for x, y in zip(xs, ys)]
    nz = np.array([x for x in xs if x], dtype=np.uint64)
    inv = O.inv(nz)
    assert all((int(x) * int(y)) % P == 1 for x, y in zip(nz, inv))


def _ext_mul_ref(D, a, b):
    w = {2: 7, 3: 2}[D]
    t = [0] * (2 * D - 1)
    for i in range(D):
        for j in range(D):
            t[i + j] += a[i] * b[j]
    for k in range(2 * D - 2, D - 1, -1):
        t[k - D] += w * t[k]
    return [x % P for x in t[:D]]


@pytest.mark.parametrize("D", [2, 3])
def test_field_ext_mul_vs_polynomial_product(D):
    rng = random.Random(D)
    n = 200
    A = [rnd(rng, D) for _ in range(n)] + [[P - 1] * D, [0] * D, [1] + [0] * (D - 1)]
    B = [rnd(rng, D) for _ in range(n)] + [[P - 1] * D, [5] * D, [0] * (D - 1) + [1]]
    out = O.mul(D, np.array(A, dtype=np.uint64), np.array(B, dtype=np.uint64))
    for x, y, z in zip(A, B, out):
        assert [int(v) for v in z] == _ext_mul_ref(D, x, y)


@pytest.mark.parametrize("D", [2, 3])
def test_field_ext_embedding_homomorphism(D):
    # Basefield embedding laws, src/algebra/embedding.rs:12-18.
    rng = random.Random(10 + D)
    a, b = rnd(rng, 50), rnd(rng, 50)
    la = O.lift(D, a).reshape(-1, D)
    lb = O.lift(D, b).reshape(-1, D)
    prod = O.mul(D, la, lb)
    assert np.array_equal(prod, O.lift(D, O.mul(1, a, b)).reshape(-1, D))


def test_u_power_identities():
    # u^2 = 7 in Field64_2, u^3 = 2 in Field64_3 (fields.rs:98, 114).
    assert [int(x) for x in O.pow_(2, [0, 1], 2).reshape(-1)] == [7, 0]
    assert [int(x) for x in O.pow_(3, [0, 1, 0], 3).reshape(-1)] == [2, 0, 0]


# ------------------------------------------------------------------ NTT
def test_root_kats():
    # cooley_tukey.rs:582-595: omega_order == GENERATOR^((p-1)/order), order = 3*2^32.
    order = 3 << 32
    assert O.ntt_root(order) == pow(7, (P - 1) // order, P)
    # 64th root of unity is a power of two: root(64) = 8^13 = 2^39, and 2^96 = -1.
    assert O.ntt_root(64) == pow(8, 13, P) == pow(2, 39, P)
    assert pow(2, 96, P) == P - 1
    for k in range(1, 33):
        w = O.ntt_root(1 << k)
        assert pow(w, 1 << k, P) == 1 and pow(w, 1 << (k - 1), P) == P - 1
    w3 = O.ntt_root(3)
    assert w3 != 1 and pow(w3, 3, P) == 1
    with pytest.raises(ValueError):
        O.ntt_root(5)


def test_next_order_and_divisors():
    # divisors(3*2^32, [2, 3]) sorted (ntt/utils.rs:69-84) then binary search.
    assert O.next_order(0) == 1
    assert O.next_order(1) == 1
    assert O.next_order(5) == 6
    assert O.next_order(7) == 8
    assert O.next_order(13) == 16
    assert O.next_order(17) == 24
    assert O.next_order(3 << 32) == 3 << 32
    assert O.next_order((3 << 32) + 1) is None


def _dft_bigint(vals, n):
    w = O.ntt_root(n)
    return [sum(vals[j] * pow(w, j * k, P) for j in range(n)) % P for k in range(n)]


@pytest.mark.parametrize("n", [2, 3, 4, 6, 8, 12])
def test_dft_direct_vs_bigint(n):
    rng = random.Random(n)
    v = rnd(rng, n)
    assert [int(x) for x in O.dft_direct(1, v)] == _dft_bigint(v, n)


@pytest.mark.parametrize("n", [1, 2, 3, 4, 6, 8, 12, 16, 24, 32, 48, 64, 96, 128, 256, 512, 1024, 2048, 4096])
def test_ntt_batch_vs_direct_dft(n):
    # test_ntt_batch_size_* of cooley_tukey.rs:842-1068: engine output == direct DFT.
    rng = random.Random(100 + n)
    batch = 3
    v = np.array(rnd(rng, n * batch), dtype=np.uint64)
    got = O.ntt_batch(1, v, n)
    for b in range(batch):
        assert np.array_equal(got[b * n:(b + 1) * n], O.dft_direct(1, v[b * n:(b + 1) * n]))


@pytest.mark.parametrize("D", [2, 3])
def test_ntt_ext_is_componentwise(D):
    rng = random.Random(7 * D)
    n = 64
    v = np.array(rnd(rng, n * D), dtype=np.uint64).reshape(n, D)
    got = O.ntt_batch(D, v, n).reshape(n, D)
    for c in range(D):
        assert np.array_equal(got[:, c], O.ntt_batch(1, np.ascontiguousarray(v[:, c]), n))


def test_evaluation_points_closed_form_and_unique():
    N, L = 64, 16
    C, Q = 16, 4
    g = O.generator(N)
    assert g == pow(7, (P - 1) // N, P)
    idx = list(range(N))
    pts = [int(x) for x in O.evaluation_points(L, N, idx)]
    for r in idx:
        c, t = divmod(r, C)
        assert pts[r] == pow(g, c + t * Q, P)
    assert len(set(pts)) == N


def _horner(coeffs, x):
    acc = 0
    for c in reversed(coeffs):
        acc = (acc * x + c) % P
    return acc


@pytest.mark.parametrize("num,L,N,mask", [(1, 1, 1, 0), (3, 4, 8, 0), (16, 16, 64, 0), (2, 5, 16, 0),
                                          (4, 3, 6, 0), (2, 6, 24, 2), (5, 32, 32, 0), (0, 4, 8, 0),
                                          (2, 0, 4, 0), (3, 7, 12, 1)])
def test_interleaved_encode_vs_brute_force(num, L, N, mask):
    # Property test of src/algebra/ntt/mod.rs:158-227: every codeword entry equals the
    # univariate evaluation of message || mask at evaluation_points(index).
    rng = random.Random(num * 1000 + L * 10 + N)
    msgs = [rnd(rng, L) for _ in range(num)]
    masks = [rnd(rng, mask) for _ in range(num)]
    cw = O.interleaved_encode(1, np.array(msgs, dtype=np.uint64).reshape(num, L),
                              N, masks=np.array(masks, dtype=np.uint64).reshape(num, mask) if mask else None)
    assert cw.shape == (N, num)
    if num == 0:
        return
    pts = [int(x) for x in O.evaluation_points(L + mask, N, list(range(N)))]
    for r in range(N):
        for j in range(num):
            assert int(cw[r, j]) == _horner(msgs[j] + masks[j], pts[r])


@pytest.mark.parametrize("D", [2, 3])
def test_interleaved_encode_ext_componentwise(D):
    rng = random.Random(33 + D)
    num, L, N = 4, 8, 32
    m = np.array(rnd(rng, num * L * D), dtype=np.uint64).reshape(num, L, D)
    cw = O.interleaved_encode(D, m.reshape(num, L * D), N).reshape(N, num, D)
    for c in range(D):
        base = O.interleaved_encode(1, np.ascontiguousarray(m[:, :, c]), N)
        assert np.array_equal(cw[:, :, c], base)


# ------------------------------------------------------------------ hashes
def test_blake3_vs_llvm_vectors():
    with open(os.path.join(GOLDEN, "blake3_vectors.json")) as f:
        cases = json.load(f)["cases"]
    for c in cases:
        data = bytes(i % 251 for i in range(c["input_len"]))
        assert O.blake3(data).hex() == c["hash"], c["input_len"]


def test_blake3_empty_hash_kat():
    # EMPTY_HASH, src/hash/blake3_engine.rs:23-25
    assert O.blake3(b"").hex() == "af1349b9f5f9a1a6a0404dea36dcc9499bcb25c9adc112b7cc9a93cae41f3262"
    assert O.hash_many(O.ENGINE_BLAKE3, 0, b"", 3) == bytes.fromhex(
        "af1349b9f5f9a1a6a0404dea36dcc9499bcb25c9adc112b7cc9a93cae41f3262") * 3


@pytest.mark.parametrize("n", [0, 1, 64, 135, 136, 137, 200, 271, 272, 273, 1000])
def test_sha3_sha2_vs_hashlib(n):
    data = bytes((i * 7 + 3) % 256 for i in range(n))
    assert O.sha3_256(data) == hashlib.sha3_256(data).digest()
    assert O.sha256(data) == hashlib.sha256(data).digest()


def test_keccak256_kat():
    assert O.keccak256(b"").hex() == "c5d2460186f7233c927e7db2dcc703c0e500b653ca82273b7bfad8045d85a470"


def test_engine_ids_kat():
    # engine_id = SHA3-256("whir::hash" || oid-bytes or name), src/hash/mod.rs:88-103,
    # constants at blake3_engine.rs:19-21 and digest_engine.rs:11-19.
    def eid(x):
        return hashlib.sha3_256(b"whir::hash" + x).hexdigest()
    assert eid(b"Blake3") == "03e01749ebcc0477924254eb482066b864a8dd4d77252464ca6f5b6f5cc05b4c"
    assert eid(bytes.fromhex("608648016503040208")) == \
        "aa802dcdabad8ea8a1430919893b96c30e31ff5734b385999108aa202d27dc12"
    assert eid(bytes.fromhex("608648016503040201")) == \
        "018eaa247cb8957ab1e9fdac885450403c5e7bda1acaaa504e4cc8f2f76bb076"
    assert eid(b"keccak") == "ddd248964e320312a66775aee8e16c88c927734be59aca09b7af6deb0ad00e8c"


def test_hash_many_is_per_message():
    data = bytes(range(256)) * 2
    for eng, fn in [(O.ENGINE_SHA3, hashlib.sha3_256), (O.ENGINE_SHA2, hashlib.sha256)]:
        out = O.hash_many(eng, 64, data, 8)
        assert out == b"".join(fn(data[64 * i:64 * (i + 1)]).digest() for i in range(8))


# ------------------------------------------------------------------ Merkle
def _merkle_ref(leaves, h):
    n = len(leaves)
    layers = 0
    while (1 << layers) < n:
        layers += 1
    level = list(leaves) + [b"\0" * 32] * ((1 << layers) - n)
    nodes = list(level)
    while len(level) > 1:
        level = [h(level[2 * i] + level[2 * i + 1]) for i in range(len(level) // 2)]
        nodes += level
    return b"".join(nodes)


@pytest.mark.parametrize("n", [1, 2, 3, 5, 8, 13, 64])
def test_merkle_commit_vs_recursion(n):
    leaves = [hashlib.sha3_256(bytes([i])).digest() for i in range(n)]
    got = O.merkle_commit(O.ENGINE_SHA3, b"".join(leaves))
    assert got == _merkle_ref(leaves, lambda x: hashlib.sha3_256(x).digest())


def test_layers_for_size_kat():
    # merkle_tree.rs:360-371
    assert [O.layers_for_size(i) for i in range(9)] == [0, 0, 1, 2, 2, 3, 3, 3, 3]


def _verify_path(root_nodes, leaves, num_leaves, indices, hints, h):
    # Restates merkle_tree::Config::verify (merkle_tree.rs:188-273) over python lists.
    layers = O.layers_for_size(num_leaves)
    pairs = sorted(set((i, leaves[i]) for i in indices))
    idx = [i for i, _ in pairs]
    hs = [x for _, x in pairs]
    hint_pos = 0
    for _ in range(layers):
        ni, nh = [], []
        k = 0
        while k < len(idx):
            a = idx[k]
            if k + 1 < len(idx) and idx[k + 1] == a ^ 1:
                nh.append(h(hs[k] + hs[k + 1]))
                k += 2
            else:
                sib = hints[hint_pos:hint_pos + 32]
                hint_pos += 32
                nh.append(h(hs[k] + sib) if a % 2 == 0 else h(sib + hs[k]))
                k += 1
            ni.append(a >> 1)
        idx, hs = ni, nh
    return idx == [0] and hs[0] == root_nodes[-32:] and hint_pos == len(hints)


def test_merkle_open_round_trip():
    # merkle_tree.rs:325-358: 256 leaves Hash([i;32]), opening [13, 42].
    n = 256
    leaves = [bytes([i]) * 32 for i in range(n)]
    h = lambda x: bytes.fromhex(O.hash_many(O.ENGINE_BLAKE3, 64, x, 1).hex())
    nodes = O.merkle_commit(O.ENGINE_BLAKE3, b"".join(leaves))
    for idx in ([13, 42], [0, 1, 2, 3], [255, 254, 7, 7, 100], [5]):
        hints = O.merkle_open(nodes, n, idx)
        assert _verify_path(nodes, leaves, n, idx, hints, h)


def test_hash_rows_encoding():
    # ArkFieldEncoder: canonical little-endian 8 bytes per base coefficient (matrix_commit.rs:130-156)
    rng = random.Random(5)
    rows, cols, D = 4, 3, 2
    m = np.array(rnd(rng, rows * cols * D), dtype=np.uint64)
    got = O.hash_rows(O.ENGINE_SHA3, D, m, rows, cols)
    ref = b"".join(hashlib.sha3_256(m[r * cols * D:(r + 1) * cols * D].astype("<u8").tobytes()).digest()
                   for r in range(rows))
    assert got == ref


# ------------------------------------------------------------------ sumcheck
def _zero_pad(v, D):
    n = len(v) // D
    p2 = 1
    while p2 < n:
        p2 *= 2
    return list(v) + [0] * ((p2 - n) * D) if n else []


@pytest.mark.parametrize("D", [1, 2])
@pytest.mark.parametrize("n", [0, 1, 2, 3, 5, 8, 13, 64, 100])
def test_sumcheck_poly_closed_form_and_padding(D, n):
    rng = random.Random(n * 3 + D)
    a = rnd(rng, n * D)
    b = rnd(rng, n * D)
    c0, c2 = O.sumcheck_poly(D, a, b)
    pa, pb = _zero_pad(a, D), _zero_pad(b, D)
    # zero-padding invariants, src/algebra/sumcheck.rs:160-173
    for x, y in [(pa, pb), (pa, b), (a, pb)]:
        e0, e2 = O.sumcheck_poly(D, x, y)
        assert np.array_equal(e0, c0) and np.array_equal(e2, c2)
    if n >= 2:
        # closed form: c0 = sum_{i<h} a_i b_i ; c2 = sum (a_{h+i}-a_i)(b_{h+i}-b_i) over padded vectors
        h = len(pa) // D // 2
        A = np.array(pa, dtype=np.uint64).reshape(-1, D)
        B = np.array(pb, dtype=np.uint64).reshape(-1, D)
        assert np.array_equal(c0, O.dot(D, A[:h], B[:h]))
        assert np.array_equal(c2, O.dot(D, O.sub(D, A[h:], A[:h]), O.sub(D, B[h:], B[:h])))
        # p(0) + p(1) = dot(a, b), where p(1) = sum a_{h+i} b_{h+i}
        s = O.dot(D, A, B)
        c1 = O.sub(D, O.sub(D, O.sub(D, s, c0), c0), c2).reshape(-1)
        p1 = O.add(D, O.add(D, c0, c1), c2).reshape(-1)
        assert np.array_equal(p1, O.dot(D, A[h:], B[h:]))


@pytest.mark.parametrize("n", [0, 1, 2, 3, 7, 16, 33])
def test_fold_zero_extend(n):
    # src/algebra/sumcheck.rs:175-188
    rng = random.Random(n + 50)
    v = rnd(rng, n)
    w = rng.randrange(P)
    f1 = O.fold(1, v, [w])
    f2 = O.fold(1, _zero_pad(v, 1), [w])
    assert np.array_equal(f1, f2)
    if n:
        m = len(_zero_pad(v, 1))
        assert f1.size == max(m // 2, 1)


@pytest.mark.parametrize("D", [1, 2, 3])
def test_fold_to_one_is_multilinear_extension(D):
    # src/protocols/sumcheck.rs:316-320; brute-force MLE with MSB-first variables.
    rng = random.Random(60 + D)
    k = 5
    v = rnd(rng, (1 << k) * D)
    pt = rnd(rng, k * D)
    cur = np.array(v, dtype=np.uint64)
    for j in range(k):
        cur = O.fold(D, cur, pt[j * D:(j + 1) * D])
    mle = O.multilinear_extend(D, v, pt)
    assert np.array_equal(cur.reshape(-1), mle)
    if D == 1:
        # brute force: sum_x v[x] * prod_i (x_i z_i + (1-x_i)(1-z_i)), x_0 = MSB
        tot = 0
        for x in range(1 << k):
            w = 1
            for i in range(k):
                bit = (x >> (k - 1 - i)) & 1
                w = w * (pt[i] if bit else (1 - pt[i])) % P
            tot += v[x] * w
        assert int(mle[0]) == tot % P


@pytest.mark.parametrize("D", [1, 2])
def test_eval_eq_brute_force_and_mle(D):
    rng = random.Random(70 + D)
    k = 6
    z = rnd(rng, k * D)
    s = rnd(rng, D)
    acc0 = rnd(rng, (1 << k) * D)
    got = O.eval_eq(D, acc0, z, s).reshape(-1, D)
    for x in range(1 << k):
        w = np.array(s, dtype=np.uint64)
        for i in range(k):
            bit = (x >> (k - 1 - i)) & 1
            zi = np.array(z[i * D:(i + 1) * D], dtype=np.uint64)
            one = np.zeros(D, dtype=np.uint64)
            one[0] = 1
            f = zi if bit else O.sub(D, one, zi).reshape(-1)
            w = O.mul(D, w, f).reshape(-1)
        exp = O.add(D, np.array(acc0[x * D:(x + 1) * D], dtype=np.uint64), w).reshape(-1)
        assert np.array_equal(got[x], exp)
    # dot(eq(z), v) = MLE(v, z)
    v = rnd(rng, (1 << k) * D)
    one = [1] + [0] * (D - 1)
    eq = O.eval_eq(D, np.zeros((1 << k) * D, dtype=np.uint64), z, one)
    assert np.array_equal(O.dot(D, eq, v), O.multilinear_extend(D, v, z))


def test_geometric_accumulate_closed_form():
    rng = random.Random(80)
    n = 200
    xs = rnd(rng, 4)
    ss = rnd(rng, 4)
    acc0 = rnd(rng, n)
    got = O.geometric_accumulate(1, acc0, ss, xs)
    for i in range(0, n, 17):
        exp = (acc0[i] + sum(s * pow(x, i, P) for s, x in zip(ss, xs))) % P
        assert int(got[i]) == exp
    # full sum: sum_i x^i = (x^n - 1)/(x - 1)
    got0 = O.geometric_accumulate(1, [0] * n, [1], [xs[0]])
    tot = sum(int(v) for v in got0) % P
    assert tot == (pow(xs[0], n, P) - 1) * pow(xs[0] - 1, P - 2, P) % P


@pytest.mark.parametrize("D", [2, 3])
def test_univariate_evaluate_vs_bigint(D):
    rng = random.Random(90 + D)
    n = 77
    c = rnd(rng, n)
    x = rnd(rng, D)
    got = O.univariate_evaluate(1, D, c, x)
    acc = [0] * D
    for ci in reversed(c):
        acc = _ext_mul_ref(D, acc, x)
        acc[0] = (acc[0] + ci) % P
    assert [int(v) for v in got] == acc


# ------------------------------------------------------------------ PoW, indices
def test_pow_threshold_kats():
    # src/protocols/proof_of_work.rs:238-245
    assert O.pow_threshold(0.0) == 2**64 - 1
    assert O.pow_threshold(60.0) == 1 << 4
    for bits in range(1, 61):
        assert O.pow_threshold(float(bits)) == 1 << (64 - bits)


def test_pow_minimal_nonce_brute_force():
    ch = hashlib.sha3_256(b"challenge").digest()
    thr = O.pow_threshold(8.0)
    got = O.pow_search(O.ENGINE_SHA3, ch, thr, 0, 1 << 20)
    n = 0
    while True:
        h = hashlib.sha3_256(ch + n.to_bytes(8, "little") + b"\0" * 24).digest()
        if int.from_bytes(h[:8], "little") <= thr:
            break
        n += 1
    assert got == n


def test_challenge_indices_mock_sponge_kats():
    # src/protocols/challenge_indices.rs:57-203 (MockSponge squeeze bytes)
    sq = bytes([0x01, 0x23, 0x45, 0x67, 0x89, 0xAB, 0xCD, 0xEF, 0x12, 0x34,
                0x56, 0x78, 0x9A, 0xBC, 0xDE, 0xF0, 0x11, 0x22, 0x33, 0x44,
                0x55, 0x66, 0x77, 0x88, 0x99])
    assert O.bytes_to_indices(sq, 5, 1 << 7) == [x % 128 for x in sq[:5]]
    two = O.bytes_to_indices(sq, 5, 1 << 13)
    assert two == [((sq[2 * i] << 8) | sq[2 * i + 1]) % (1 << 13) for i in range(5)]
    three = O.bytes_to_indices(sq, 5, 1 << 20)
    assert three == [int.from_bytes(sq[3 * i:3 * i + 3], "big") % (1 << 20) for i in range(5)]


# ------------------------------------------------------------------ large sizes (OpenMP paths)
# The oracle splits its large loops over threads (chunked Horner, chunked
# geometric powers, split eq tree, partial sums).  These sizes take those paths;
# each result is pinned by big-int arithmetic or a closed form, and must not
# depend on the thread count.
@pytest.fixture
def oracle_threads():
    before = O.threads()
    yield
    O.threads(before)


@pytest.mark.parametrize("nth", [1, 8])
def test_large_ntt_deltas_and_sampled_outputs(nth, oracle_threads):
    O.threads(nth)
    n = 1 << 17
    w = O.ntt_root(n)
    rng = random.Random(500)
    v = np.array(rnd(rng, n), dtype=np.uint64)
    got = O.ntt_batch(1, v, n)
    vi = [int(x) for x in v]
    for k in (0, 1, 12345, n - 1):
        wk = pow(w, k, P)
        acc = 0
        for x in reversed(vi):
            acc = (acc * wk + x) % P
        assert int(got[k]) == acc
    j = 777
    d = np.zeros(n, dtype=np.uint64)
    d[j] = 1
    gd = O.ntt_batch(1, d, n)
    for k in (0, 3, 99999, n - 1):
        assert int(gd[k]) == pow(w, j * k, P)


@pytest.mark.parametrize("nth", [1, 8])
def test_large_hash_many_and_rows(nth, oracle_threads):
    O.threads(nth)
    rng = np.random.default_rng(501)
    data = rng.integers(0, 256, size=1000 * 48, dtype=np.uint8).tobytes()
    out = O.hash_many(O.ENGINE_SHA3, 48, data, 1000)
    for i in (0, 1, 511, 999):
        assert out[32 * i:32 * i + 32] == hashlib.sha3_256(data[48 * i:48 * i + 48]).digest()
    m = rng.integers(0, P, size=(700, 6), dtype=np.uint64)
    rows = O.hash_rows(O.ENGINE_SHA3, 2, m, 700, 3)
    for r in (0, 350, 699):
        assert rows[32 * r:32 * r + 32] == hashlib.sha3_256(m[r].astype("<u8").tobytes()).digest()


@pytest.mark.parametrize("nth", [1, 8])
def test_large_sumcheck_fold_dot(nth, oracle_threads):
    O.threads(nth)
    rng = random.Random(502)
    n = 12000  # ragged: padded to 16384, half = 8192, 3808 upper entries
    a, b = rnd(rng, n), rnd(rng, n)
    c0, c2 = O.sumcheck_poly(1, a, b)
    h = 8192
    pa, pb = a + [0] * (2 * h - n), b + [0] * (2 * h - n)
    assert int(c0[0]) == sum(pa[i] * pb[i] for i in range(h)) % P
    assert int(c2[0]) == sum((pa[h + i] - pa[i]) * (pb[h + i] - pb[i]) for i in range(h)) % P
    assert int(O.dot(1, a, b)[0]) == sum(x * y for x, y in zip(a, b)) % P
    w = rng.randrange(P)
    f = O.fold(1, a, [w])
    assert f.size == h
    for i in (0, 100, n - h - 1, n - h, h - 1):
        assert int(f[i]) == (pa[i] + w * (pa[h + i] - pa[i])) % P


@pytest.mark.parametrize("nth", [1, 8])
def test_large_eval_eq_sampled(nth, oracle_threads):
    O.threads(nth)
    rng = random.Random(503)
    k = 10
    z, s = rnd(rng, k), rng.randrange(P)
    got = O.eval_eq(1, np.zeros(1 << k, dtype=np.uint64), z, [s])
    for x in (0, 1, 63, 64, 513, 1023):
        e = s
        for i in range(k):
            e = e * (z[i] if (x >> (k - 1 - i)) & 1 else 1 - z[i]) % P
        assert int(got[x]) == e


@pytest.mark.parametrize("nth", [1, 8])
def test_large_geometric_and_univariate(nth, oracle_threads):
    O.threads(nth)
    rng = random.Random(504)
    n = 10000  # three chunks of 4096, the last ragged
    xs, ss = rnd(rng, 3), rnd(rng, 3)
    got = O.geometric_accumulate(1, [0] * n, ss, xs)
    for i in (0, 4095, 4096, 8191, 8192, n - 1):
        assert int(got[i]) == sum(s * pow(x, i, P) for s, x in zip(ss, xs)) % P
    tot = sum(int(v) for v in O.geometric_accumulate(1, [0] * n, [1], [xs[0]])) % P
    assert tot == (pow(xs[0], n, P) - 1) * pow(xs[0] - 1, P - 2, P) % P
    m = (1 << 17) + 1234  # three Horner chunks of 2^16, the last ragged
    c = rnd(rng, m)
    x = rng.randrange(P)
    acc = 0
    for ci in reversed(c):
        acc = (acc * x + ci) % P
    assert int(O.univariate_evaluate(1, 1, c, [x])[0]) == acc
