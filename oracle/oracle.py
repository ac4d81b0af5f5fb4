"""Python handle on the CPU oracle (oracle/whir_oracle.c).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker.  The product package
(whir_amd) never imports this module.

Element layout: numpy uint64 arrays of shape (n, D) (or flat n*D), canonical
coefficients lowest degree first; D = 1 (Field64), 2 (Field64_2), 3 (Field64_3).

How each oracle function is pinned (each check is a test under tests/):

  field arithmetic      Python big-int arithmetic mod p on random and edge values;
                        extension mul checked against polynomial product mod u^D - W.
  ntt_batch             direct O(n^2) DFT (itself checked against Python big-int
                        sums); root(order) KATs: omega_order = 7^((p-1)/(3*2^32)),
                        root(64) = 8^13.
  interleaved_encode    brute-force univariate evaluation at evaluation_points
                        (the reference property test, src/algebra/ntt/mod.rs:158-227).
  evaluation_points     closed form generator^(c + t*Q); uniqueness.
  blake3                LLVM's independent BLAKE3 (libclang-cpp.so) vectors committed
                        in tests/golden/blake3_vectors.json; EMPTY_HASH KAT of
                        src/hash/blake3_engine.rs:23-25.
  sha3_256 / sha256     hashlib.
  keccak256             keccak256("") KAT.
  merkle commit/open    Python recursion over hashlib SHA3; sibling-path re-verification.
  sumcheck poly / fold  closed forms: c0 = sum a_i b_i over the low half, the
                        zero-padding invariants of src/algebra/sumcheck.rs:160-188,
                        and fold-to-one == multilinear extension (MSB first,
                        src/protocols/sumcheck.rs:316-320).
  eval_eq               dot(eq_weights(z), v) == multilinear_extend(v, z); brute-force
                        product formula on small k.
  geometric_accumulate  closed-form geometric sums (x != 1) and brute force.
  univariate_evaluate   Python big-int Horner.
  pow_search            brute force with hashlib SHA3.
  bytes_to_indices      MockSponge KATs of src/protocols/challenge_indices.rs:57-203.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

P = 0xFFFFFFFF00000001
_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")

_lib = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


u64p = ctypes.POINTER(ctypes.c_uint64)
u8p = ctypes.POINTER(ctypes.c_uint8)
szp = ctypes.POINTER(ctypes.c_size_t)
sz = ctypes.c_size_t
ci = ctypes.c_int


def _declare(L):
    L.oracle_field_op.argtypes = [ci, ci, u64p, u64p, u64p, sz]
    L.oracle_ntt_root.argtypes = [ctypes.c_uint64, u64p]
    L.oracle_next_order.argtypes = [ctypes.c_uint64, u64p]
    L.oracle_generator.argtypes = [ctypes.c_uint64, u64p]
    L.oracle_evaluation_points.argtypes = [ctypes.c_uint64, ctypes.c_uint64, u64p, sz, u64p]
    L.oracle_ntt_batch.argtypes = [ci, u64p, sz, sz]
    L.oracle_dft_direct.argtypes = [ci, u64p, u64p, sz]
    L.oracle_interleaved_encode.argtypes = [ci, u64p, sz, sz, u64p, sz, sz, u64p]
    L.oracle_blake3.argtypes = [u8p, sz, u8p]
    L.oracle_sha3_256.argtypes = [u8p, sz, u8p]
    L.oracle_keccak256.argtypes = [u8p, sz, u8p]
    L.oracle_sha256.argtypes = [u8p, sz, u8p]
    L.oracle_keccak_f.argtypes = [u64p]
    L.oracle_hash_many.argtypes = [ci, sz, u8p, sz, u8p]
    L.oracle_layers_for_size.argtypes = [sz]
    L.oracle_layers_for_size.restype = sz
    L.oracle_merkle_commit.argtypes = [ci, u8p, sz, u8p]
    L.oracle_merkle_open.argtypes = [u8p, sz, szp, sz, u8p]
    L.oracle_merkle_open.restype = ctypes.c_long
    L.oracle_hash_rows.argtypes = [ci, ci, u64p, sz, sz, u8p]
    L.oracle_sumcheck_poly.argtypes = [ci, u64p, sz, u64p, sz, u64p, u64p]
    L.oracle_fold.argtypes = [ci, u64p, sz, u64p]
    L.oracle_fold.restype = sz
    L.oracle_eval_eq.argtypes = [ci, u64p, u64p, sz, u64p]
    L.oracle_multilinear_extend.argtypes = [ci, u64p, sz, u64p, u64p]
    L.oracle_geometric_accumulate.argtypes = [ci, u64p, sz, u64p, u64p, sz]
    L.oracle_univariate_evaluate.argtypes = [ci, ci, u64p, sz, u64p, u64p]
    L.oracle_lift.argtypes = [ci, u64p, sz, u64p]
    L.oracle_mixed_scalar_mul_add.argtypes = [ci, u64p, u64p, u64p, sz]
    L.oracle_dot.argtypes = [ci, u64p, u64p, sz, u64p]
    L.oracle_pow_search.argtypes = [ci, u8p, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64, u64p]
    L.oracle_bytes_to_indices.argtypes = [u8p, sz, sz, szp]
    L.oracle_pow_threshold.argtypes = [ctypes.c_double]
    L.oracle_pow_threshold.restype = ctypes.c_uint64
    L.oracle_threads.argtypes = [ci]
    L.oracle_threads.restype = ci


def threads(n: int = 0) -> int:
    """Set (n > 0) and return the oracle's OpenMP thread count; results do not
    depend on it."""
    return lib().oracle_threads(n)


def _p64(a: np.ndarray):
    assert a.dtype == np.uint64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(u64p)


def _p8(a: np.ndarray):
    assert a.dtype == np.uint8 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(u8p)


def _arr(x, D=None) -> np.ndarray:
    a = np.ascontiguousarray(np.asarray(x, dtype=np.uint64))
    return a


# ---------------------------------------------------------------- field ops
def field_op(D: int, op: int, a, b) -> np.ndarray:
    a = _arr(a).reshape(-1)
    b = _arr(b).reshape(-1)
    n = a.size // D
    out = np.zeros(n * D, dtype=np.uint64)
    lib().oracle_field_op(D, op, _p64(a), _p64(b), _p64(out), n)
    return out.reshape(n, D) if D > 1 else out


def add(D, a, b):
    return field_op(D, 0, a, b)


def sub(D, a, b):
    return field_op(D, 1, a, b)


def mul(D, a, b):
    return field_op(D, 2, a, b)


def pow_(D, a, e: int):
    return field_op(D, 3, a, np.array([e], dtype=np.uint64))


def inv(a):
    return field_op(1, 4, a, a)


# ---------------------------------------------------------------- NTT
def ntt_root(order: int) -> int:
    out = ctypes.c_uint64()
    if lib().oracle_ntt_root(order, ctypes.byref(out)):
        raise ValueError(f"no subgroup of order {order}")
    return out.value


def next_order(size: int):
    out = ctypes.c_uint64()
    if lib().oracle_next_order(size, ctypes.byref(out)):
        return None
    return out.value


def generator(N: int) -> int:
    out = ctypes.c_uint64()
    if lib().oracle_generator(N, ctypes.byref(out)):
        raise ValueError("unsupported codeword length")
    return out.value


def evaluation_points(masked_len: int, N: int, indices) -> np.ndarray:
    idx = np.ascontiguousarray(np.asarray(indices, dtype=np.uint64))
    out = np.zeros(idx.size, dtype=np.uint64)
    if lib().oracle_evaluation_points(masked_len, N, _p64(idx), idx.size, _p64(out)):
        raise ValueError("bad evaluation_points arguments")
    return out


def ntt_batch(D: int, values, size: int) -> np.ndarray:
    v = _arr(values).reshape(-1).copy()
    if lib().oracle_ntt_batch(D, _p64(v), v.size // D, size):
        raise ValueError("bad ntt size")
    return v


def dft_direct(D: int, values) -> np.ndarray:
    v = _arr(values).reshape(-1)
    out = np.zeros_like(v)
    if lib().oracle_dft_direct(D, _p64(v), _p64(out), v.size // D):
        raise ValueError("bad dft size")
    return out


def interleaved_encode(D: int, messages, N: int, masks=None) -> np.ndarray:
    """messages: (num_messages, L*D) uint64 -> (N, num_messages*D)."""
    m = _arr(messages)
    num = m.shape[0] if m.ndim > 1 else (1 if m.size else 0)
    m = m.reshape(-1)
    L = m.size // (D * num) if num else 0
    if masks is None:
        mk = np.zeros(1, dtype=np.uint64)
        ml = 0
    else:
        mk = _arr(masks).reshape(-1)
        ml = mk.size // (D * num) if num else 0
        if mk.size == 0:
            mk = np.zeros(1, dtype=np.uint64)
    out = np.zeros(max(N * num * D, 1), dtype=np.uint64)
    if lib().oracle_interleaved_encode(D, _p64(m) if m.size else _p64(np.zeros(1, np.uint64)),
                                       num, L, _p64(mk), ml, N, _p64(out)):
        raise ValueError("bad interleaved_encode arguments")
    return out[: N * num * D].reshape(N, num * D)


# ---------------------------------------------------------------- hashes
ENGINE_COPY, ENGINE_SHA2, ENGINE_KECCAK, ENGINE_SHA3, ENGINE_BLAKE3 = 0, 1, 2, 3, 4


def _h(fn, data: bytes) -> bytes:
    buf = np.frombuffer(bytes(data) + b"\0", dtype=np.uint8).copy()
    out = np.zeros(32, dtype=np.uint8)
    fn(_p8(buf), len(data), _p8(out))
    return out.tobytes()


def blake3(data: bytes) -> bytes:
    return _h(lib().oracle_blake3, data)


def sha3_256(data: bytes) -> bytes:
    return _h(lib().oracle_sha3_256, data)


def keccak256(data: bytes) -> bytes:
    return _h(lib().oracle_keccak256, data)


def sha256(data: bytes) -> bytes:
    return _h(lib().oracle_sha256, data)


def hash_many(engine: int, size: int, data: bytes, count: int) -> bytes:
    buf = np.frombuffer(bytes(data) + b"\0", dtype=np.uint8).copy()
    out = np.zeros(32 * max(count, 1), dtype=np.uint8)
    if lib().oracle_hash_many(engine, size, _p8(buf), count, _p8(out)):
        raise ValueError("unsupported size for engine")
    return out[: 32 * count].tobytes()


def layers_for_size(n: int) -> int:
    return lib().oracle_layers_for_size(n)


def merkle_commit(engine: int, leaves: bytes) -> bytes:
    n = len(leaves) // 32
    layers = layers_for_size(n)
    num_nodes = (1 << (layers + 1)) - 1
    buf = np.frombuffer(bytes(leaves) + b"\0", dtype=np.uint8).copy()
    out = np.zeros(32 * num_nodes, dtype=np.uint8)
    if lib().oracle_merkle_commit(engine, _p8(buf), n, _p8(out)):
        raise ValueError("merkle commit failed")
    return out.tobytes()


def merkle_open(nodes: bytes, num_leaves: int, indices) -> bytes:
    idx = np.ascontiguousarray(np.asarray(indices, dtype=np.uintp))
    layers = layers_for_size(num_leaves)
    buf = np.frombuffer(nodes if isinstance(nodes, (bytes, bytearray)) else bytes(nodes), dtype=np.uint8)
    out = np.zeros(32 * max(1, len(idx) * max(layers, 1)), dtype=np.uint8)
    nh = lib().oracle_merkle_open(_p8(buf), num_leaves, idx.ctypes.data_as(szp), idx.size, _p8(out))
    return out[: 32 * nh].tobytes()


def hash_rows(engine: int, D: int, matrix, rows: int, cols: int) -> bytes:
    m = _arr(matrix).reshape(-1)
    if m.size == 0:
        m = np.zeros(1, dtype=np.uint64)
    out = np.zeros(32 * max(rows, 1), dtype=np.uint8)
    if lib().oracle_hash_rows(engine, D, _p64(m), rows, cols, _p8(out)):
        raise ValueError("hash_rows failed")
    return out[: 32 * rows].tobytes()


# ---------------------------------------------------------------- sumcheck
def sumcheck_poly(D: int, a, b):
    a = _arr(a).reshape(-1)
    b = _arr(b).reshape(-1)
    c0 = np.zeros(D, dtype=np.uint64)
    c2 = np.zeros(D, dtype=np.uint64)
    z = np.zeros(1, dtype=np.uint64)
    lib().oracle_sumcheck_poly(D, _p64(a) if a.size else _p64(z), a.size // D,
                               _p64(b) if b.size else _p64(z), b.size // D, _p64(c0), _p64(c2))
    return c0, c2


def fold(D: int, v, w) -> np.ndarray:
    v = _arr(v).reshape(-1).copy()
    w = _arr(w).reshape(-1)
    if v.size == 0:
        return v
    n = lib().oracle_fold(D, _p64(v), v.size // D, _p64(w))
    return v[: n * D].copy()


def eval_eq(D: int, acc, point, scalar) -> np.ndarray:
    acc = _arr(acc).reshape(-1).copy()
    pt = _arr(point).reshape(-1)
    k = pt.size // D
    if pt.size == 0:
        pt = np.zeros(1, dtype=np.uint64)
    s = _arr(scalar).reshape(-1)
    lib().oracle_eval_eq(D, _p64(acc), _p64(pt), k, _p64(s))
    return acc


def multilinear_extend(D: int, evals, point) -> np.ndarray:
    e = _arr(evals).reshape(-1)
    pt = _arr(point).reshape(-1)
    k = pt.size // D
    if pt.size == 0:
        pt = np.zeros(1, dtype=np.uint64)
    out = np.zeros(D, dtype=np.uint64)
    lib().oracle_multilinear_extend(D, _p64(e), k, _p64(pt), _p64(out))
    return out


def geometric_accumulate(D: int, acc, scalars, points) -> np.ndarray:
    acc = _arr(acc).reshape(-1).copy()
    s = _arr(scalars).reshape(-1)
    x = _arr(points).reshape(-1)
    m = s.size // D
    z = np.zeros(1, dtype=np.uint64)
    lib().oracle_geometric_accumulate(D, _p64(acc) if acc.size else _p64(z), acc.size // D,
                                      _p64(s) if s.size else z.ctypes.data_as(u64p),
                                      _p64(x) if x.size else z.ctypes.data_as(u64p), m)
    return acc


def univariate_evaluate(Ds: int, D: int, coeffs, x) -> np.ndarray:
    c = _arr(coeffs).reshape(-1)
    n = c.size // Ds
    if c.size == 0:
        c = np.zeros(1, dtype=np.uint64)
    xx = _arr(x).reshape(-1)
    out = np.zeros(D, dtype=np.uint64)
    lib().oracle_univariate_evaluate(Ds, D, _p64(c), n, _p64(xx), _p64(out))
    return out


def lift(D: int, src) -> np.ndarray:
    s = _arr(src).reshape(-1)
    out = np.zeros(s.size * D, dtype=np.uint64)
    if s.size:
        lib().oracle_lift(D, _p64(s), s.size, _p64(out))
    return out


def mixed_scalar_mul_add(D: int, acc, w, v) -> np.ndarray:
    acc = _arr(acc).reshape(-1).copy()
    v = _arr(v).reshape(-1)
    ww = _arr(w).reshape(-1)
    if v.size:
        lib().oracle_mixed_scalar_mul_add(D, _p64(acc), _p64(ww), _p64(v), v.size)
    return acc


def dot(D: int, a, b) -> np.ndarray:
    a = _arr(a).reshape(-1)
    b = _arr(b).reshape(-1)
    out = np.zeros(D, dtype=np.uint64)
    z = np.zeros(1, dtype=np.uint64)
    lib().oracle_dot(D, _p64(a) if a.size else _p64(z), _p64(b) if b.size else _p64(z),
                     a.size // D, _p64(out))
    return out


# ---------------------------------------------------------------- PoW / indices
def pow_threshold(bits: float) -> int:
    """proof_of_work::threshold, src/protocols/proof_of_work.rs:29-39 (libm exp2, as the
    reference's f64::exp2)."""
    assert 0.0 <= bits <= 60.0
    return int(lib().oracle_pow_threshold(float(bits)))


def pow_search(engine: int, challenge: bytes, threshold: int, start: int = 0, limit: int = 1 << 40):
    ch = np.frombuffer(bytes(challenge), dtype=np.uint8).copy()
    out = ctypes.c_uint64()
    found = lib().oracle_pow_search(engine, _p8(ch), threshold, start, limit, ctypes.byref(out))
    return out.value if found else None


def bytes_to_indices(entropy: bytes, count: int, num_leaves: int) -> list:
    buf = np.frombuffer(bytes(entropy) + b"\0", dtype=np.uint8).copy()
    out = np.zeros(max(count, 1), dtype=np.uintp)
    lib().oracle_bytes_to_indices(_p8(buf), count, num_leaves, out.ctypes.data_as(szp))
    return [int(x) for x in out[:count]]
