"""whir_amd -- Python binding of libwhir_amd (MI355X WHIR prover hot path).

A thin ctypes layer over the C ABI declared in include/whir_amd.h.  Names and
argument meanings follow the reference crate's Rust interfaces:

  ReedSolomon        next_order, generator, evaluation_points, interleaved_rs_encode
                     (src/algebra/ntt/mod.rs:65-134)
  HashEngine         Engine(...).hash_many / engine_id / supports_size (src/hash/mod.rs:58-103)
  merkle / matrix    merkle_commit, merkle_open, matrix_commit
                     (src/protocols/merkle_tree.rs:82-186, matrix_commit.rs:211-232)
  sumcheck           compute_sumcheck_polynomial, fold, fold_and_compute_polynomial
                     (src/algebra/sumcheck.rs:12-111)
  covectors          eval_eq, geometric_accumulate, mixed_univariate_evaluate, lift,
                     mixed_scalar_mul_add, dot (src/algebra/multilinear.rs, src/algebra/mod.rs)
  proof of work      pow_threshold, pow_grind (src/protocols/proof_of_work.rs:29-166)

Field elements are numpy uint64 arrays of canonical values, `degree` words per
element.  Every call runs on the GPU through libwhir_amd.so; there is no CPU
path.  Loading fails loudly when the library has not been built.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libwhir_amd.so")

P = 0xFFFFFFFF00000001
U64_MAX = 2**64 - 1

HASH_COPY, HASH_SHA2, HASH_KECCAK, HASH_SHA3, HASH_BLAKE3 = 0, 1, 2, 3, 4
FORM_MLE, FORM_COVECTOR, FORM_UNIVARIATE = 0, 1, 2  # linear form kinds of whir_prove_forms_dev
ENGINE_NAMES = {HASH_COPY: "copy", HASH_SHA2: "sha2", HASH_KECCAK: "keccak", HASH_SHA3: "sha3",
                HASH_BLAKE3: "Blake3"}


class WhirError(RuntimeError):
    """A nonzero status from libwhir_amd (the reference would panic here)."""


_lib = None

u64p = ctypes.POINTER(ctypes.c_uint64)
u8p = ctypes.POINTER(ctypes.c_uint8)
szp = ctypes.POINTER(ctypes.c_size_t)
vp = ctypes.c_void_p
sz = ctypes.c_size_t
ci = ctypes.c_int
u64 = ctypes.c_uint64

# Symbol table: name -> argtypes.  Also the list tests use to check exports.
SIGNATURES = {
    "whir_abi_version": [],
    "whir_device_count": [ctypes.POINTER(ci)],
    "whir_set_device": [ci],
    "whir_dev_alloc": [sz, ctypes.POINTER(vp)],
    "whir_dev_free": [vp],
    "whir_host_alloc": [sz, ctypes.POINTER(vp)],
    "whir_host_free": [vp],
    "whir_commit_host": [vp, vp, vp, vp, ctypes.POINTER(vp)],
    "whir_memcpy_h2d": [vp, vp, sz],
    "whir_memcpy_d2h": [vp, vp, sz],
    "whir_memcpy_d2d": [vp, vp, sz],
    "whir_synchronize": [],
    "whir_profile_enable": [ci],
    "whir_profile_reset": [],
    "whir_profile_report": [ctypes.c_char_p, sz, szp],
    "whir_rs_next_order": [u64, u64p],
    "whir_rs_generator": [u64, u64p],
    "whir_rs_evaluation_points": [u64, u64, u64p, sz, u64p],
    "whir_rs_interleaved_encode": [ci, ctypes.POINTER(u64p), sz, sz, u64p, sz, sz, u64p],
    "whir_rs_interleaved_encode_dev": [ci, vp, sz, sz, vp, sz, sz, vp, ci],
    "whir_hash_engine_id": [ci, u8p],
    "whir_hash_supports_size": [ci, sz, ctypes.POINTER(ci)],
    "whir_hash_preferred_batch_size": [ci, ctypes.POINTER(sz)],
    "whir_hash_many": [ci, sz, u8p, sz, u8p],
    "whir_hash_many_dev": [ci, sz, vp, sz, vp],
    "whir_merkle_num_nodes": [sz, szp],
    "whir_merkle_commit": [ci, u8p, sz, u8p],
    "whir_merkle_commit_dev": [ci, vp, sz],
    "whir_matrix_commit": [ci, ci, u64p, sz, sz, u8p],
    "whir_hash_rows_cols_dev": [ci, ci, vp, sz, sz, vp],
    "whir_matrix_commit_cols_dev": [ci, ci, vp, sz, sz, vp],
    "whir_merkle_open": [u8p, sz, szp, sz, u8p, szp],
    "whir_sumcheck_polynomial": [ci, u64p, u64p, sz, u64p],
    "whir_sumcheck_polynomial_dev": [ci, vp, vp, sz, u64p],
    "whir_fold": [ci, u64p, sz, u64p, szp],
    "whir_fold_dev": [ci, vp, sz, u64p, szp],
    "whir_fold_and_compute_polynomial": [ci, u64p, u64p, sz, u64p, u64p, szp],
    "whir_fold_and_compute_polynomial_dev": [ci, vp, vp, sz, u64p, u64p, szp],
    "whir_fold_and_compute_polynomial_dev_noalloc": [ci, vp, vp, sz, u64p, u64p, szp],
    "whir_eval_eq": [ci, u64p, u64p, sz, u64p],
    "whir_eval_eq_dev": [ci, vp, u64p, sz, u64p],
    "whir_geometric_accumulate": [ci, u64p, sz, u64p, u64p, sz],
    "whir_geometric_accumulate_dev": [ci, vp, sz, u64p, u64p, sz],
    "whir_mixed_univariate_evaluate": [ci, ci, u64p, sz, u64p, u64p],
    "whir_mixed_univariate_evaluate_dev": [ci, ci, vp, sz, u64p, u64p],
    "whir_lift": [ci, u64p, sz, u64p],
    "whir_lift_dev": [ci, vp, sz, vp],
    "whir_mixed_scalar_mul_add": [ci, u64p, u64p, u64p, sz],
    "whir_mixed_scalar_mul_add_dev": [ci, vp, u64p, vp, sz],
    "whir_dot": [ci, u64p, u64p, sz, u64p],
    "whir_pow_threshold": [ctypes.c_double, u64p],
    "whir_pow_grind": [ci, u8p, u64, u64p, ctypes.POINTER(ci)],
    "whir_pow_grind_bounded": [ci, u8p, u64, u64, u64p, ctypes.POINTER(ci)],
    # prover (whir_amd/prover.py)
    "whir_config_new": [vp, sz, ctypes.POINTER(vp)],
    "whir_config_disable_pow": [vp],
    "whir_config_describe": [vp, ctypes.c_char_p, sz, szp],
    "whir_config_free": [vp],
    "whir_transcript_new": [u8p, u8p, u8p, sz, ctypes.POINTER(vp)],
    "whir_transcript_from_vtable": [vp, ctypes.POINTER(vp)],
    "whir_transcript_proof": [vp, u8p, sz, szp, u8p, sz, szp],
    "whir_transcript_free": [vp],
    "whir_commit_dev": [vp, vp, vp, ctypes.POINTER(vp)],
    "whir_prove_dev": [vp, vp, vp, vp, u64p, sz, u64p, u64p],
    "whir_ntt_autotune": [ci],
    "whir_ntt_set_schedule": [ci, sz, sz, sz, ci, ci, ci, ci, ci],
    "whir_ntt_clear_schedules": [],
    "whir_ntt_schedules": [ctypes.c_char_p, sz, szp],
    "whir_prove_forms_dev": [vp, vp, ctypes.POINTER(vp), sz, ctypes.POINTER(vp), sz, ctypes.POINTER(ci), sz, u64p,
                             ctypes.POINTER(vp), u64p, u64p],
    "whir_commit_batch_dev": [vp, vp, ctypes.POINTER(vp), sz, ctypes.POINTER(vp)],
    "whir_prove_batch_dev": [vp, vp, ctypes.POINTER(vp), sz, ctypes.POINTER(vp), sz, u64p, sz, u64p, u64p],
    "whir_witness_free": [vp],
    "whir_config_memory": [vp, szp],
    "whir_executor_new": [ci, sz, ctypes.POINTER(vp)],
    "whir_executor_free": [vp],
    "whir_executor_submit": [vp, ci, vp, vp, vp, u64p, sz, u64p, ctypes.POINTER(ctypes.c_uint64)],
    "whir_executor_submit_batch": [vp, vp, vp, ctypes.POINTER(vp), sz, u64p, sz, u64p,
                                   ctypes.POINTER(ctypes.c_uint64)],
    "whir_executor_wait": [vp, ctypes.c_uint64, ctypes.POINTER(ctypes.c_double), u64p],
    "whir_executor_submit_prove": [vp, vp, vp, vp, vp, u64p, sz, u64p, ctypes.POINTER(ctypes.c_uint64)],
    "whir_commit_many_dev": [ctypes.POINTER(vp), ctypes.POINTER(vp), ctypes.POINTER(vp), sz, ctypes.POINTER(vp)],
    "whir_executor_clock": [vp, ctypes.POINTER(ctypes.c_double)],
    "whir_mle_evaluate_dev": [ci, vp, sz, u64p, u64p],
}
VOID_FUNCS = {"whir_config_free", "whir_transcript_free", "whir_witness_free", "whir_executor_free"}


def lib():
    """The loaded libwhir_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise WhirError(f"{LIB_PATH} is missing: build it with `python -m whir_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = None if name in VOID_FUNCS else ci
        L.whir_last_error.argtypes = []
        L.whir_last_error.restype = ctypes.c_char_p
        _lib = L
    return _lib


def _call(name, *args):
    st = getattr(lib(), name)(*args)
    if st != 0:
        raise WhirError(f"{name}: {lib().whir_last_error().decode()}")


def _a64(x, n_words=None) -> np.ndarray:
    a = np.ascontiguousarray(np.asarray(x, dtype=np.uint64)).reshape(-1)
    if a.size == 0:
        a = np.zeros(1, dtype=np.uint64)
    return a


def _p64(a: np.ndarray):
    return a.ctypes.data_as(u64p)


def _p8(a: np.ndarray):
    return a.ctypes.data_as(u8p)


def _bytes_arr(b: bytes) -> np.ndarray:
    return np.frombuffer(bytes(b) + b"\0", dtype=np.uint8).copy()


def device_count() -> int:
    n = ci()
    _call("whir_device_count", ctypes.byref(n))
    return n.value


# ------------------------------------------------------------------ ReedSolomon
def next_order(size: int):
    out = u64()
    if lib().whir_rs_next_order(size, ctypes.byref(out)) != 0:
        return None
    return out.value


def generator(codeword_length: int) -> int:
    out = u64()
    _call("whir_rs_generator", codeword_length, ctypes.byref(out))
    return out.value


def evaluation_points(masked_message_length: int, codeword_length: int, indices) -> np.ndarray:
    idx = np.ascontiguousarray(np.asarray(indices, dtype=np.uint64))
    out = np.zeros(max(idx.size, 1), dtype=np.uint64)
    _call("whir_rs_evaluation_points", masked_message_length, codeword_length,
          _p64(idx if idx.size else out), idx.size, _p64(out))
    return out[: idx.size]


def interleaved_rs_encode(messages, masks, codeword_length: int, degree: int = 1) -> np.ndarray:
    """messages: list of arrays (message_length*degree words each); masks: flat array
    (num_messages * mask_length * degree words) or None.  Returns the row-major
    codeword_length x num_messages matrix as a (codeword_length, num_messages*degree) array."""
    msgs = [np.ascontiguousarray(np.asarray(m, dtype=np.uint64)).reshape(-1) for m in messages]
    M = len(msgs)
    L = (msgs[0].size // degree) if M else 0
    if any(m.size != L * degree for m in msgs):
        raise WhirError("interleaved_encode: messages of unequal length")
    mk = _a64(masks if masks is not None else [])
    ml = (0 if masks is None else np.asarray(masks).size // max(M * degree, 1))
    keep = [m if m.size else np.zeros(1, dtype=np.uint64) for m in msgs]
    ptrs = (u64p * max(M, 1))(*[_p64(m) for m in keep])
    out = np.zeros(max(codeword_length * M * degree, 1), dtype=np.uint64)
    _call("whir_rs_interleaved_encode", degree, ptrs, M, L, _p64(mk), ml, codeword_length, _p64(out))
    return out[: codeword_length * M * degree].reshape(codeword_length, M * degree)


def ntt_autotune(enable: bool = True) -> None:
    """Tune the schedule of each new encode shape on its first use (see
    whir_ntt_autotune); outputs do not depend on the schedule."""
    _call("whir_ntt_autotune", int(bool(enable)))


def ntt_set_schedule(degree: int, log_n: int, log_c: int, num_messages: int, log_c1: int, variant1: int,
                     tile1: int, variant2: int, tile2: int) -> None:
    _call("whir_ntt_set_schedule", degree, log_n, log_c, num_messages, log_c1, variant1, tile1, variant2, tile2)


def ntt_clear_schedules() -> None:
    _call("whir_ntt_clear_schedules")


def ntt_schedules() -> list:
    """The schedule table: one dict per shape (tuned or pinned)."""
    import json
    n = sz()
    _call("whir_ntt_schedules", None, 0, ctypes.byref(n))
    buf = ctypes.create_string_buffer(n.value + 1)
    _call("whir_ntt_schedules", buf, n.value + 1, ctypes.byref(n))
    return json.loads(buf.value.decode())


# ------------------------------------------------------------------ HashEngine
class Engine:
    """HashEngine (src/hash/mod.rs:58-86) backed by the GPU."""

    def __init__(self, engine: int):
        self.engine = engine

    def name(self) -> str:
        return ENGINE_NAMES[self.engine]

    def engine_id(self) -> bytes:
        out = np.zeros(32, dtype=np.uint8)
        _call("whir_hash_engine_id", self.engine, _p8(out))
        return out.tobytes()

    def supports_size(self, size: int) -> bool:
        r = ci()
        _call("whir_hash_supports_size", self.engine, size, ctypes.byref(r))
        return bool(r.value)

    def preferred_batch_size(self) -> int:
        r = sz()
        _call("whir_hash_preferred_batch_size", self.engine, ctypes.byref(r))
        return r.value

    def hash_many(self, size: int, data: bytes) -> bytes:
        if size == 0:
            raise WhirError("hash_many: use count explicitly for size 0 via hash_many_count")
        count = len(data) // size
        return self.hash_many_count(size, data, count)

    def hash_many_count(self, size: int, data: bytes, count: int) -> bytes:
        if len(data) != size * count:
            raise WhirError(f"Input length should be size * output.len() = {size} * {count}")
        buf = _bytes_arr(data)
        out = np.zeros(32 * max(count, 1), dtype=np.uint8)
        _call("whir_hash_many", self.engine, size, _p8(buf), count, _p8(out))
        return out[: 32 * count].tobytes()


Blake3 = lambda: Engine(HASH_BLAKE3)  # noqa: E731
Sha3 = lambda: Engine(HASH_SHA3)  # noqa: E731
Sha2 = lambda: Engine(HASH_SHA2)  # noqa: E731
Keccak = lambda: Engine(HASH_KECCAK)  # noqa: E731


# ------------------------------------------------------------------ Merkle / matrix
def merkle_num_nodes(num_leaves: int) -> int:
    out = sz()
    _call("whir_merkle_num_nodes", num_leaves, ctypes.byref(out))
    return out.value


def merkle_commit(engine: int, leaves: bytes) -> bytes:
    n = len(leaves) // 32
    nn = merkle_num_nodes(n)
    buf = _bytes_arr(leaves)
    out = np.zeros(32 * nn, dtype=np.uint8)
    _call("whir_merkle_commit", engine, _p8(buf), n, _p8(out))
    return out.tobytes()


def matrix_commit(engine: int, degree: int, matrix, rows: int, cols: int) -> bytes:
    m = _a64(matrix)
    nn = merkle_num_nodes(rows)
    out = np.zeros(32 * nn, dtype=np.uint8)
    _call("whir_matrix_commit", engine, degree, _p64(m), rows, cols, _p8(out))
    return out.tobytes()


def merkle_open(nodes: bytes, num_leaves: int, indices) -> bytes:
    idx = np.ascontiguousarray(np.asarray(indices, dtype=np.uintp))
    if idx.size == 0:
        idx = np.zeros(1, dtype=np.uintp)
        count = 0
    else:
        count = idx.size
    layers = max(1, (num_leaves - 1).bit_length())
    hints = np.zeros(32 * max(1, count * layers), dtype=np.uint8)
    nh = sz()
    buf = _bytes_arr(nodes)
    _call("whir_merkle_open", _p8(buf), num_leaves, idx.ctypes.data_as(szp), count, _p8(hints), ctypes.byref(nh))
    return hints[: 32 * nh.value].tobytes()


# ------------------------------------------------------------------ sumcheck
def compute_sumcheck_polynomial(degree: int, a, b):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.uint64)).reshape(-1)
    b = np.ascontiguousarray(np.asarray(b, dtype=np.uint64)).reshape(-1)
    n = min(a.size, b.size) // degree
    out = np.zeros(2 * degree, dtype=np.uint64)
    _call("whir_sumcheck_polynomial", degree, _p64(_a64(a)), _p64(_a64(b)), n, _p64(out))
    return out[:degree].copy(), out[degree:].copy()


def fold(degree: int, values, weight) -> np.ndarray:
    v = np.ascontiguousarray(np.asarray(values, dtype=np.uint64)).reshape(-1).copy()
    n = v.size // degree
    if n == 0:
        return v
    w = _a64(weight)
    nl = sz()
    _call("whir_fold", degree, _p64(v), n, _p64(w), ctypes.byref(nl))
    return v[: nl.value * degree].copy()


def fold_and_compute_polynomial(degree: int, a, b, weight):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.uint64)).reshape(-1).copy()
    b = np.ascontiguousarray(np.asarray(b, dtype=np.uint64)).reshape(-1).copy()
    n = a.size // degree
    w = _a64(weight)
    out = np.zeros(2 * degree, dtype=np.uint64)
    nl = sz()
    _call("whir_fold_and_compute_polynomial", degree, _p64(_a64(a) if n == 0 else a),
          _p64(_a64(b) if n == 0 else b), n, _p64(w), _p64(out), ctypes.byref(nl))
    m = nl.value
    return a[: m * degree].copy(), b[: m * degree].copy(), out[:degree].copy(), out[degree:].copy()


def eval_eq(degree: int, acc, point, scalar) -> np.ndarray:
    acc = np.ascontiguousarray(np.asarray(acc, dtype=np.uint64)).reshape(-1).copy()
    pt = _a64(point)
    k = np.asarray(point).size // degree
    _call("whir_eval_eq", degree, _p64(acc), _p64(pt), k, _p64(_a64(scalar)))
    return acc


def geometric_accumulate(degree: int, acc, scalars, points) -> np.ndarray:
    acc = np.ascontiguousarray(np.asarray(acc, dtype=np.uint64)).reshape(-1).copy()
    n = acc.size // degree
    m = np.asarray(scalars).size // degree
    if n == 0:
        return acc
    _call("whir_geometric_accumulate", degree, _p64(acc), n, _p64(_a64(scalars)), _p64(_a64(points)), m)
    return acc


def mixed_univariate_evaluate(src_degree: int, degree: int, coeffs, point) -> np.ndarray:
    c = np.asarray(coeffs, dtype=np.uint64).reshape(-1)
    n = c.size // src_degree
    out = np.zeros(degree, dtype=np.uint64)
    _call("whir_mixed_univariate_evaluate", src_degree, degree, _p64(_a64(c)), n, _p64(_a64(point)), _p64(out))
    return out


def lift(degree: int, src) -> np.ndarray:
    s = np.asarray(src, dtype=np.uint64).reshape(-1)
    out = np.zeros(max(s.size * degree, 1), dtype=np.uint64)
    _call("whir_lift", degree, _p64(_a64(s)), s.size, _p64(out))
    return out[: s.size * degree]


def mixed_scalar_mul_add(degree: int, acc, weight, v) -> np.ndarray:
    acc = np.ascontiguousarray(np.asarray(acc, dtype=np.uint64)).reshape(-1).copy()
    vv = np.asarray(v, dtype=np.uint64).reshape(-1)
    _call("whir_mixed_scalar_mul_add", degree, _p64(_a64(acc) if acc.size == 0 else acc), _p64(_a64(weight)),
          _p64(_a64(vv)), vv.size)
    return acc


def dot(degree: int, a, b) -> np.ndarray:
    a = np.asarray(a, dtype=np.uint64).reshape(-1)
    b = np.asarray(b, dtype=np.uint64).reshape(-1)
    out = np.zeros(degree, dtype=np.uint64)
    _call("whir_dot", degree, _p64(_a64(a)), _p64(_a64(b)), a.size // degree, _p64(out))
    return out


# ------------------------------------------------------------------ proof of work
def pow_threshold(bits: float) -> int:
    out = u64()
    _call("whir_pow_threshold", float(bits), ctypes.byref(out))
    return out.value


def pow_grind(engine: int, challenge: bytes, threshold: int):
    """Minimal nonce, or None for the no-op threshold u64::MAX (no transcript)."""
    ch = _bytes_arr(challenge)
    n = u64()
    performed = ci()
    _call("whir_pow_grind", engine, _p8(ch), threshold, ctypes.byref(n), ctypes.byref(performed))
    return n.value if performed.value else None


def pow_grind_bounded(engine: int, challenge: bytes, threshold: int, max_nonces: int):
    ch = _bytes_arr(challenge)
    n = u64()
    found = ci()
    _call("whir_pow_grind_bounded", engine, _p8(ch), threshold, max_nonces, ctypes.byref(n), ctypes.byref(found))
    return n.value if found.value else None


# ------------------------------------------------------------------ device buffers
class DeviceBuffer:
    """A device allocation from whir_dev_alloc, freed on close()/garbage collection."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = vp()
        _call("whir_dev_alloc", self.nbytes, ctypes.byref(p))
        self.ptr = p.value

    @classmethod
    def from_array(cls, a) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(max(a.nbytes, 1))
        if a.nbytes:
            _call("whir_memcpy_h2d", b.ptr, a.ctypes.data, a.nbytes)
        return b

    def to_array(self, dtype=np.uint64, count=None) -> np.ndarray:
        itemsize = np.dtype(dtype).itemsize
        count = self.nbytes // itemsize if count is None else count
        out = np.zeros(max(count, 1), dtype=dtype)
        if count:
            _call("whir_memcpy_d2h", out.ctypes.data, self.ptr, count * itemsize)
        return out[:count]

    def close(self):
        if self.ptr:
            lib().whir_dev_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PinnedArray:
    """A numpy uint64 array in pinned host memory (whir_host_alloc)."""

    def __init__(self, n: int):
        self.n = int(n)
        p = vp()
        _call("whir_host_alloc", max(self.n, 1) * 8, ctypes.byref(p))
        self.ptr = p.value
        self.array = np.ctypeslib.as_array(ctypes.cast(self.ptr, u64p), shape=(max(self.n, 1),))[: self.n]

    def close(self):
        if self.ptr:
            self.array = None
            lib().whir_host_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def memcpy_h2d(buf: DeviceBuffer, a) -> None:
    """Copy a host array into the start of buf (synchronous)."""
    a = np.ascontiguousarray(a)
    if a.nbytes > buf.nbytes:
        raise ValueError("memcpy_h2d: array larger than the buffer")
    if a.nbytes:
        _call("whir_memcpy_h2d", buf.ptr, a.ctypes.data, a.nbytes)


def synchronize():
    _call("whir_synchronize")


def profile_enable(on: bool = True):
    _call("whir_profile_enable", 1 if on else 0)


def profile_reset():
    _call("whir_profile_reset")


def profile_report() -> dict:
    """{kernel family: {"launches", "total_ms", "bytes"}} from HIP events on the launching stream."""
    import json
    n = sz()
    _call("whir_profile_report", None, 0, ctypes.byref(n))
    buf = ctypes.create_string_buffer(n.value + 1)
    _call("whir_profile_report", buf, n.value + 1, ctypes.byref(n))
    return json.loads(buf.value.decode())


def rs_encode_dev(degree: int, d_msgs: DeviceBuffer, num_messages: int, message_length: int,
                  codeword_length: int, d_out: DeviceBuffer, row_major: bool = False):
    _call("whir_rs_interleaved_encode_dev", degree, d_msgs.ptr, num_messages, message_length, None, 0,
          codeword_length, d_out.ptr, 1 if row_major else 0)


def hash_rows_cols_dev(engine: int, degree: int, d_cw: DeviceBuffer, rows: int, cols: int, d_leaves: DeviceBuffer):
    _call("whir_hash_rows_cols_dev", engine, degree, d_cw.ptr, rows, cols, d_leaves.ptr)


def matrix_commit_cols_dev(engine: int, degree: int, d_cw: DeviceBuffer, rows: int, cols: int,
                           d_nodes: DeviceBuffer):
    """Leaves and every Merkle layer of a column-major device codeword (the commit
    irs_commit.rs:292-350 performs) into d_nodes (merkle_num_nodes(rows) hashes)."""
    _call("whir_matrix_commit_cols_dev", engine, degree, d_cw.ptr, rows, cols, d_nodes.ptr)


def merkle_commit_dev(engine: int, d_nodes: DeviceBuffer, num_leaves: int):
    _call("whir_merkle_commit_dev", engine, d_nodes.ptr, num_leaves)
