"""WHIR prover binding: whir::Config (src/protocols/whir/config.rs:16-152), commit
(src/protocols/whir/mod.rs:79-92) and prove (src/protocols/whir/prover.rs:50-316)
running on the GPU through libwhir_amd.so."""
from __future__ import annotations

import ctypes
import json
from dataclasses import dataclass

import numpy as np

from . import (DeviceBuffer, HASH_BLAKE3, _call, _p64, _p8, ci, lib, sz, u64p, vp)


class _Params(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in (
        "degree", "security_level", "pow_bits", "initial_folding_factor", "folding_factor",
        "unique_decoding", "starting_log_inv_rate", "batch_size", "hash_engine")]


@dataclass
class ProtocolParameters:
    """ProtocolParameters (src/parameters.rs:8-24) plus the field degree
    (1 = Goldilocks1, 2 = Goldilocks2, 3 = Goldilocks3)."""
    degree: int = 2
    security_level: int = 100
    pow_bits: int = 20
    initial_folding_factor: int = 4
    folding_factor: int = 4
    unique_decoding: bool = False
    starting_log_inv_rate: int = 1
    batch_size: int = 1
    hash_engine: int = HASH_BLAKE3

    def _c(self):
        return _Params(self.degree, self.security_level, self.pow_bits, self.initial_folding_factor,
                       self.folding_factor, int(self.unique_decoding), self.starting_log_inv_rate,
                       self.batch_size, self.hash_engine)


class Config:
    """whir::Config::new(size, params)."""

    def __init__(self, size: int, params: ProtocolParameters):
        self.params = params
        self.size = size
        self.degree = params.degree
        self.num_variables = size.bit_length() - 1
        h = vp()
        p = params._c()
        _call("whir_config_new", ctypes.byref(p), size, ctypes.byref(h))
        self.h = h.value

    def disable_pow(self):
        _call("whir_config_disable_pow", self.h)
        return self

    def describe(self) -> dict:
        n = sz()
        _call("whir_config_describe", self.h, None, 0, ctypes.byref(n))
        buf = ctypes.create_string_buffer(n.value + 1)
        _call("whir_config_describe", self.h, buf, n.value + 1, ctypes.byref(n))
        return json.loads(buf.value.decode())

    def __del__(self):
        try:
            if self.h:
                lib().whir_config_free(self.h)
                self.h = None
        except Exception:
            pass


class Transcript:
    """The built-in Keccak duplex transcript (DomainSeparator protocol/session ids)."""

    def __init__(self, protocol_id: bytes, session_id: bytes, instance: bytes = b""):
        assert len(protocol_id) == 64 and len(session_id) == 32
        self._keep = [np.frombuffer(bytes(x) + b"\0", dtype=np.uint8).copy()
                      for x in (protocol_id, session_id, instance)]
        h = vp()
        _call("whir_transcript_new", _p8(self._keep[0]), _p8(self._keep[1]), _p8(self._keep[2]), len(instance),
              ctypes.byref(h))
        self.h = h.value

    def proof(self):
        """(narg_string, hints) of Proof (src/transcript/mod.rs:63-70)."""
        na, nh = sz(), sz()
        _call("whir_transcript_proof", self.h, None, 0, ctypes.byref(na), None, 0, ctypes.byref(nh))
        a = np.zeros(max(na.value, 1), dtype=np.uint8)
        b = np.zeros(max(nh.value, 1), dtype=np.uint8)
        _call("whir_transcript_proof", self.h, _p8(a), a.size, ctypes.byref(na), _p8(b), b.size, ctypes.byref(nh))
        return a[: na.value].tobytes(), b[: nh.value].tobytes()

    def __del__(self):
        try:
            if self.h:
                lib().whir_transcript_free(self.h)
                self.h = None
        except Exception:
            pass


class _Vtable(ctypes.Structure):
    _F = ctypes.CFUNCTYPE
    _u8p = ctypes.POINTER(ctypes.c_uint8)
    _fields_ = [("ctx", vp),
                ("prover_field", _F(ci, vp, ci, u64p, sz)),
                ("verifier_field", _F(ci, vp, ci, u64p)),
                ("verifier_bytes", _F(ci, vp, _u8p, sz)),
                ("prover_hash", _F(ci, vp, _u8p)),
                ("prover_u64", _F(ci, vp, ctypes.c_uint64)),
                ("hint_hash", _F(ci, vp, _u8p)),
                ("hint_field_vec", _F(ci, vp, ci, u64p, sz))]


class CallbackTranscript:
    """A transcript whose typed operations (ProverState, src/transcript/mod.rs:183-267)
    run in Python, through whir_transcript_from_vtable: the GPU prover calls
    impl.prover_field(D, words), impl.verifier_field(D) -> D words,
    impl.verifier_bytes(n) -> n bytes, impl.prover_hash(h32), impl.prover_u64(v),
    impl.hint_hash(h32) and impl.hint_field_vec(D, words) in transcript order.  An
    exception in a callback fails the prover call (whir_last_error names the
    operation); the first one is kept in `error`."""

    def __init__(self, impl):
        self.impl = impl
        self.error = None

        def guard(fn):
            def run(*a):
                try:
                    fn(*a)
                    return 0
                except Exception as e:  # reported through the status, kept for the caller
                    if self.error is None:
                        self.error = e
                    return 1
            return run

        def prover_field(_, D, p, count):
            impl.prover_field(D, np.ctypeslib.as_array(p, shape=(D * count,)).copy() if count else
                              np.zeros(0, np.uint64))

        def verifier_field(_, D, out):
            v = np.asarray(impl.verifier_field(D), dtype=np.uint64).reshape(-1)
            assert v.size == D
            np.ctypeslib.as_array(out, shape=(D,))[:] = v

        def verifier_bytes(_, out, n):
            b = bytes(impl.verifier_bytes(n))
            assert len(b) == n
            if n:
                ctypes.memmove(out, b, n)

        def prover_hash(_, h):
            impl.prover_hash(ctypes.string_at(h, 32))

        def prover_u64(_, v):
            impl.prover_u64(int(v))

        def hint_hash(_, h):
            impl.hint_hash(ctypes.string_at(h, 32))

        def hint_field_vec(_, D, p, count):
            impl.hint_field_vec(D, np.ctypeslib.as_array(p, shape=(D * count,)).copy() if count else
                                np.zeros(0, np.uint64))

        fns = [prover_field, verifier_field, verifier_bytes, prover_hash, prover_u64, hint_hash, hint_field_vec]
        self._vt = _Vtable(None, *[ftype(guard(fn)) for fn, (_, ftype) in zip(fns, _Vtable._fields_[1:])])
        h = vp()
        _call("whir_transcript_from_vtable", ctypes.byref(self._vt), ctypes.byref(h))
        self.h = h.value

    def __del__(self):
        try:
            if self.h:
                lib().whir_transcript_free(self.h)
                self.h = None
        except Exception:
            pass


class Witness:
    def __init__(self, h):
        self.h = h

    def __del__(self):
        try:
            if self.h:
                lib().whir_witness_free(self.h)
                self.h = None
        except Exception:
            pass


def commit(cfg: Config, t: Transcript, d_vector: DeviceBuffer) -> Witness:
    h = vp()
    _call("whir_commit_dev", cfg.h, t.h, d_vector.ptr, ctypes.byref(h))
    return Witness(h.value)


def commit_host(cfg: Config, t: Transcript, host_vector, d_vector: DeviceBuffer) -> Witness:
    """whir::Config::commit of a vector in host memory (a numpy uint64 array or a
    PinnedArray): uploaded into d_vector a chunk of messages at a time, each chunk's
    encode overlapping the later chunks' copies."""
    from . import PinnedArray
    arr = None if isinstance(host_vector, PinnedArray) else np.ascontiguousarray(host_vector, dtype=np.uint64)
    ptr = host_vector.ptr if arr is None else arr.ctypes.data
    h = vp()
    _call("whir_commit_host", cfg.h, t.h, ptr, d_vector.ptr, ctypes.byref(h))  # returns after the copies
    del arr
    return Witness(h.value)


def commit_many(cfgs, transcripts, d_vectors):
    """whir::Config::commit of many independent instances in shared launches (a HIP
    graph per batch layout); the same witnesses and transcripts as commit() on each."""
    n = len(cfgs)
    assert len(transcripts) == n and len(d_vectors) == n
    c = (vp * max(n, 1))(*[x.h for x in cfgs])
    t = (vp * max(n, 1))(*[x.h for x in transcripts])
    v = (vp * max(n, 1))(*[x.ptr for x in d_vectors])
    out = (vp * max(n, 1))()
    _call("whir_commit_many_dev", c, t, v, n, out)
    return [Witness(out[k]) for k in range(n)]


def prove(cfg: Config, t: Transcript, d_vector: DeviceBuffer, w: Witness, points, evaluations) -> np.ndarray:
    """points: (num_points, num_variables * degree) words; evaluations: (num_points, degree)."""
    D, nv = cfg.degree, cfg.num_variables
    pts = np.ascontiguousarray(np.asarray(points, dtype=np.uint64)).reshape(-1)
    ev = np.ascontiguousarray(np.asarray(evaluations, dtype=np.uint64)).reshape(-1)
    npts = pts.size // max(nv * D, 1) if nv else len(evaluations) if len(evaluations) else 0
    if pts.size == 0:
        pts = np.zeros(1, dtype=np.uint64)
    if ev.size == 0:
        ev = np.zeros(1, dtype=np.uint64)
    out = np.zeros(max(nv * D, 1), dtype=np.uint64)
    _call("whir_prove_dev", cfg.h, t.h, d_vector.ptr, w.h, _p64(pts), npts, _p64(ev), _p64(out))
    return out[: nv * D]


def commit_batch(cfg: Config, t: Transcript, d_vectors) -> Witness:
    """whir::Config::commit of batch_size device vectors in one commitment."""
    ptrs = (vp * max(len(d_vectors), 1))(*[d.ptr for d in d_vectors])
    h = vp()
    _call("whir_commit_batch_dev", cfg.h, t.h, ptrs, len(d_vectors), ctypes.byref(h))
    return Witness(h.value)


def prove_batch(cfg: Config, t: Transcript, d_vectors, witnesses, points, evaluations) -> np.ndarray:
    """whir::Config::prove over several commitments (batch_size vectors each, in
    commitment order).  points: (num_points, num_variables * degree) words;
    evaluations: (num_points, num_vectors, degree), vector j at point k."""
    D, nv = cfg.degree, cfg.num_variables
    pts = np.ascontiguousarray(np.asarray(points, dtype=np.uint64)).reshape(-1)
    ev = np.ascontiguousarray(np.asarray(evaluations, dtype=np.uint64)).reshape(-1)
    npts = pts.size // (nv * D) if nv else len(evaluations)
    assert ev.size == npts * len(d_vectors) * D, "evaluations: num_points x num_vectors values"
    if pts.size == 0:
        pts = np.zeros(1, dtype=np.uint64)
    if ev.size == 0:
        ev = np.zeros(1, dtype=np.uint64)
    vptrs = (vp * max(len(d_vectors), 1))(*[d.ptr for d in d_vectors])
    wptrs = (vp * max(len(witnesses), 1))(*[w.h for w in witnesses])
    out = np.zeros(max(nv * D, 1), dtype=np.uint64)
    _call("whir_prove_batch_dev", cfg.h, t.h, vptrs, len(d_vectors), wptrs, len(witnesses), _p64(pts), npts,
          _p64(ev), _p64(out))
    return out[: nv * D]


class Univariate:
    """UnivariateEvaluation linear form at an extension point x (degree words)."""

    def __init__(self, x):
        self.x = np.ascontiguousarray(np.asarray(x, dtype=np.uint64)).reshape(-1)


def prove_forms(cfg: Config, t: Transcript, d_vectors, witnesses, forms, evaluations) -> np.ndarray:
    """whir::Config::prove with linear forms in order: a DeviceBuffer is a Covector
    (vector_size elements of degree words), a Univariate a UnivariateEvaluation,
    anything else an MLE point (num_variables * degree words).  evaluations:
    (num_forms, num_vectors, degree)."""
    from . import FORM_COVECTOR, FORM_MLE, FORM_UNIVARIATE
    D, nv = cfg.degree, cfg.num_variables
    kinds = [FORM_COVECTOR if isinstance(f, DeviceBuffer) else FORM_UNIVARIATE if isinstance(f, Univariate)
             else FORM_MLE for f in forms]
    pts = []
    for f in forms:
        if isinstance(f, Univariate):
            assert f.x.size == D, "univariate points are degree words"
            pts.append(f.x)
        elif not isinstance(f, DeviceBuffer):
            p = np.ascontiguousarray(np.asarray(f, dtype=np.uint64)).reshape(-1)
            assert p.size == nv * D, "MLE points must be num_variables x degree words"
            pts.append(p)
    flat = np.concatenate(pts) if pts else np.zeros(1, dtype=np.uint64)
    covs = [f for f in forms if isinstance(f, DeviceBuffer)]
    ev = np.ascontiguousarray(np.asarray(evaluations, dtype=np.uint64)).reshape(-1)
    assert ev.size == len(forms) * len(d_vectors) * D, "evaluations: num_forms x num_vectors values"
    if ev.size == 0:
        ev = np.zeros(1, dtype=np.uint64)
    karr = (ci * max(len(kinds), 1))(*kinds)
    vptrs = (vp * max(len(d_vectors), 1))(*[d.ptr for d in d_vectors])
    wptrs = (vp * max(len(witnesses), 1))(*[w.h for w in witnesses])
    cptrs = (vp * max(len(covs), 1))(*[c.ptr for c in covs])
    out = np.zeros(max(nv * D, 1), dtype=np.uint64)
    _call("whir_prove_forms_dev", cfg.h, t.h, vptrs, len(d_vectors), wptrs, len(witnesses), karr, len(forms),
          _p64(flat), cptrs, _p64(ev), _p64(out))
    return out[: nv * D]


def mle_evaluate(degree: int, d_vector: DeviceBuffer, num_variables: int, point) -> np.ndarray:
    pt = np.ascontiguousarray(np.asarray(point, dtype=np.uint64)).reshape(-1)
    if pt.size == 0:
        pt = np.zeros(1, dtype=np.uint64)
    out = np.zeros(degree, dtype=np.uint64)
    _call("whir_mle_evaluate_dev", degree, d_vector.ptr, num_variables, _p64(pt), _p64(out))
    return out


def config_memory(cfg: Config) -> int:
    """Estimated peak device bytes of one commit + prove under cfg."""
    n = sz()
    _call("whir_config_memory", cfg.h, ctypes.byref(n))
    return n.value


class Executor:
    """Batched / served proving: `workers` threads, each on its own HIP stream, run
    commit + prove (or commit only) for submitted jobs, admitted by an HBM budget
    (bytes; 0 = 80% of the free device memory).  Results are byte-identical to the
    same proofs made one at a time."""

    COMMIT_AND_PROVE = 0
    COMMIT_ONLY = 1

    def __init__(self, workers: int = 8, hbm_budget: int = 0):
        h = vp()
        _call("whir_executor_new", workers, hbm_budget, ctypes.byref(h))
        self.h = h.value
        self._keep = {}

    def submit(self, cfg: Config, t: Transcript, d_vector: DeviceBuffer, points=None, evaluations=None,
               kind: int = COMMIT_AND_PROVE) -> int:
        D, nv = cfg.degree, cfg.num_variables
        pts = np.ascontiguousarray(np.asarray(points if points is not None else [], dtype=np.uint64)).reshape(-1)
        ev = np.ascontiguousarray(np.asarray(evaluations if evaluations is not None else [],
                                             dtype=np.uint64)).reshape(-1)
        npts = ev.size // D
        assert pts.size == npts * nv * D, "points must be num_points x num_variables x degree words"
        pts = pts if pts.size else np.zeros(1, dtype=np.uint64)
        ev = ev if ev.size else np.zeros(1, dtype=np.uint64)
        jid = ctypes.c_uint64()
        _call("whir_executor_submit", self.h, kind, cfg.h, t.h, d_vector.ptr, _p64(pts), npts, _p64(ev),
              ctypes.byref(jid))
        # the transcript, config and vector must outlive the job
        self._keep[jid.value] = (cfg, t, d_vector, kind)
        return jid.value

    def submit_batch(self, cfg: Config, t: Transcript, d_vectors, points=None, evaluations=None) -> int:
        """Several commitments of batch_size vectors proven together (as prove_batch);
        evaluations: (num_points, num_vectors, degree)."""
        D, nv = cfg.degree, cfg.num_variables
        pts = np.ascontiguousarray(np.asarray(points if points is not None else [], dtype=np.uint64)).reshape(-1)
        ev = np.ascontiguousarray(np.asarray(evaluations if evaluations is not None else [],
                                             dtype=np.uint64)).reshape(-1)
        npts = pts.size // (nv * D) if nv else 0
        assert ev.size == npts * len(d_vectors) * D, "evaluations must be num_points x num_vectors x degree words"
        pts = pts if pts.size else np.zeros(1, dtype=np.uint64)
        ev = ev if ev.size else np.zeros(1, dtype=np.uint64)
        ptrs = (vp * max(len(d_vectors), 1))(*[d.ptr for d in d_vectors])
        jid = ctypes.c_uint64()
        _call("whir_executor_submit_batch", self.h, cfg.h, t.h, ptrs, len(d_vectors), _p64(pts), npts, _p64(ev),
              ctypes.byref(jid))
        self._keep[jid.value] = (cfg, t, list(d_vectors), self.COMMIT_AND_PROVE)
        return jid.value

    def submit_prove(self, cfg: Config, t: Transcript, d_vector: DeviceBuffer, witness: Witness, points=None,
                     evaluations=None) -> int:
        """Prove only, for a vector already committed to in t (witness from commit or
        commit_many)."""
        D, nv = cfg.degree, cfg.num_variables
        pts = np.ascontiguousarray(np.asarray(points if points is not None else [], dtype=np.uint64)).reshape(-1)
        ev = np.ascontiguousarray(np.asarray(evaluations if evaluations is not None else [],
                                             dtype=np.uint64)).reshape(-1)
        npts = ev.size // D
        assert pts.size == npts * nv * D, "points must be num_points x num_variables x degree words"
        pts = pts if pts.size else np.zeros(1, dtype=np.uint64)
        ev = ev if ev.size else np.zeros(1, dtype=np.uint64)
        jid = ctypes.c_uint64()
        _call("whir_executor_submit_prove", self.h, cfg.h, t.h, d_vector.ptr, witness.h, _p64(pts), npts, _p64(ev),
              ctypes.byref(jid))
        self._keep[jid.value] = (cfg, t, (d_vector, witness), self.COMMIT_AND_PROVE)
        return jid.value

    def wait(self, job_id: int):
        """Blocks until the job ends; returns ((submitted, started, finished) seconds,
        folding randomness (num_variables * degree words; empty for commit-only jobs)).
        Raises WhirError if the job failed."""
        times = (ctypes.c_double * 3)()
        cfg, _, _, kind = self._keep[job_id]
        words = cfg.num_variables * cfg.degree if kind == self.COMMIT_AND_PROVE else 0
        ep = np.zeros(max(words, 1), dtype=np.uint64)
        try:
            _call("whir_executor_wait", self.h, job_id, times, _p64(ep) if words else None)
        finally:
            self._keep.pop(job_id, None)
        return (times[0], times[1], times[2]), ep[:words]

    def clock(self) -> float:
        c = ctypes.c_double()
        _call("whir_executor_clock", self.h, ctypes.byref(c))
        return c.value

    def close(self):
        if self.h:
            lib().whir_executor_free(self.h)
            self.h = None
            self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
