"""whir_commit_host: a commit whose vector arrives from host memory, uploaded a chunk
of messages at a time with each chunk's encode overlapping the later copies.  The
device copy, the witness (through the proof made from it) and the transcript equal
those of uploading the vector first and calling whir_commit_dev, for pinned and
pageable host memory."""
import hashlib

import numpy as np
import pytest

import whir_amd as W
from oracle import oracle as O
from whir_amd import prover as GP

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nv,D,lir,pinned", [(20, 2, 1, True), (20, 2, 1, False), (16, 3, 1, True),
                                             (14, 1, 2, True), (10, 2, 1, False), (22, 2, 2, True)])
def test_commit_host_equals_upload_then_commit(nv, D, lir, pinned):
    rng = np.random.default_rng(nv * 7 + D)
    vec = rng.integers(0, O.P, size=1 << nv, dtype=np.uint64)
    cfg = GP.Config(1 << nv, GP.ProtocolParameters(degree=D, security_level=32, pow_bits=0,
                                                   starting_log_inv_rate=lir)).disable_pow()
    pts = rng.integers(0, O.P, size=nv * D, dtype=np.uint64)
    pid = hashlib.sha3_512(f"commit host {nv} {D} {lir}".encode()).digest()
    sid = hashlib.sha3_256(b"commit host").digest()
    dv = W.DeviceBuffer.from_array(vec)
    ev = GP.mle_evaluate(D, dv, nv, pts)
    t0 = GP.Transcript(pid, sid)
    w0 = GP.commit(cfg, t0, dv)
    after0 = t0.proof()
    GP.prove(cfg, t0, dv, w0, pts, ev)

    host = W.PinnedArray(vec.size) if pinned else vec.copy()
    if pinned:
        host.array[:] = vec
    dv2 = W.DeviceBuffer(vec.nbytes)
    for rep in range(2):  # the second call reuses the copy stream and its events
        t1 = GP.Transcript(pid, sid)
        w1 = GP.commit_host(cfg, t1, host, dv2)
        assert t1.proof() == after0, "transcript after the host commit differs"
        assert np.array_equal(dv2.to_array(), vec), "device copy of the vector differs"
        GP.prove(cfg, t1, dv2, w1, pts, ev)
        assert t1.proof() == t0.proof(), "proof from the host commit's witness differs"
