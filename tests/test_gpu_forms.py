"""Covector and UnivariateEvaluation linear forms (src/algebra/linear_form/) next
to MultilinearExtension points, in the reference tests' shapes (whir/mod.rs:215-231:
MLE points then the covector (0, 1, 2, ...)): GPU proofs equal the oracle's byte for
byte and its verifier accepts them."""
import random

import numpy as np
import pytest

import whir_amd as W
from oracle import oracle as O
from oracle import whir_protocol as WP
from whir_amd import prover as GP

pytestmark = pytest.mark.gpu


def make(nv, D, spec, B, NW, seed):
    vecs = [np.random.default_rng(seed + j).integers(0, O.P, size=1 << nv, dtype=np.uint64) for j in range(B * NW)]
    rng = random.Random(seed)
    forms = []
    for k, f in enumerate(spec):
        if f == "m":
            forms.append([np.array([rng.randrange(O.P) for _ in range(D)], dtype=np.uint64) for _ in range(nv)])
        elif f == "u":
            forms.append(WP.Univariate(np.array([rng.randrange(O.P) for _ in range(D)], dtype=np.uint64)))
        elif k == len(spec) - 1:
            v = np.zeros((1 << nv) * D, dtype=np.uint64)
            v[0::D] = np.arange(1 << nv, dtype=np.uint64)
            forms.append(WP.Covector(v))
        else:
            forms.append(WP.Covector(np.random.default_rng(seed + 50 + k).integers(0, O.P, size=(1 << nv) * D,
                                                                                     dtype=np.uint64)))
    evals = [[WP.form_evaluate(D, f, v) for v in vecs] for f in forms]
    return vecs, forms, evals


CASES = [(12, 2, "mmc", 1, 1), (10, 3, "c", 1, 1), (11, 2, "cm", 1, 1), (10, 2, "mcm", 2, 2),
         (14, 2, "mc", 1, 2), (9, 1, "cc", 3, 1), (10, 2, "umc", 1, 1), (11, 3, "u", 2, 1), (12, 2, "mucu", 1, 2)]


@pytest.mark.parametrize("nv,D,spec,B,NW", CASES)
def test_gpu_covector_forms_equal_oracle(nv, D, spec, B, NW):
    vecs, forms, evals = make(nv, D, spec, B, NW, seed=nv * 7 + D)
    params = dict(security_level=32, pow_bits=0, initial_folding_factor=4, folding_factor=4,
                  unique_decoding=(D == 1), starting_log_inv_rate=1, batch_size=B)
    pid, sid = WP.protocol_ids(f"gpu forms {nv} {D} {spec} {B} {NW}")
    ocfg = WP.WhirConfig.new(1 << nv, WP.ProtocolParameters(**params), D)
    ocfg.disable_pow()
    t = WP.Transcript(pid, sid)
    ows = [WP.commit_batch(t, ocfg, vecs[w * B:(w + 1) * B]) for w in range(NW)]
    oep, _ = WP.prove_batch(t, ocfg, vecs, ows, forms, evals)

    cfg = GP.Config(1 << nv, GP.ProtocolParameters(degree=D, **params)).disable_pow()
    dvs = [W.DeviceBuffer.from_array(v) for v in vecs]
    gforms = [W.DeviceBuffer.from_array(np.asarray(f.vector, dtype=np.uint64)) if isinstance(f, WP.Covector)
              else GP.Univariate(f.point) if isinstance(f, WP.Univariate) else np.concatenate(f) for f in forms]
    gt = GP.Transcript(pid, sid)
    gws = [GP.commit_batch(cfg, gt, dvs[w * B:(w + 1) * B]) for w in range(NW)]
    gev = np.array([[np.asarray(e, dtype=np.uint64) for e in row] for row in evals], dtype=np.uint64)
    gep = GP.prove_forms(cfg, gt, dvs, gws, gforms, gev)
    assert gt.proof() == (bytes(t.narg), bytes(t.hints))
    assert np.array_equal(gep, np.concatenate(oep))
    WP.verify_batch(WP.TranscriptReader(pid, sid, *gt.proof()), ocfg, NW, evals, forms)


def test_gpu_forms_argument_errors():
    nv, D = 10, 2
    cfg = GP.Config(1 << nv, GP.ProtocolParameters(degree=D, security_level=32, pow_bits=0))
    dv = W.DeviceBuffer.from_array(np.arange(1 << nv, dtype=np.uint64))
    t = GP.Transcript(b"\0" * 64, b"\0" * 32)
    w = GP.commit(cfg, t, dv)
    import ctypes
    kinds = (ctypes.c_int * 1)(7)
    with pytest.raises(W.WhirError, match="unknown linear form kind"):
        W._call("whir_prove_forms_dev", cfg.h, t.h, (ctypes.c_void_p * 1)(dv.ptr), 1, (ctypes.c_void_p * 1)(w.h),
                1, kinds, 1, None, None, W._p64(np.zeros(2, np.uint64)), None)
