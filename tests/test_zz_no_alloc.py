"""whir_fold_and_compute_polynomial_dev_noalloc (the allocation-free sumcheck round)
against whir_fold_and_compute_polynomial_dev on the same inputs: identical folded
vectors, (c0, c2) and new length, on two calls in a row (the second smaller, so it
runs on the scratch the first left behind), for every degree."""
import ctypes

import numpy as np
import pytest

import whir_amd as W
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def run(name, D, a, b, w):
    da, db = W.DeviceBuffer.from_array(a), W.DeviceBuffer.from_array(b)
    out = np.zeros(2 * D, dtype=np.uint64)
    m = ctypes.c_size_t()
    n = a.size // D
    W._call(name, D, da.ptr, db.ptr, n, W._p64(np.ascontiguousarray(w)), W._p64(out), ctypes.byref(m))
    W.synchronize()
    return m.value, out, da.to_array(count=m.value * D), db.to_array(count=m.value * D)


@pytest.mark.parametrize("D", [1, 2, 3])
def test_noalloc_round_equals_allocating_round(D):
    rng = np.random.default_rng(100 + D)
    # 2^21 + 3 elements: a ragged length (zero padding to 2^22) over 2048 blocks of
    # partials; then 5000 (one block) and 2^16 (a few blocks) on the scratch left behind.
    for n in ((1 << 21) + 3, 5000, 1 << 16, 2, 1):
        a = rng.integers(0, O.P, size=n * D, dtype=np.uint64)
        b = rng.integers(0, O.P, size=n * D, dtype=np.uint64)
        w = rng.integers(0, O.P, size=D, dtype=np.uint64)
        want = run("whir_fold_and_compute_polynomial_dev", D, a, b, w)
        got = run("whir_fold_and_compute_polynomial_dev_noalloc", D, a, b, w)
        assert got[0] == want[0], f"new length differs at n={n}"
        for x, y, what in zip(got[1:], want[1:], ("(c0, c2)", "folded a", "folded b")):
            assert np.array_equal(x, y), f"{what} differs at n={n}, D={D}"


def test_noalloc_round_rejects_bad_arguments():
    out = np.zeros(2, dtype=np.uint64)
    m = ctypes.c_size_t()
    with pytest.raises(W.WhirError, match="degree"):
        W._call("whir_fold_and_compute_polynomial_dev_noalloc", 4, None, None, 0, W._p64(out), W._p64(out),
                ctypes.byref(m))
    with pytest.raises(W.WhirError, match="length above 2\\^62"):
        buf = W.DeviceBuffer(64)
        W._call("whir_fold_and_compute_polynomial_dev_noalloc", 1, buf.ptr, buf.ptr, (1 << 62) + 1,
                W._p64(out), W._p64(out), ctypes.byref(m))
