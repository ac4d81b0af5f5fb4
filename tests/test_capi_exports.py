"""CPU checks of the C ABI library: it loads and exports every symbol that
include/whir_amd.h declares (no compute calls: there is no GPU here)."""
import os
import re

import pytest

import whir_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    syms = set()
    for f in os.listdir(os.path.join(ROOT, "include")):
        if f.endswith(".h"):
            text = open(os.path.join(ROOT, "include", f)).read()
            syms |= set(re.findall(r"\b(whir_[a-z0-9_]+)\s*\(", text))
    return syms


def test_header_declares_entry_points():
    syms = declared_symbols()
    assert "whir_rs_interleaved_encode" in syms and "whir_hash_many" in syms
    assert len(syms) >= 40


def test_library_exports_every_declared_symbol():
    L = whir_amd.lib()
    missing = [s for s in sorted(declared_symbols()) if not hasattr(L, s)]
    assert not missing, missing


def test_binding_covers_every_declared_symbol():
    declared = declared_symbols() - {"whir_last_error"}
    assert declared <= set(whir_amd.SIGNATURES), sorted(declared - set(whir_amd.SIGNATURES))


def test_host_only_entry_points_without_gpu():
    # These never touch the device: next_order / generator / evaluation_points / ids / thresholds.
    assert whir_amd.next_order(5) == 6
    assert whir_amd.next_order((3 << 32) + 1) is None
    P = whir_amd.P
    assert whir_amd.generator(1 << 20) == pow(7, (P - 1) >> 20, P)
    pts = whir_amd.evaluation_points(16, 64, list(range(64)))
    g = whir_amd.generator(64)
    assert [int(x) for x in pts] == [pow(g, (r // 16) + (r % 16) * 4, P) for r in range(64)]
    assert whir_amd.Engine(whir_amd.HASH_BLAKE3).engine_id().hex() == \
        "03e01749ebcc0477924254eb482066b864a8dd4d77252464ca6f5b6f5cc05b4c"
    assert whir_amd.Engine(whir_amd.HASH_SHA3).engine_id().hex() == \
        "aa802dcdabad8ea8a1430919893b96c30e31ff5734b385999108aa202d27dc12"
    assert whir_amd.Engine(whir_amd.HASH_SHA2).engine_id().hex() == \
        "018eaa247cb8957ab1e9fdac885450403c5e7bda1acaaa504e4cc8f2f76bb076"
    assert whir_amd.Engine(whir_amd.HASH_KECCAK).engine_id().hex() == \
        "ddd248964e320312a66775aee8e16c88c927734be59aca09b7af6deb0ad00e8c"
    assert whir_amd.pow_threshold(0.0) == 2**64 - 1
    assert whir_amd.pow_threshold(60.0) == 16
    assert all(whir_amd.pow_threshold(float(b)) == 1 << (64 - b) for b in range(1, 61))
    with pytest.raises(whir_amd.WhirError):
        whir_amd.pow_threshold(61.0)


def test_merkle_open_host_logic_matches_oracle():
    from oracle import oracle as O
    n = 100
    leaves = b"".join(bytes([i % 256]) * 32 for i in range(n))
    nodes = O.merkle_commit(O.ENGINE_SHA3, leaves)
    for idx in ([3, 4, 99], [0], [50, 51, 50, 7]):
        assert whir_amd.merkle_open(nodes, n, idx) == O.merkle_open(nodes, n, idx)
