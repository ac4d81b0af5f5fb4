"""The library's host transcript (whir_amd/csrc/transcript.cpp: the duplex sponge that
the GPU prover's Fiat-Shamir steps run on) against the oracle's (Transcript in
oracle/whir_protocol.py), operation by operation: random scripts of prover messages,
challenges, hints and byte squeezes whose sizes cross the sponge rate (136 bytes) in
every position.  The driver tests/transcript_check.cpp is compiled for the host."""
import os
import random
import shutil
import subprocess

import numpy as np
import pytest

from oracle import whir_protocol as WP

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def _cxx():
    for c in ("/opt/rocm/llvm/bin/clang++", shutil.which("clang++") or "", shutil.which("g++") or ""):
        if c and os.path.exists(c):
            return c
    return None


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    if _cxx() is None:
        pytest.skip("needs a host C++ compiler")
    exe = tmp_path_factory.mktemp("tc") / "transcript_check"
    src = os.path.join(ROOT, "whir_amd", "csrc")
    subprocess.run([_cxx(), "-O2", "-std=c++17", "-w", "-I", src, "-I", os.path.join(ROOT, "include"),
                    "-I", "/opt/rocm/include", "-D__HIP_PLATFORM_AMD__", os.path.join(HERE, "transcript_check.cpp"),
                    os.path.join(src, "transcript.cpp"), os.path.join(src, "keccak_host.cpp"), "-o", str(exe)],
                   check=True)
    return str(exe)


def _script(seed):
    rng = random.Random(seed)
    ops = []
    for _ in range(60):
        k = rng.randrange(7)
        if k == 0:
            D, n = rng.choice((1, 2, 3)), rng.choice((1, 2, 17, 300))
            ops.append(("pf", D, [rng.randrange(WP.P) for _ in range(D * n)]))
        elif k == 1:
            ops.append(("vf", rng.choice((1, 2, 3))))
        elif k == 2:
            ops.append(("vb", rng.choice((1, 3, 24, 135, 136, 137, 300))))
        elif k == 3:
            ops.append(("ph", bytes(rng.randrange(256) for _ in range(32))))
        elif k == 4:
            ops.append(("pu", rng.randrange(1 << 64)))
        elif k == 5:
            n = rng.choice((1, 5, 40))
            ops.append(("hh", bytes(rng.randrange(256) for _ in range(32 * n))))
        else:
            D, n = rng.choice((1, 2, 3)), rng.choice((0, 1, 16))
            ops.append(("hv", D, [rng.randrange(WP.P) for _ in range(D * n)]))
    return ops


def _encode(ops, pid, sid):
    lines = [pid.hex() + " " + sid.hex()]
    for op in ops:
        if op[0] in ("pf", "hv"):
            D, w = op[1], op[2]
            lines.append(f"{op[0]} {D} {len(w) // D} " + " ".join(map(str, w)))
        elif op[0] in ("vf", "vb", "pu"):
            lines.append(f"{op[0]} {op[1]}")
        elif op[0] == "ph":
            lines.append("ph " + op[1].hex())
        else:
            lines.append(f"hh {len(op[1]) // 32} " + op[1].hex())
    return "\n".join(lines) + "\n"


def _oracle(ops, pid, sid):
    t = WP.Transcript(pid, sid)
    out = []
    for op in ops:
        if op[0] == "pf":
            t.prover_field(op[1], np.array(op[2], dtype=np.uint64))
        elif op[0] == "hv":
            t.hint_field_vec(op[1], np.array(op[2], dtype=np.uint64))
        elif op[0] == "vf":
            out.append("".join(f"{int(x)} " for x in t.verifier_field(op[1])))
        elif op[0] == "vb":
            out.append(t.verifier_bytes(op[1]).hex())
        elif op[0] == "ph":
            t.prover_hash(op[1])
        elif op[0] == "pu":
            t.prover_u64(op[1])
        else:
            for i in range(len(op[1]) // 32):
                t.hint_hash(op[1][32 * i:32 * i + 32])
    out.append(bytes(t.narg).hex())
    out.append(bytes(t.hints).hex())
    return out


@pytest.mark.parametrize("seed", range(6))
def test_host_transcript_matches_oracle(driver, seed):
    pid, sid = bytes(range(64)), bytes(range(100, 132))
    ops = _script(seed)
    r = subprocess.run([driver], input=_encode(ops, pid, sid), capture_output=True, text=True, check=True)
    got = [line for line in r.stdout.split("\n")][:-1]
    exp = _oracle(ops, pid, sid)
    assert len(got) == len(exp)
    for i, (g, e) in enumerate(zip(got, exp)):
        assert g == e, f"seed {seed}: output {i} differs"
