"""The GPU's Goldilocks formulations (whir_amd/csrc/gl.h, carry-chain add/sub,
128-bit product, reduction, mul_pow2), compiled for the host and checked against
unsigned __int128 arithmetic over all pairs of edge values and random pairs."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def _cxx():
    for c in ("/opt/rocm/llvm/bin/clang++", shutil.which("clang++") or ""):
        if c and os.path.exists(c):
            return c
    return None


@pytest.mark.skipif(_cxx() is None, reason="needs clang++ (ROCm llvm) for __builtin_addc/__builtin_subc")
def test_gpu_field_formulas_match_int128(tmp_path):
    exe = tmp_path / "gl_carry_check"
    subprocess.run([_cxx(), "-O2", "-std=c++17", "-w", os.path.join(HERE, "gl_carry_check.cpp"), "-o", str(exe)],
                   check=True)
    r = subprocess.run([str(exe), "4000000"], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("ok ")
