"""The guarded FP64 division and square root used by the step kernels
(titan_amd/csrc/fastmath.h) against the device's IEEE `/` and `sqrt`, bit for
bit (a zero quotient may differ in sign), on 2^25 random inputs per case:
sqrt over [2^-767, 2^1000], quotients a / b over the fast range with +-0
numerators, and unit vectors t / |t| whose
components span tiny, ordinary and huge magnitudes so that both the fast path
and the fallback run."""
from __future__ import annotations

import subprocess

import pytest

from titan_amd import build


@pytest.mark.gpu
def test_fast_div_sqrt_bitwise_equal_to_ieee():
    exe = build.FASTMATH_CHECK_BIN
    if not exe.exists():
        raise RuntimeError(f"{exe} is not built; run __graft_entry__.build()")
    print("fastmath_check 33554432", flush=True)
    res = subprocess.run([str(exe), "33554432"], capture_output=True, text=True, timeout=120)
    print(res.stdout)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "sqrt mismatches 0" in res.stdout
    assert "div mismatches 0" in res.stdout
    assert "unit_axis mismatches 0" in res.stdout
