"""The C ABI library on the CPU: it loads, exports every function that
include/titan_mi355x.h declares, its host-only entry points agree with the
oracle, and its error paths report cleanly without a GPU."""
from __future__ import annotations

import ctypes
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

import titan_amd
from oracle import oracle as orc

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "titan_mi355x.h"


def declared_functions() -> list[str]:
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^[A-Za-z_][\w\s\*]*?\b(titan_\w+)\s*\(", text, flags=re.M)))


def test_library_exports_every_declared_symbol():
    names = declared_functions()
    assert len(names) >= 20
    out = subprocess.run(["nm", "-D", "--defined-only", str(titan_amd.LIB_PATH)], capture_output=True,
                         text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    assert sorted(titan_amd.EXPORTS) == names
    L = titan_amd.lib()
    for n in names:
        assert getattr(L, n) is not None


@pytest.mark.parametrize("dims", [(5, 5, 5), (3, 4, 7), (1, 6, 2), (12, 3, 1)])
def test_lattice_generator_matches_oracle(dims):
    a = titan_amd.lattice((0.5, -1.0, 10.0), (5.0, 4.0, 3.0), *dims)
    b = orc.lattice((0.5, -1.0, 10.0), (5.0, 4.0, 3.0), *dims)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)


def test_lattice_spring_count_closed_form():
    L = titan_amd.lib()
    for n in (2, 5, 17, 100):
        assert L.titan_lattice_spring_count(n, n, n) == 3 * n * n * (n - 1) + 6 * n * (n - 1) ** 2 + 4 * (n - 1) ** 3
    assert L.titan_lattice_spring_count(0, 3, 3) == 0
    assert L.titan_lattice(0, 0, 0, 1, 1, 1, 0, 2, 2, None, None, None, None) == titan_amd.TITAN_E_INVALID


@pytest.mark.skipif(Path("/dev/kfd").exists(), reason="checks the no-GPU error path")
def test_context_creation_without_gpu_fails_cleanly():
    L = titan_amd.lib()
    h = ctypes.c_void_p(1234)
    rc = L.titan_ctx_create(0, 0, ctypes.byref(h))
    assert rc in (titan_amd.TITAN_E_HIP, titan_amd.TITAN_E_INVALID)
    assert h.value is None
    assert len(L.titan_last_error(None)) > 0
    with pytest.raises(titan_amd.TitanError):
        titan_amd.Context()


def test_null_context_is_rejected():
    L = titan_amd.lib()
    assert L.titan_step(None, 1) == titan_amd.TITAN_E_INVALID
    assert L.titan_sync(None) == titan_amd.TITAN_E_INVALID
    assert L.titan_body_destroy(None, 0) == titan_amd.TITAN_E_INVALID
    assert L.titan_ctx_destroy(None) == titan_amd.TITAN_OK
    assert L.titan_last_error(None)  # never NULL
