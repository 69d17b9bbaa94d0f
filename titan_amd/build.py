"""Build recipe for the native libraries (run on the CPU; hipcc cross-compiles gfx950).

    python -m titan_amd.build            # libtitan_mi355x.so + the C++ facade library

Outputs, in-tree so that they travel with the repository snapshot:
    titan_amd/libtitan_mi355x.so   HIP kernels + C ABI (include/titan_mi355x.h)
    titan_amd/libtitan.so          C++ titan::Simulation facade over the C ABI
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
CPP = PKG / "cpp"
INCLUDE = ROOT / "include"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# -ffp-contract=off: every FP64 multiply and add is rounded separately, as
# the reference's source expressions are written, so passive spring networks
# reproduce the oracle bit for bit.  Packed FP32 instructions are disabled for
# all device code.
HIP_FLAGS = [
    "-O3", "-std=c++17", f"--offload-arch={ARCH}", "-fPIC",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-command-line-argument",
]

ENGINE_SOURCES = [CSRC / "kernels.hip", CSRC / "engine.hip", CSRC / "builder.hip", CSRC / "tiles.cpp",
                  CSRC / "lattice.cpp", CSRC / "serve.cpp"]
FACADE_SOURCES = [CPP / "sim.cpp"]

LIB = PKG / "libtitan_mi355x.so"
FACADE_LIB = PKG / "libtitan.so"


def _run(cmd: list[str], verbose: bool) -> None:
    if verbose:
        print(" ".join(str(c) for c in cmd), flush=True)
    res = subprocess.run([str(c) for c in cmd], capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"build step failed ({res.returncode}): {cmd[0]} ... {cmd[-1]}")
    if verbose and res.stderr.strip():
        # the host pass warns about the packed-fp32 feature flag; keep it quiet
        lines = [l for l in res.stderr.splitlines() if "packed-fp32-ops" not in l]
        if lines:
            print("\n".join(lines), file=sys.stderr)


def build_engine(verbose: bool = False) -> Path:
    objs = []
    bdir = PKG / "_build"
    bdir.mkdir(exist_ok=True)
    for src in ENGINE_SOURCES:
        obj = bdir / (src.name + ".o")
        lang = ["-x", "hip"] if src.suffix == ".hip" else []
        _run([HIPCC, *HIP_FLAGS, *lang, f"-I{INCLUDE}", f"-I{CSRC}", "-c", src, "-o", obj], verbose)
        objs.append(obj)
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp], verbose)
    os.replace(tmp, LIB)
    return LIB


def build_facade(verbose: bool = False) -> Path:
    tmp = FACADE_LIB.with_suffix(".so.tmp")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-Wall", "-Wextra",
          f"-I{INCLUDE}", f"-I{CPP}", *FACADE_SOURCES, "-o", tmp,
          f"-L{PKG}", "-ltitan_mi355x", "-Wl,-rpath,$ORIGIN"], verbose)
    os.replace(tmp, FACADE_LIB)
    return FACADE_LIB


FACADE_TESTS_SRC = ROOT / "tests" / "cpp" / "facade_tests.cpp"
FACADE_TESTS_BIN = ROOT / "tests" / "cpp" / "facade_tests"
# the same tests built as a program that passes the reference's RK2 build option
FACADE_TESTS_RK2_BIN = ROOT / "tests" / "cpp" / "facade_tests_rk2"


def build_facade_tests(verbose: bool = False, out: Path | None = None, defines=()) -> Path:
    """The reference's physics tests restated against the C++ facade."""
    out = out or FACADE_TESTS_BIN
    tmp = out.with_name(out.name + ".tmp")
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-Wextra", *[f"-D{d}" for d in defines], f"-I{CPP}",
          f"-I{INCLUDE}", FACADE_TESTS_SRC, "-o", tmp, f"-L{PKG}", "-ltitan", "-ltitan_mi355x",
          f"-Wl,-rpath,{PKG}"], verbose)
    os.replace(tmp, out)
    return out


FASTMATH_CHECK_SRC = ROOT / "tests" / "cpp" / "fastmath_check.hip"
FASTMATH_CHECK_BIN = ROOT / "tests" / "cpp" / "fastmath_check"


def build_fastmath_check(verbose: bool = False, out: Path | None = None) -> Path:
    """Bitwise check of the guarded FP64 division/sqrt against IEEE `/` and `sqrt`."""
    out = out or FASTMATH_CHECK_BIN
    tmp = out.with_name(out.name + ".tmp")
    _run([HIPCC, *HIP_FLAGS, "-x", "hip", f"-I{CSRC}", FASTMATH_CHECK_SRC, "-o", tmp], verbose)
    os.replace(tmp, out)
    return out


def build_all(verbose: bool = False) -> None:
    build_engine(verbose)
    build_facade(verbose)
    if FACADE_TESTS_SRC.exists():
        build_facade_tests(verbose)
        build_facade_tests(verbose, out=FACADE_TESTS_RK2_BIN, defines=("RK2",))
    if FASTMATH_CHECK_SRC.exists():
        build_fastmath_check(verbose)
    shutil.rmtree(PKG / "_build", ignore_errors=True)


if __name__ == "__main__":
    build_all(verbose=True)
    print("built", LIB)
