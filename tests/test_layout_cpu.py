"""Host-side checks of the device layout builders (no GPU): the builders are
compiled with g++ into a small driver that compares every tile's halo and
every mass's incidence list (other endpoint as arena slot or LDS row, spring
parameters, ascending spring order, padding) with brute force on random
graphs and a Morton-ordered lattice."""
import subprocess
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent


def test_tile_builder_invariants(tmp_path):
    exe = tmp_path / "tiles_check"
    subprocess.run(["g++", "-O2", "-std=c++17", f"-I{ROOT / 'titan_amd' / 'csrc'}",
                    str(ROOT / "tests" / "cpp" / "tiles_check.cpp"), str(ROOT / "titan_amd" / "csrc" / "tiles.cpp"),
                    "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0 and out.stdout.strip().endswith("OK"), out.stdout
