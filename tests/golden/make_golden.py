"""Writes the golden fixtures of this directory: the README scene (the
reference's docs/README example, 125 masses, 1036 springs, damping 0.1,
ground plane mu_k 0.8 / mu_s 1.0, dt 1e-4) stepped by pyref.py, the plain
Python restatement of the reference's step written apart from the C++
oracle, with positions and velocities stored as hexadecimal floats so that
they round-trip exactly.  The lattice comes from pyref's own restatement of
Lattice::Lattice.  The reference ships no golden vectors of its own (only
loose physical checks in test/physics); tests/test_golden.py checks that
the C++ oracle reproduces these, tests/test_golden_gpu.py that the HIP path
does.

Run from the repository root: python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE))

# (name, starting height of the lattice centre, steps, integrator)
CASES = [
    ("readme_cz10_euler_100", 10.0, 100, 0),
    ("readme_cz2.45_euler_1000", 2.45, 1000, 0),
    ("readme_cz2.45_verlet_300", 2.45, 300, 1),
    ("readme_cz2.45_rk2_300", 2.45, 300, 2),
]


def make(case):
    import pyref

    name, cz, steps, integ = case
    pos, left, right, rest = pyref.lattice((0.0, 0.0, cz), (5.0, 5.0, 5.0), 5, 5, 5)
    w = pyref.World(pos, [0.1] * len(pos), left, right, [1e4] * len(left), rest, [0.1] * len(left),
                    planes=[(0.0, 0.0, 1.0, 0.0, 0.8, 1.0)], dt=1e-4, integrator=integ)
    w.step(steps)
    return {
        "case": name, "center_z": cz, "steps": steps, "integrator": ["euler", "verlet", "rk2"][integ],
        "dt": w.dt, "time": float(w.t).hex(),
        "pos": [float(x).hex() for p in w.pos for x in p],
        "vel": [float(x).hex() for v in w.vel for x in v],
    }


def main():
    for case in CASES:
        data = make(case)
        (HERE / f"{case[0]}.json").write_text(json.dumps(data, indent=0) + "\n")
        print(f"wrote {case[0]}.json")


if __name__ == "__main__":
    main()
