"""Writes the golden fixtures of this directory: the README scene (the
reference's docs/README example, 125 masses, 1036 springs, damping 0.1,
ground plane mu_k 0.8 / mu_s 1.0, dt 1e-4) stepped by the CPU oracle, with
positions and velocities stored as hexadecimal floats so that they round-trip
exactly.  The reference ships no golden vectors of its own (only loose
physical checks in test/physics), so these pin the oracle and the HIP path
to each other across changes: tests/test_golden.py checks that the oracle
still produces them, tests/test_golden_gpu.py that the HIP path does.

Run from the repository root: python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent.parent))
sys.path.insert(0, str(HERE.parent))

# (name, starting height of the lattice centre, steps, integrator)
CASES = [
    ("readme_cz10_euler_100", 10.0, 100, 0),
    ("readme_cz2.45_euler_1000", 2.45, 1000, 0),
    ("readme_cz2.45_verlet_300", 2.45, 300, 1),
    ("readme_cz2.45_rk2_300", 2.45, 300, 2),
]


def make(case):
    from helpers import to_oracle
    from titan_amd import scenes

    name, cz, steps, integ = case
    sc = scenes.readme_lattice(cz)
    o = to_oracle(sc, threads=1, integrator=integ)
    o.run(steps)
    return {
        "case": name, "center_z": cz, "steps": steps, "integrator": ["euler", "verlet", "rk2"][integ],
        "dt": sc.dt, "time": float(o.t).hex(),
        "pos": [float(x).hex() for x in o.pos.reshape(-1)],
        "vel": [float(x).hex() for x in o.vel.reshape(-1)],
    }


def main():
    for case in CASES:
        data = make(case)
        (HERE / f"{case[0]}.json").write_text(json.dumps(data, indent=0) + "\n")
        print(f"wrote {case[0]}.json")


if __name__ == "__main__":
    main()
