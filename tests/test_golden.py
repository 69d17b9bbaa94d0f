"""The oracle still reproduces the committed golden fixtures (tests/golden,
written by tests/golden/make_golden.py) bit for bit.  CPU only."""
from __future__ import annotations

import json
from pathlib import Path

import numpy as np
import pytest

from helpers import to_oracle
from titan_amd import scenes

GOLDEN = sorted((Path(__file__).resolve().parent / "golden").glob("*.json"))
INTEG = {"euler": 0, "verlet": 1, "rk2": 2}


def load(path):
    d = json.loads(path.read_text())
    d["pos"] = np.array([float.fromhex(x) for x in d["pos"]]).reshape(-1, 3)
    d["vel"] = np.array([float.fromhex(x) for x in d["vel"]]).reshape(-1, 3)
    d["time"] = float.fromhex(d["time"])
    return d


def test_fixtures_exist():
    assert len(GOLDEN) >= 4


@pytest.mark.parametrize("path", GOLDEN, ids=[p.stem for p in GOLDEN])
def test_oracle_reproduces_fixture(path):
    g = load(path)
    sc = scenes.readme_lattice(g["center_z"])
    o = to_oracle(sc, threads=1, integrator=INTEG[g["integrator"]])
    o.run(g["steps"])
    np.testing.assert_array_equal(o.pos, g["pos"])
    np.testing.assert_array_equal(o.vel, g["vel"])
    assert o.t == g["time"]
