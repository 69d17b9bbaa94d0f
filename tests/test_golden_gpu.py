"""The HIP path reproduces the committed golden fixtures (tests/golden) bit
for bit, in every layout, without running the oracle."""
from __future__ import annotations

import numpy as np
import pytest

from test_golden import GOLDEN, INTEG, load
from titan_amd import LAYOUT_GATHER, LAYOUT_RESIDENT, LAYOUT_TILED, Context, scenes

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("layout", [LAYOUT_GATHER, LAYOUT_TILED, LAYOUT_RESIDENT], ids=["gather", "tiled", "resident"])
@pytest.mark.parametrize("path", GOLDEN, ids=[p.stem for p in GOLDEN])
def test_hip_path_reproduces_fixture(path, layout):
    g = load(path)
    print(f"case golden {path.stem} layout {layout}", flush=True)
    sc = scenes.readme_lattice(g["center_z"])
    with Context() as ctx:
        ctx.set_integrator(INTEG[g["integrator"]])
        (bid,) = sc.load(ctx, layout=layout)
        ctx.step(g["steps"])
        got = ctx.download(bid)
        t = ctx.time()
    np.testing.assert_array_equal(got["pos"], g["pos"])
    np.testing.assert_array_equal(got["vel"], g["vel"])
    assert t == g["time"]
