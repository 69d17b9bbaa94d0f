"""The README case (5x5x5 lattice, 125 masses, 1,036 springs, damping 0.1,
contact plane z = 0 with mu_k 0.8 / mu_s 1.0, dt = 1e-4, FP64) over 10,000
steps: the energy trace of the HIP path matches the oracle's.

Energy is kinetic + gravitational + spring potential + the contact plane's
penalty potential (NORMAL * disp^2 / 2 below the plane), evaluated by one
numpy function on both runs' states every 250 steps.  The bar is 1e-10
relative per sample; passive networks agree bit for bit, so the traces are
identical.  Two starts: centre z = 10 (the README's, free fall for the whole
run) and z = 2.45 (bottom face 5 cm into the plane, contact and friction
from the first step)."""
from __future__ import annotations

import numpy as np
import pytest

from titan_amd import Context, scenes
from helpers import to_oracle

pytestmark = pytest.mark.gpu

NORMAL = 20000.0
N_STEPS, EVERY = 10_000, 250


def energy(scene, pos, vel):
    b = scene.bodies[0]
    m = np.asarray(b["m"], float)
    ke = 0.5 * np.sum(m * np.sum(vel * vel, axis=1))
    pe = np.sum(m * 9.81 * pos[:, 2])
    d = pos[b["left"]] - pos[b["right"]]
    ext = np.sqrt(np.sum(d * d, axis=1)) - b["rest"]
    se = 0.5 * np.sum(b["k"] * ext * ext)
    disp = np.minimum(pos[:, 2], 0.0)
    ce = 0.5 * NORMAL * np.sum(disp * disp)
    return ke + pe + se + ce


@pytest.mark.parametrize("cz", [10.0, 2.45])
def test_energy_trace_10k_steps_matches_oracle(cz):
    sc = scenes.readme_lattice(cz)
    print(f"case energy trace cz={cz}, {N_STEPS} steps, sample every {EVERY}", flush=True)
    o = to_oracle(sc)
    gpu, orc = [], []
    with Context() as ctx:
        (bid,) = sc.load(ctx)
        for _ in range(N_STEPS // EVERY):
            ctx.step(EVERY)
            st = ctx.download(bid)
            o.run(EVERY)
            gpu.append(energy(sc, st["pos"], st["vel"]))
            orc.append(energy(sc, o.pos, o.vel))
    gpu, orc = np.array(gpu), np.array(orc)
    rel = np.abs(gpu - orc) / np.abs(orc)
    print(f"energy: first {orc[0]:.6f} J, last {orc[-1]:.6f} J, max rel diff GPU vs oracle {rel.max():.3e}")
    assert rel.max() <= 1e-10
    np.testing.assert_array_equal(gpu, orc)
