"""The C++ oracle against a second, independent restatement of the
reference's step (tests/golden/pyref.py: plain Python floats, a serial run of
the spring scatter in spring order, then the mass pass), bit for bit, on
small scenes that take every branch the oracle has: contact with friction
(kinetic and static), balls inside and outside, fixed masses and springs to
them, integer drag, external force, local contact planes, balls, constraint
planes and directions on chosen masses, springs with a missing endpoint, the
three actuation kinds, and each integrator.  Also: the lattice generator of
titan_amd.scenes against the restated Lattice::Lattice.  CPU only."""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import pytest

from oracle import oracle as orc
from titan_amd import scenes

sys.path.insert(0, str(Path(__file__).resolve().parent / "golden"))
import pyref  # noqa: E402


def small_scene(seed: int, integrator: int):
    """A jittered 4x3x3 lattice sitting 2 cm into a frictional ground plane,
    plus everything else the mass pass can apply."""
    rng = np.random.default_rng(seed)
    pos, left, right, _ = pyref.lattice((0.0, 0.0, 0.48), (1.0, 1.0, 1.0), 4, 3, 3)
    pos = np.array(pos) + rng.normal(0.0, 1e-3, size=(len(pos), 3))
    n, s = len(pos), len(left)
    left, right = np.array(left), np.array(right)
    left[5] = -1  # a spring with a missing endpoint does nothing (src/sim.cu:1052)
    rest = np.array([np.linalg.norm(pos[r] - pos[l]) if l >= 0 else 1.0 for l, r in zip(left, right)])
    rest *= rng.uniform(0.97, 1.03, size=s)
    k = rng.uniform(2e3, 2e4, size=s)
    damping = rng.uniform(0.0, 0.5, size=s)
    types = rng.integers(0, 5, size=s)
    omega = rng.uniform(1.0, 20.0, size=s)
    act_a = np.where(types == pyref.GENERAL, rng.uniform(-0.2, 0.2, size=s), 0.0)
    act_phi = np.where(types == pyref.GENERAL, rng.uniform(0.0, 6.28, size=s), 0.0)
    m = rng.uniform(0.05, 0.2, size=n)
    vel = rng.normal(0.0, 0.3, size=(n, 3))
    vel[3] = 0.0  # a mass at rest, on the ground: the static friction branch
    fixed = np.zeros(n, np.uint8)
    fixed[[7, 20]] = 1
    drag = rng.integers(0, 3, size=n).astype(np.int32)
    ext = rng.normal(0.0, 0.5, size=(n, 3))
    # local constraints: (kind, vector, number) per mass, in insertion order
    local = [[] for _ in range(n)]
    local[1] = [(pyref.BALL, (0.2, 0.1, 0.5), 0.4), (pyref.CONTACT_PLANE, (0.0, 1.0, 0.0), -0.3)]
    local[9] = [(pyref.CONSTRAINT_PLANE, (0.0, 0.0, 1.0), 0.3)]
    local[12] = [(pyref.DIRECTION, (1.0, 0.0, 0.0), 0.2), (pyref.BALL, (5.0, 5.0, 5.0), 0.1)]
    local[30] = [(pyref.CONTACT_PLANE, (0.0, 0.0, 1.0), 0.1), (pyref.CONSTRAINT_PLANE, (0.6, 0.0, 0.8), 0.5)]
    planes = [(0.0, 0.0, 1.0, 0.0, 0.8, 1.0), (0.6, 0.0, 0.8, -0.2, 0.5, 0.7)]
    balls = [(0.1, 0.0, 0.3, 0.25), (3.0, 3.0, 3.0, 0.5)]
    py = pyref.World(pos, m, left, right, k, rest, damping, vel=vel, planes=planes, balls=balls, fixed=fixed,
                     drag=drag, ext=ext, types=types, omega=omega, act_a=act_a, act_phi=act_phi, local=local,
                     dt=2e-4, integrator=integrator)
    lc_ptr, lc_type, lc_vec, lc_num = [0], [], [], []
    for lst in local:
        for kind, v, num in lst:
            lc_type.append(kind)
            lc_vec.append(v)
            lc_num.append(num)
        lc_ptr.append(len(lc_type))
    o = orc.Scene(pos, m, vel=vel, ext=ext, fixed=fixed, drag=drag, left=left, right=right, k=k, rest=rest,
                  damping=damping, type=types, omega=omega, act_a=act_a, act_phi=act_phi, planes=planes,
                  balls=balls, lc_ptr=lc_ptr, lc_type=lc_type, lc_vec=np.array(lc_vec), lc_num=lc_num,
                  dt=2e-4, threads=1, integrator=integrator)
    return py, o


@pytest.mark.parametrize("integrator", [orc.EULER, orc.VERLET, orc.RK2], ids=["euler", "verlet", "rk2"])
@pytest.mark.parametrize("seed", [0, 1])
def test_oracle_equals_independent_restatement(seed, integrator):
    py, o = small_scene(seed, integrator)
    for chunk in (1, 39, 160):  # 200 steps, compared after each chunk
        py.step(chunk)
        o.run(chunk)
        np.testing.assert_array_equal(o.pos, np.array(py.pos))
        np.testing.assert_array_equal(o.vel, np.array(py.vel))
        np.testing.assert_array_equal(o.acc, np.array(py.acc))
        np.testing.assert_array_equal(o.T, np.array(py.T))
        assert o.t == py.t


def test_small_scene_takes_every_branch():
    """The scenes above are not vacuous: masses touch the plane with and
    without sliding, a ball holds a mass, drag and the local kinds act."""
    py, _ = small_scene(0, orc.EULER)
    hits = {"kinetic": 0, "static": 0, "ball": 0}
    n_ref = py.contact_plane

    def spy_plane(pos, vel, force, n, offset, mu_k, mu_s):
        if pyref.dot(pos, n) - offset < 0 and (mu_s > 0 or mu_k > 0):
            vp = pyref.sub(vel, pyref.mul(n, pyref.dot(vel, n)))
            hits["kinetic" if pyref.norm(vp) > 1e-16 else "static"] += 1
        return n_ref(pos, vel, force, n, offset, mu_k, mu_s)

    b_ref = py.ball

    def spy_ball(pos, force, c, r):
        if pyref.norm(pyref.sub(pos, c)) <= r:
            hits["ball"] += 1
        return b_ref(pos, force, c, r)

    py.contact_plane, py.ball = spy_plane, spy_ball
    py.step(5)
    assert hits["kinetic"] > 0 and hits["static"] > 0 and hits["ball"] > 0, hits


@pytest.mark.parametrize("dims", [(5, 5, 5), (3, 4, 2), (1, 3, 4), (2, 1, 1)])
def test_scene_lattice_matches_restated_generator(dims):
    nx, ny, nz = dims
    pos, left, right, rest = scenes.lattice((0.3, -1.0, 2.0), (1.5, 2.0, 2.5), nx, ny, nz)
    p2, l2, r2, s2 = pyref.lattice((0.3, -1.0, 2.0), (1.5, 2.0, 2.5), nx, ny, nz)
    np.testing.assert_array_equal(np.asarray(pos), np.array(p2).reshape(-1, 3))
    np.testing.assert_array_equal(np.asarray(left), l2)
    np.testing.assert_array_equal(np.asarray(right), r2)
    np.testing.assert_array_equal(np.asarray(rest), s2)
