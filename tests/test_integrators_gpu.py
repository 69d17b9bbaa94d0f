"""The reference's VERLET and RK2 build options (src/sim.cu:1141-1160,
1572-1582) against the oracle.

Passive springs must be bit-identical to the oracle under each integrator
and in every layout: positions, velocities, accelerations, per-mass times
and the context clock (RK2 advances it by dt / 2 twice per step).  RK2's
second stage starts from the state its first stage saved after the
constraints acted, which constraint planes and directions change; actuated
springs see the stage's own time.  Actuated cases are held to the 1e-10
tolerance of the path."""
from __future__ import annotations

import sys

import numpy as np
import pytest

from titan_amd import (INTEGRATOR_EULER, INTEGRATOR_RK2, INTEGRATOR_VERLET, LAYOUT_GATHER,
                       LAYOUT_RESIDENT, LAYOUT_TILED, Context, scenes)
from helpers import normwise_rel, to_oracle

pytestmark = pytest.mark.gpu

INTEGS = pytest.mark.parametrize("integ", [INTEGRATOR_VERLET, INTEGRATOR_RK2], ids=["verlet", "rk2"])
LAYOUTS = pytest.mark.parametrize("layout", [LAYOUT_GATHER, LAYOUT_TILED, LAYOUT_RESIDENT],
                                  ids=["gather", "tiled", "resident"])


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _engine(sc, integ, layout, calls):
    with Context() as ctx:
        ctx.set_integrator(integ)
        assert ctx.integrator() == integ
        ids = sc.load(ctx, layout=layout)
        for n in calls:
            ctx.step(n)
        out = [ctx.download(i, acc=True, T=True) for i in ids]
        t = ctx.time()
    return {k: np.concatenate([o[k] for o in out]) for k in out[0]}, t


def _oracle(sc, integ, n):
    o = to_oracle(sc, integrator=integ)
    o.run(n)
    return o


@INTEGS
@LAYOUTS
def test_contact_lattice_bitexact(integ, layout):
    sc = scenes.readme_lattice(2.45)
    b = sc.bodies[0]
    rng = np.random.default_rng(5)
    b["vel"] = rng.normal(0, 0.3, b["pos"].shape)
    calls = [1, 64, 70, 3]
    _say(f"case integrator {integ} layout {layout}: README lattice in contact, calls {calls}")
    eng, t = _engine(sc, integ, layout, calls)
    orc = _oracle(sc, integ, sum(calls))
    np.testing.assert_array_equal(eng["pos"], orc.pos)
    np.testing.assert_array_equal(eng["vel"], orc.vel)
    np.testing.assert_array_equal(eng["acc"], orc.acc)
    np.testing.assert_array_equal(eng["T"], orc.T)
    assert t == orc.t


@INTEGS
@pytest.mark.parametrize("layout", [LAYOUT_GATHER, LAYOUT_TILED], ids=["gather", "tiled"])
def test_constraints_that_change_velocity_bitexact(integ, layout):
    # constraint planes and directions project the velocity before the
    # integrator runs: RK2's saved state must be the projected one
    sc = scenes.readme_lattice(2.0)
    b = sc.bodies[0]
    n = len(b["pos"])
    rng = np.random.default_rng(6)
    cnt = rng.integers(0, 3, n)
    b["lc_ptr"] = np.concatenate([[0], np.cumsum(cnt)]).astype(np.int64)
    tot = int(cnt.sum())
    b["lc_type"] = rng.choice([0, 3, 1], size=tot).astype(np.int32)  # constraint plane, direction, contact
    v = rng.normal(size=(tot, 3))
    b["lc_vec"] = v / np.sqrt((v * v).sum(1))[:, None]
    b["lc_num"] = rng.uniform(0.0, 0.4, tot)
    b["vel"] = rng.normal(0, 0.3, (n, 3))
    b["fixed"] = (np.arange(n) % 19 == 0).astype(np.uint8)
    b["drag"] = (np.arange(n) % 4).astype(np.int32)
    b["ext"] = rng.normal(0, 0.2, (n, 3)) * (np.arange(n) % 2 == 0)[:, None]
    sc.balls = [(0.5, 0.5, 1.5, 0.8)]
    _say(f"case integrator {integ} layout {layout}: velocity constraints, fixed, drag, extern, ball; 150 steps")
    eng, _ = _engine(sc, integ, layout, [150])
    orc = _oracle(sc, integ, 150)
    np.testing.assert_array_equal(eng["pos"], orc.pos)
    np.testing.assert_array_equal(eng["vel"], orc.vel)
    np.testing.assert_array_equal(eng["T"], orc.T)


@INTEGS
def test_actuated_springs_within_tolerance_and_alike_across_layouts(integ):
    rng = np.random.default_rng(7)
    b = scenes.actuate(scenes.lattice_body((0.0, 0.0, 0.58), (1.0, 1.0, 1.0), 7, damping=0.1), 0.3, rng)
    sc = scenes.Scene(bodies=[b], planes=[(0.0, 0.0, 1.0, 0.0, 0.8, 1.0)], dt=1e-4)
    _say(f"case integrator {integ}: actuated 7^3 lattice, 120 steps, three layouts")
    runs = [_engine(sc, integ, lay, [120])[0] for lay in (LAYOUT_GATHER, LAYOUT_TILED, LAYOUT_RESIDENT)]
    for r in runs[1:]:
        np.testing.assert_array_equal(r["pos"], runs[0]["pos"])
        np.testing.assert_array_equal(r["vel"], runs[0]["vel"])
    orc = _oracle(sc, integ, 120)
    assert normwise_rel(runs[0]["pos"], orc.pos) <= 1e-10
    assert normwise_rel(runs[0]["vel"], orc.vel) <= 1e-10


@pytest.mark.parametrize("integ", [INTEGRATOR_EULER, INTEGRATOR_VERLET, INTEGRATOR_RK2],
                         ids=["euler", "verlet", "rk2"])
def test_resident_bodies_in_the_arena_bitexact(integ):
    # more resident bodies than compute units: each steps in its arena state
    # buffers, under each integrator
    rng = np.random.default_rng(9)
    bodies = []
    for q in range(300):
        b = scenes.lattice_body((2.0 * (q % 20), 2.0 * (q // 20), 0.6), (1.0, 1.0, 1.0), 3 + q % 4, damping=0.1)
        b["pos"] = b["pos"] + rng.normal(0, 1e-3, b["pos"].shape)
        b["vel"] = rng.normal(0, 0.3, b["pos"].shape)
        bodies.append(b)
    sc = scenes.Scene(bodies=bodies, planes=[(0.0, 0.0, 1.0, 0.0, 0.8, 1.0)], dt=1e-4)
    calls = [1, 64, 33]
    _say(f"case integrator {integ}: 300 resident bodies (27..216 masses), calls {calls}")
    with Context() as ctx:
        ctx.set_integrator(integ)
        ids = sc.load(ctx, layout=LAYOUT_RESIDENT)
        for n in calls:
            ctx.step(n)
        got = [ctx.download(i, acc=True, T=True) for i in ids]
    for q in range(0, 300, 23):
        o = to_oracle(scenes.Scene(bodies=[bodies[q]], planes=sc.planes, dt=sc.dt), integrator=integ)
        o.run(sum(calls))
        np.testing.assert_array_equal(got[q]["pos"], o.pos)
        np.testing.assert_array_equal(got[q]["vel"], o.vel)
        np.testing.assert_array_equal(got[q]["acc"], o.acc)
        np.testing.assert_array_equal(got[q]["T"], o.T)


def test_rk2_budgets_clock_and_breakpoint():
    a = scenes.readme_lattice(2.45)
    c = scenes.lattice_body((-9.0, 0.0, 2.2), (4.0, 4.0, 4.0), (9, 9, 9), damping=0.2)
    sc = scenes.Scene(bodies=[a.bodies[0], c], planes=a.planes, dt=a.dt)
    _say("case RK2 step_bodies budgets, body clocks and step_until")
    with Context() as ctx:
        ctx.set_integrator(INTEGRATOR_RK2)
        ids = sc.load(ctx, layout=[LAYOUT_RESIDENT, LAYOUT_TILED])
        ctx.step_bodies(ids, [5, 11])
        ctx.step_bodies([ids[1]], [2])
        k = ctx.step_until(0.0031)
        got = [ctx.download(i, T=True) for i in ids]
        clocks = [ctx.body_time(i) for i in ids]
        t = ctx.time()
    # the breakpoint rule with the RK2 clock
    T, n = 0.0, 0
    while not (0.0031 <= T):
        T += 0.5 * sc.dt
        T += 0.5 * sc.dt
        n += 1
    assert k == n and t == T
    for q, steps in enumerate((5 + n, 13 + n)):
        o = to_oracle(scenes.Scene(bodies=[sc.bodies[q]], planes=sc.planes, dt=sc.dt), integrator=INTEGRATOR_RK2)
        o.run(steps)
        np.testing.assert_array_equal(got[q]["pos"], o.pos)
        np.testing.assert_array_equal(got[q]["vel"], o.vel)
        np.testing.assert_array_equal(got[q]["T"], o.T)
        assert clocks[q] == (o.t, steps)


def test_integrator_is_chosen_before_bodies():
    sc = scenes.readme_lattice(2.45)
    with Context() as ctx:
        assert ctx.integrator() == INTEGRATOR_EULER
        sc.load(ctx)
        with pytest.raises(RuntimeError):
            ctx.set_integrator(INTEGRATOR_RK2)
        with pytest.raises(RuntimeError):
            ctx.set_integrator(7)
