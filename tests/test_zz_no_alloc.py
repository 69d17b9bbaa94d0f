"""titan_step_noalloc against titan_step.

The no-allocation entry point runs the same kernels with their scratch
buffers (actuation factors, RK2 stage state) in the library's static device
array, launches every step directly and never rebuilds the ball grid, so its
states must equal titan_step's bit for bit: on two calls in a row (the
second on a smaller input, the array still holding the first call's
contents), when mixed with titan_step on one context, and a context that
needs more scratch than the array holds is refused before any work."""
from __future__ import annotations

import sys

import numpy as np
import pytest

from titan_amd import (INTEGRATOR_EULER, INTEGRATOR_RK2, LAYOUT_GATHER, LAYOUT_RESIDENT, LAYOUT_TILED,
                       TITAN_E_INVALID, Context, TitanError, scenes)

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _big():
    # three actuated robots (tiled, gather, resident), ten balls (the grid's
    # range), a ground plane with friction: 16 B of factors per actuator and,
    # under RK2, 48 B of stage state per mass slot in the scratch array
    sc = scenes.robot_population(3, side_lo=4, side_hi=11, seed=17)
    sides = [round(len(b["pos"]) ** (1 / 3)) for b in sc.bodies]
    assert len(set(sides)) == 3
    sc.planes = [(0.0, 0.0, 1.0, 0.0, 0.8, 1.0)]
    rng = np.random.default_rng(3)
    sc.balls = [(float(x), float(y), 0.2, 0.25) for x, y in rng.uniform(-0.5, 4.5, (10, 2))]
    return sc, [LAYOUT_TILED, LAYOUT_GATHER, LAYOUT_RESIDENT]


def _small():
    sc = scenes.robot_population(1, side_lo=5, side_hi=5, seed=29)
    return sc, [LAYOUT_RESIDENT]


def _run(sc, layouts, integ, plan):
    """plan: [(noalloc, steps), ...] applied in order on one context."""
    with Context() as ctx:
        ctx.set_integrator(integ)
        ids = sc.load(ctx, layout=layouts)
        for noalloc, n in plan:
            (ctx.step_noalloc if noalloc else ctx.step)(n)
        out = [ctx.download(i, acc=True) for i in ids]
    return {k: np.concatenate([o[k] for o in out]) for k in out[0]}


def _same(a, b):
    for k in ("pos", "vel", "acc"):
        np.testing.assert_array_equal(a[k], b[k], err_msg=k)


def test_noalloc_equals_step_on_two_calls_in_a_row():
    sc, lay = _big()
    _say(f"case noalloc call 1: {sc.n_masses} masses, {sc.n_actuated} actuated, 10 balls, RK2, 150 steps")
    got = _run(sc, lay, INTEGRATOR_RK2, [(True, 150)])
    want = _run(sc, lay, INTEGRATOR_RK2, [(False, 150)])
    _same(got, want)
    sc2, lay2 = _small()
    _say(f"case noalloc call 2: {sc2.n_masses} masses, {sc2.n_actuated} actuated, Euler, 70 steps")
    got = _run(sc2, lay2, INTEGRATOR_EULER, [(True, 70)])
    want = _run(sc2, lay2, INTEGRATOR_EULER, [(False, 70)])
    _same(got, want)


def test_noalloc_mixed_with_step_on_one_context():
    sc, lay = _big()
    _say("case titan_step (builds the ball grid, captures a graph), then noalloc, then step")
    got = _run(sc, lay, INTEGRATOR_EULER, [(False, 70), (True, 90), (False, 5), (True, 1)])
    want = _run(sc, lay, INTEGRATOR_EULER, [(False, 166)])
    _same(got, want)


def test_noalloc_refuses_a_context_beyond_its_scratch():
    # RK2 keeps 48 B of stage state per mass slot: 1.5M slots need 72 MB,
    # more than the 64 MiB array
    n = 1_500_000
    _say(f"case noalloc size limit: {n} masses under RK2")
    pos = np.zeros((n, 3))
    pos[:, 0] = np.arange(n) * 1e-3
    pos[:, 2] = 5.0
    with Context() as ctx:
        ctx.set_integrator(INTEGRATOR_RK2)
        bid = ctx.add_body(pos, 0.1)
        with pytest.raises(TitanError) as err:
            ctx.step_noalloc(3)
        assert err.value.code == TITAN_E_INVALID and "scratch" in str(err.value)
        assert ctx.body_time(bid)[1] == 0  # refused before any work
        ctx.step(1)
        assert ctx.body_time(bid)[1] == 1
