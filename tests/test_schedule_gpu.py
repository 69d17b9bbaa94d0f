"""Schedules and the autotuner (titan_set_schedule, titan_autotune).

A schedule only changes how bodies are laid out on the device, so every
schedule must give bit-identical states; the autotuner must return one of its
candidates, make it the context's schedule, and bodies created afterwards must
follow it; bad schedules are rejected."""
from __future__ import annotations

import numpy as np
import pytest

from titan_amd import AUTOTUNE_CANDIDATES, LAYOUT_GATHER, LAYOUT_TILED, Context, scenes

pytestmark = pytest.mark.gpu


def _robots(n=12, seed=3):
    return scenes.robot_population(n, side_lo=3, side_hi=12, seed=seed)


def _run(sc, sched, steps=60):
    with Context() as ctx:
        ctx.set_dt(sc.dt)
        ctx.set_global_acc(sc.g)
        ctx.set_schedule(*sched)
        ids = [ctx.add_body(**b) for b in sc.bodies]
        layouts = [ctx.body_layout(i) for i in ids]
        ctx.step(steps)
        st = [ctx.download(i) for i in ids]
    return st, layouts


def test_every_schedule_gives_identical_states():
    sc = _robots()
    print(f"case schedules on {len(sc.bodies)} robots ({sc.n_masses} masses)", flush=True)
    ref, lay = _run(sc, (0, 2304))
    assert all(l == LAYOUT_TILED for l in lay)
    for sched in [(10**9, 2304), (256, 682), (512, 2304), (0, 682), (0, 64)]:
        st, lay = _run(sc, sched)
        n = [len(b["pos"]) for b in sc.bodies]
        assert lay == [LAYOUT_TILED if k >= sched[0] else LAYOUT_GATHER for k in n]
        for a, b in zip(st, ref):
            np.testing.assert_array_equal(a["pos"], b["pos"])
            np.testing.assert_array_equal(a["vel"], b["vel"])


def test_autotune_picks_a_candidate_and_applies_it():
    sc = _robots(8, seed=5)
    with Context() as ctx:
        ctx.set_dt(sc.dt)
        ctx.set_global_acc(sc.g)
        best, rates = ctx.autotune(sc.bodies, replicas=2, steps=4)
        assert len(rates) == len(AUTOTUNE_CANDIDATES) == 12
        assert all(r > 0 for _, r in rates)
        won = max(rates, key=lambda cr: cr[1])[0]
        want = (2**63 - 1 if won[0] < 0 else won[0], won[1])
        assert best == want and ctx.schedule() == best
        bid = ctx.add_body(**sc.bodies[0])
        n = len(sc.bodies[0]["pos"])
        assert ctx.body_layout(bid) == (LAYOUT_TILED if n >= best[0] else LAYOUT_GATHER)


def test_bad_schedule_rejected():
    with Context() as ctx:
        for bad in [(0, 10), (0, 5000), (-1, 2304)]:
            with pytest.raises(RuntimeError):
                ctx.set_schedule(*bad)
        assert ctx.schedule() == (0, 2304)
