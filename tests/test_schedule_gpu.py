"""Schedules and the autotuner (titan_set_schedule, titan_autotune).

A schedule only changes how bodies are laid out on the device, so every
schedule must give bit-identical states; the autotuner must return one of its
candidates, make it the context's schedule, and bodies created afterwards must
follow it; bad schedules are rejected."""
from __future__ import annotations

import numpy as np
import pytest

from titan_amd import AUTOTUNE_CANDIDATES, LAYOUT_GATHER, LAYOUT_RESIDENT, LAYOUT_TILED, Context, scenes

pytestmark = pytest.mark.gpu


def _robots(n=12, seed=3):
    return scenes.robot_population(n, side_lo=3, side_hi=12, seed=seed)


def _layout(n, sched):
    mn, _, rs = sched
    return LAYOUT_RESIDENT if n <= rs else LAYOUT_TILED if n >= mn else LAYOUT_GATHER


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
    ref, lay = _run(sc, (0, 2304, 0))
    assert all(l == LAYOUT_TILED for l in lay)
    n = [len(b["pos"]) for b in sc.bodies]
    assert min(n) <= 1024 < max(n)
    for sched in [(10**9, 2304, 0), (256, 682, 0), (512, 2304, 0), (0, 682, 0), (0, 64, 0),
                  (0, 2304, 1024), (10**9, 682, 400), (0, 2304, 1), (0, 2304, 32767)]:
        st, lay = _run(sc, sched)
        assert lay == [_layout(k, sched) for k in n]
        for a, b in zip(st, ref):
            np.testing.assert_array_equal(a["pos"], b["pos"])
            np.testing.assert_array_equal(a["vel"], b["vel"])


def test_autotune_picks_a_candidate_and_applies_it():
    sc = _robots(8, seed=5)
    with Context() as ctx:
        ctx.set_dt(sc.dt)
        ctx.set_global_acc(sc.g)
        best, rates = ctx.autotune(sc.bodies, replicas=2, steps=4)
        assert len(rates) == len(AUTOTUNE_CANDIDATES) == 48
        assert all(r > 0 for _, r in rates)
        won = max(rates, key=lambda cr: cr[1])[0]
        want = (2**63 - 1 if won[0] < 0 else won[0], won[1], won[2])
        assert best == want and ctx.schedule() == best
        for b in sc.bodies[:3]:
            bid = ctx.add_body(**b)
            assert ctx.body_layout(bid) == _layout(len(b["pos"]), best)


def test_bad_schedule_rejected():
    with Context() as ctx:
        for bad in [(0, 10), (0, 5000), (-1, 2304), (0, 2304, 32768), (0, 2304, -1)]:
            with pytest.raises(RuntimeError):
                ctx.set_schedule(*bad)
        assert ctx.schedule() == (0, 2304, 256)


def test_partial_schedule_keeps_the_other_fields():
    with Context() as ctx:
        ctx.set_schedule(tile_rows=682)
        assert ctx.schedule() == (0, 682, 256)
        ctx.set_schedule(resident_max_masses=0)
        assert ctx.schedule() == (0, 682, 0)
        ctx.set_schedule(512)
        assert ctx.schedule() == (512, 682, 0)
