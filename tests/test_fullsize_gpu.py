"""Parity with the oracle at the benchmark workloads' full sizes.

The other GPU tests compare the HIP path with the oracle on scenes the
oracle finishes in a fraction of a second; these run the benchmark's own
scenes at their real shapes:

* the 100^3 lattice (1,000,000 masses, 12,731,796 springs) for the 100-step
  acceptance gate and on to 1000 steps, in the tiled and gather layouts (a
  launch this large is issued directly, not graph-captured): bit for bit;
  and a 60^3 lattice, just under the size whose launches are still
  captured in 64-step graphs, with graphs and with them turned off;
* the 4096-robot population (2.53M masses, 26.2M springs, 7.87M actuated)
  for one 100-step rollout under the default schedule and under the
  all-tiled one: within 1e-10 normwise relative error per body against the
  oracle (actuated springs take the device's sin), identical between the
  schedules, and bit for bit on a passive copy of the population;
* the 2M-mass kNN-12 cloud with 1000 balls and a ground plane for a few
  steps, in the tiled and gather layouts: bit for bit.
"""
from __future__ import annotations

import os
import sys

import numpy as np
import pytest

from titan_amd import LAYOUT_GATHER, LAYOUT_TILED, Context, scenes
from helpers import normwise_rel, to_oracle

pytestmark = pytest.mark.gpu

TOL_ACT = 1e-10


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _assert_same(got, o):
    np.testing.assert_array_equal(got["pos"], o.pos)
    np.testing.assert_array_equal(got["vel"], o.vel)


@pytest.fixture(scope="module")
def lattice100():
    sc = scenes.big_lattice(100, seed=0, dt=1e-5)
    o = to_oracle(sc)
    o.run(100)
    at100 = (o.pos.copy(), o.vel.copy())
    return sc, o, at100


def test_million_mass_lattice_100_and_1000_steps(lattice100):
    sc, o, at100 = lattice100
    _say(f"case 100^3 lattice {sc.n_masses} masses {sc.n_springs} springs: 100 then 1000 steps, tiled and gather")
    ctxs = {}
    for name, lay in (("tiled", LAYOUT_TILED), ("gather", LAYOUT_GATHER)):
        ctx = Context()
        (bid,) = sc.load(ctx, layout=lay)
        ctx.step(100)  # 100 direct launches (3907 tiles: above the graph limit)
        got = ctx.download(bid)
        np.testing.assert_array_equal(got["pos"], at100[0], err_msg=name)
        np.testing.assert_array_equal(got["vel"], at100[1], err_msg=name)
        ctxs[name] = (ctx, bid)
        _say(f"  {name}: 100 steps bit-identical")
    o.run(900)
    for name, (ctx, bid) in ctxs.items():
        ctx.step(900)
        got = ctx.download(bid, acc=True)
        _assert_same(got, o)
        np.testing.assert_array_equal(got["acc"], o.acc)
        ctx.close()
        _say(f"  {name}: 1000 steps bit-identical")


@pytest.mark.parametrize("graphs", ["1", "0"])
def test_lattice_under_the_graph_limit(graphs, monkeypatch):
    # 216,000 masses: 844 tiles and 844 gather workgroups per step, so the
    # bulk of 200 steps replays captured 64-step graphs (or, with
    # TITAN_GRAPHS=0, runs as direct launches)
    sc = scenes.big_lattice(60, seed=3, dt=1e-5)
    _say(f"case 60^3 lattice {sc.n_masses} masses, 200 steps, TITAN_GRAPHS={graphs}, tiled and gather")
    monkeypatch.setenv("TITAN_GRAPHS", graphs)
    o = to_oracle(sc)
    o.run(200)
    for lay in (LAYOUT_TILED, LAYOUT_GATHER):
        with Context() as ctx:
            (bid,) = sc.load(ctx, layout=lay)
            ctx.step(200)
            _assert_same(ctx.download(bid), o)


def _per_body(sc, cat):
    """Split concatenated (N, 3) arrays back into the scene's bodies."""
    out, at = [], 0
    for b in sc.bodies:
        n = len(b["pos"])
        out.append(cat[at:at + n])
        at += n
    return out


def _run_population(sc, sched, steps):
    with Context() as ctx:
        ctx.set_dt(sc.dt)
        ctx.set_global_acc(sc.g)
        if sched is not None:
            ctx.set_schedule(*sched)
        ids = [ctx.add_body(**b) for b in sc.bodies]
        ctx.step(steps)
        st = [ctx.download(i) for i in ids]
        kernel = ctx.kernel_name()
    return st, kernel


def test_robot_population_4096_rollout():
    sc = scenes.robot_population(4096, seed=0)
    _say(f"case 4096 robots: {sc.n_masses} masses, {sc.n_springs} springs, {sc.n_actuated} actuated, 100 steps")
    o = to_oracle(sc)
    o.run(100)
    want_pos, want_vel = _per_body(sc, o.pos), _per_body(sc, o.vel)
    runs = {}
    for name, sched in (("default", None), ("all tiled", (0, 2304, 0))):
        st, kernel = _run_population(sc, sched, 100)
        worst = 0.0
        for q, s in enumerate(st):
            worst = max(worst, normwise_rel(s["pos"], want_pos[q]), normwise_rel(s["vel"], want_vel[q]))
        _say(f"  {name} ({kernel}): worst per-body normwise relative error {worst:.3e}")
        assert worst <= TOL_ACT
        runs[name] = st
    for a, b in zip(runs["default"], runs["all tiled"]):
        np.testing.assert_array_equal(a["pos"], b["pos"])
        np.testing.assert_array_equal(a["vel"], b["vel"])


def test_robot_population_4096_passive_bitexact():
    sc = scenes.robot_population(4096, seed=0)
    for b in sc.bodies:
        for key in ("type", "omega", "act_a", "act_phi"):
            b.pop(key)
    _say(f"case 4096 robots, passive copy: {sc.n_masses} masses, 100 steps")
    o = to_oracle(sc)
    o.run(100)
    st, kernel = _run_population(sc, None, 100)
    got = {k: np.concatenate([s[k] for s in st]) for k in ("pos", "vel")}
    _assert_same(got, o)
    _say(f"  {kernel}: bit-identical")


def test_knn_cloud_2m_with_1000_balls():
    sc = scenes.knn_scene(2_000_000, n_balls=1000, seed=0)
    steps = 6
    _say(f"case kNN cloud {sc.n_masses} masses, {sc.n_springs} springs, 1000 balls, {steps} steps, tiled and gather")
    o = to_oracle(sc)
    o.run(steps)
    for name, lay in (("tiled", LAYOUT_TILED), ("gather", LAYOUT_GATHER)):
        with Context() as ctx:
            (bid,) = sc.load(ctx, layout=lay)
            ctx.step(steps)
            got = ctx.download(bid)
        _assert_same(got, o)
        _say(f"  {name}: bit-identical")
    # the balls act: some masses start inside one
    pos = sc.bodies[0]["pos"]
    inside = 0
    for cx, cy, cz, r in sc.balls[:50]:
        inside += int(np.count_nonzero(np.sum((pos - (cx, cy, cz)) ** 2, axis=1) <= r * r))
    assert inside > 0
