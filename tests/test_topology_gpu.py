"""Topology edits mid-run (titan_body_remove; the facade's deleteSpring,
deleteMass and deleteContainer after start, src/sim.cu:352-563).

The engine removes springs and masses from a live body in place; the oracle
gets the same edit at the same step (a removed spring's endpoints become -1,
which the reference skips, src/sim.cu:1052; a removed mass becomes fixed and
its springs are removed, as the reference stops updating it and skips the
springs of an invalidated mass).  Passive networks must stay bit-identical
to the oracle in every layout, before and after the edit, and after a
second edit; actuated ones within 1e-10."""
from __future__ import annotations

import sys

import numpy as np
import pytest

from titan_amd import (LAYOUT_GATHER, LAYOUT_RESIDENT, LAYOUT_TILED, TITAN_E_INVALID, TITAN_E_NOTFOUND, Context,
                       TitanError, build, scenes)
from helpers import normwise_rel, to_oracle

pytestmark = pytest.mark.gpu
LAYOUTS3 = pytest.mark.parametrize("layout", [LAYOUT_GATHER, LAYOUT_TILED, LAYOUT_RESIDENT],
                                   ids=["gather", "tiled", "resident"])


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _oracle_remove(o, masses, springs):
    """The same edit on the oracle scene (arrays it reads at every step)."""
    for s in springs:
        o.left[s] = -1
        o.right[s] = -1
    if len(masses):
        if o.fixed is None:
            o.fixed = np.zeros(o.n, np.uint8)
        ms = np.asarray(masses)
        o.fixed[ms] = 1
        touch = np.isin(o.left, ms) | np.isin(o.right, ms)
        o.left[touch] = -1
        o.right[touch] = -1


def _scene(n, per_k):
    rng = np.random.default_rng(7)
    b = scenes.lattice_body((0.0, 0.0, 0.45), (1.0, 1.0, 1.0), n, damping=0.1)
    if per_k:
        b["k"] = rng.uniform(5e3, 2e4, len(b["left"]))
    return scenes.Scene(bodies=[b], planes=[(0.0, 0.0, 1.0, 0.0, 0.8, 1.0)], dt=1e-4)


def _check(ctx, bid, o, exact=True):
    got = ctx.download(bid, acc=True, T=True)
    if exact:
        np.testing.assert_array_equal(got["pos"], o.pos)
        np.testing.assert_array_equal(got["vel"], o.vel)
        np.testing.assert_array_equal(got["T"], o.T)
    else:
        assert normwise_rel(got["pos"], o.pos) <= 1e-10
        assert normwise_rel(got["vel"], o.vel) <= 1e-10


@LAYOUTS3
@pytest.mark.parametrize("per_k", [False, True], ids=["shared_k", "per_spring_k"])
def test_remove_springs_and_masses_mid_run(layout, per_k):
    sc = _scene(9, per_k)
    S = sc.n_springs
    rng = np.random.default_rng(11)
    springs1 = rng.choice(S, size=S // 20, replace=False)
    masses1 = np.array([40, 364, 700])
    springs2 = np.concatenate([springs1[:5], rng.choice(S, size=30, replace=False)])  # some already removed
    _say(f"case remove mid-run: {sc.n_masses} masses, {S} springs, layout {layout}, per-spring k {per_k}")
    o = to_oracle(sc)
    with Context() as ctx:
        (bid,) = sc.load(ctx, layout=layout)
        assert ctx.body_layout(bid) == layout
        ctx.step(60)
        o.run(60)
        _check(ctx, bid, o)
        ctx.remove(bid, masses=masses1, springs=springs1)
        _oracle_remove(o, masses1, springs1)
        ctx.step(90)  # graph-sized: one 64-step replay
        o.run(90)
        _check(ctx, bid, o)
        ctx.remove(bid, masses=[364, 5], springs=springs2)
        _oracle_remove(o, [364, 5], springs2)
        ctx.step(40)
        o.run(40)
        _check(ctx, bid, o)
        frozen = ctx.download(bid)["pos"][masses1]
        ctx.step(3)
        np.testing.assert_array_equal(ctx.download(bid)["pos"][masses1], frozen)


def test_remove_from_a_split_form_body_and_an_actuated_one():
    # the README lattice is resident in the split form by default: a removal
    # makes it walk its incidences instead, with the same results
    sc = scenes.readme_lattice(2.45)
    _say("case remove from the README lattice (resident, split form)")
    o = to_oracle(sc)
    with Context() as ctx:
        (bid,) = sc.load(ctx)
        assert ctx.body_layout(bid) == LAYOUT_RESIDENT
        ctx.step(30)
        o.run(30)
        ctx.remove(bid, masses=[62], springs=[0, 1, 2, 500, 1035])
        _oracle_remove(o, [62], [0, 1, 2, 500, 1035])
        ctx.step(70)
        o.run(70)
        _check(ctx, bid, o)
    rob = scenes.robot_population(2, side_lo=7, side_hi=8, seed=3)
    _say(f"case remove from actuated robots: {rob.n_masses} masses, {rob.n_actuated} actuated")
    o = to_oracle(rob)
    with Context() as ctx:
        ctx.set_dt(rob.dt)
        ids = [ctx.add_body(**b, layout=LAYOUT_TILED) for b in rob.bodies]
        ctx.step(50)
        o.run(50)
        n0 = len(rob.bodies[0]["pos"])
        s0 = len(rob.bodies[0]["left"])
        ctx.remove(ids[0], masses=[3], springs=np.arange(0, s0, 11))
        ctx.remove(ids[1], springs=np.arange(5, len(rob.bodies[1]["left"]), 13))
        _oracle_remove(o, [3], np.arange(0, s0, 11))
        _oracle_remove(o, [], s0 + np.arange(5, len(rob.bodies[1]["left"]), 13))
        ctx.step(80)
        o.run(80)
        got = np.concatenate([ctx.download(i)["pos"] for i in ids])
        assert normwise_rel(got, o.pos) <= 1e-10
        assert got.shape[0] == n0 + len(rob.bodies[1]["pos"])


def test_remove_rejects_bad_indices():
    sc = _scene(4, False)
    with Context() as ctx:
        (bid,) = sc.load(ctx)
        for kw in (dict(masses=[64]), dict(masses=[-1]), dict(springs=[sc.n_springs])):
            with pytest.raises(TitanError) as e:
                ctx.remove(bid, **kw)
            assert e.value.code == TITAN_E_INVALID
        with pytest.raises(TitanError) as e:
            ctx.remove(bid + 1, springs=[0])
        assert e.value.code == TITAN_E_NOTFOUND


def test_facade_deletions_mid_run_match_the_oracle():
    # tests/cpp/facade_tests.cpp delete_parity: deleteSpring, deleteMass and
    # deleteContainer after 60 steps, then 90 more
    res = __import__("subprocess").run([str(build.FACADE_TESTS_BIN), "delete_parity"], capture_output=True,
                                       text=True, timeout=240)
    assert res.returncode == 0, res.stdout[-2000:] + res.stderr[-2000:]
    T, rows = None, []
    for line in res.stdout.splitlines():
        if line.startswith("T "):
            T = float.fromhex(line.split()[1])
        elif line.startswith("M "):
            rows.append([float.fromhex(x) for x in line.split()[1:]])
    got = np.array(rows)
    a = scenes.lattice_body((0.0, 0.0, 0.45), (1.0, 1.0, 1.0), 10, damping=0.1)
    b = scenes.lattice_body((3.0, 0.0, 0.45), (1.0, 1.0, 1.0), 3, damping=0.1)
    sc = scenes.Scene(bodies=[a, b], planes=[(0.0, 0.0, 1.0, 0.0, 0.8, 1.0)], dt=1e-4)
    o = to_oracle(sc)
    o.run(60)
    na, sa = len(a["pos"]), len(a["left"])
    springs = list(range(0, sa, 37)) + list(range(sa, sc.n_springs))
    masses = [123, 456] + list(range(na, sc.n_masses))
    _oracle_remove(o, masses, springs)
    o.run(90)
    live = np.setdiff1d(np.arange(na), [123, 456])
    assert got.shape == (len(live), 7)
    assert T == o.t
    np.testing.assert_array_equal(got[:, 0:3], o.pos[live])
    np.testing.assert_array_equal(got[:, 3:6], o.vel[live])
    np.testing.assert_array_equal(got[:, 6], o.T[live])
