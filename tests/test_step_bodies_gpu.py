"""Per-body stepping (titan_step_bodies) against the oracle.

Bodies advanced by their own step counts must match the oracle run for
exactly those counts, bit for bit (passive springs), whether several bodies
share each launch or one body is launched over its own tiles or chunks;
bodies that are not listed must not move; a later whole-context step must
continue every body from its own state; the acceleration read back is the
one of the body's last step; each body's clock counts its own steps."""
from __future__ import annotations

import numpy as np
import pytest

from titan_amd import LAYOUT_GATHER, LAYOUT_RESIDENT, LAYOUT_TILED, TITAN_E_NOMEM, Context, TitanError, scenes
from helpers import to_oracle

pytestmark = pytest.mark.gpu


def _bodies():
    a = scenes.readme_lattice(2.45)
    b = scenes.lattice_body((9.0, 0.0, 3.0), (2.0, 3.0, 2.0), (4, 6, 5), damping=0.1)
    c = scenes.lattice_body((-9.0, 0.0, 2.2), (4.0, 4.0, 4.0), (9, 9, 9), damping=0.2)  # 729 masses: 3 tiles
    return scenes.Scene(bodies=[a.bodies[0], b, c], planes=a.planes, dt=a.dt)


def _oracle(scene, i, n):
    o = to_oracle(scenes.Scene(bodies=[scene.bodies[i]], planes=scene.planes, dt=scene.dt))
    o.run(n)
    return o


def _check(ctx, bid, scene, i, n):
    got = ctx.download(bid, acc=True, T=True)
    o = _oracle(scene, i, n)
    np.testing.assert_array_equal(got["pos"], o.pos)
    np.testing.assert_array_equal(got["vel"], o.vel)
    if n > 0:
        np.testing.assert_array_equal(got["acc"], o.acc)
    T, k = ctx.body_time(bid)
    t = 0.0
    for _ in range(n):
        t += scene.dt
    assert k == n and T == t


MIXED = [LAYOUT_RESIDENT, LAYOUT_GATHER, LAYOUT_TILED]


@pytest.mark.parametrize("layout", [0, LAYOUT_GATHER, LAYOUT_TILED, MIXED],
                         ids=["auto", "gather", "tiled", "mixed"])
def test_step_bodies_own_counts_then_all(layout):
    sc = _bodies()
    print(f"case step_bodies layout {layout}: masses {[len(b['pos']) for b in sc.bodies]}", flush=True)
    with Context() as ctx:
        ids = sc.load(ctx, layout=layout)
        auto = [LAYOUT_RESIDENT, LAYOUT_RESIDENT, LAYOUT_TILED]  # the default schedule: <= 256 masses resident
        want = MIXED if layout == MIXED else auto if layout == 0 else [layout] * 3
        assert [ctx.body_layout(i) for i in ids] == want
        # several bodies per launch, uneven budgets; body 1 listed with 0
        ctx.step_bodies([ids[0], ids[1], ids[2]], [7, 0, 13])
        for i, n in zip(range(3), (7, 0, 13)):
            _check(ctx, ids[i], sc, i, n)
        # one body per launch (odd and even counts), the others stay put
        ctx.step_bodies([ids[1]], [5])
        ctx.step_bodies([ids[2]], [4])
        ctx.step_bodies([ids[1]], [64 + 3])
        for i, n in zip(range(3), (7, 72, 17)):
            _check(ctx, ids[i], sc, i, n)
        # the whole context: every body continues from its own state
        ctx.step(70)
        for i, n in zip(range(3), (77, 142, 87)):
            _check(ctx, ids[i], sc, i, n)
        assert ctx.time() == _clock(sc.dt, 70)  # the context clock counts whole-context steps


def _clock(dt, n):
    t = 0.0
    for _ in range(n):
        t += dt
    return t


def test_step_bodies_rejects_bad_input():
    sc = _bodies()
    with Context() as ctx:
        ids = sc.load(ctx)
        with pytest.raises(RuntimeError):
            ctx.step_bodies([ids[0], ids[0]], [1, 2])
        with pytest.raises(RuntimeError):
            ctx.step_bodies([ids[0]], [-1])
        with pytest.raises(RuntimeError):
            ctx.step_bodies([99], [1])
        ctx.step_bodies([ids[0]], [2])
        _check(ctx, ids[0], sc, 0, 2)


@pytest.mark.parametrize("layout", [LAYOUT_TILED, LAYOUT_GATHER])
def test_reserve_then_create_grows_nothing(layout):
    # titan_reserve grows the arenas once; bodies within the reservation are
    # then created without allocating (the device bytes stay put) and step
    # exactly as in a context that grew on demand
    sc = _bodies()
    masses = sum(len(b["pos"]) for b in sc.bodies)
    springs = sum(len(b["left"]) for b in sc.bodies)
    print(f"case reserve {masses} masses {springs} springs layout {layout}", flush=True)
    with Context() as ctx:
        sc.load_globals(ctx)
        ctx.reserve(masses, springs, 0, len(sc.bodies))
        before = ctx.mem_used()
        ids = [ctx.add_body(**b, layout=layout) for b in sc.bodies]
        assert ctx.mem_used() == before
        ctx.step(70)
        for i, bid in enumerate(ids):
            _check(ctx, bid, sc, i, 70)
    with Context() as ctx:
        with pytest.raises(Exception):
            ctx.reserve(-1, 0, 0, 1)
        with pytest.raises(Exception):
            ctx.reserve(10, 5, 6, 1)  # more actuated springs than springs
    # beyond the HBM budget: refused, and nothing grows
    with Context(hbm_budget=64 << 20) as ctx:
        before = ctx.mem_used()
        with pytest.raises(TitanError) as err:
            ctx.reserve(4_000_000, 30_000_000, 1_000_000, 100)
        assert err.value.code == TITAN_E_NOMEM
        assert ctx.mem_used() == before
        ctx.reserve(masses, springs, 0, len(sc.bodies))  # a reservation that fits
        ids = [ctx.add_body(**b, layout=layout) for b in sc.bodies]
