"""The device-side layout builder (builder.hip) against the host builders
(tiles.cpp): for the same body, every item of the device layout (incidence
names, rest lengths, constants, degrees, flags, state, masses, tiles, halos,
chunk table, mass order) must be the same bytes whichever built it, and the
two bodies must then step identically.  TITAN_DEVICE_BUILD, read when a
context is created, sets the size from which tiled bodies are built on the
device (0: never)."""
from __future__ import annotations

import sys
import time

import numpy as np
import pytest

from titan_amd import LAYOUT_TILED, Context, scenes
from helpers import normwise_rel, to_oracle

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _ctx(monkeypatch, device_min: int, sc=None, tile_rows=None) -> Context:
    monkeypatch.setenv("TITAN_DEVICE_BUILD", str(device_min))
    ctx = Context()
    monkeypatch.delenv("TITAN_DEVICE_BUILD")
    if sc is not None:
        ctx.set_dt(sc.dt)
        ctx.set_global_acc(sc.g)
        for p in sc.planes:
            ctx.add_plane(p[:3], p[3], p[4], p[5])
    if tile_rows is not None:
        ctx.set_schedule(tile_rows=tile_rows)
    return ctx


def _compare(monkeypatch, body, sc=None, steps=0, tile_rows=None):
    """Build `body` on the device and on the host; compare every layout item
    byte for byte, then `steps` steps of each.  Returns the device context's
    download after the steps."""
    dev = _ctx(monkeypatch, 1, sc, tile_rows)
    host = _ctx(monkeypatch, 0, sc, tile_rows)
    try:
        t0 = time.perf_counter()
        bd = dev.add_body(**body, layout=LAYOUT_TILED)
        t1 = time.perf_counter()
        bh = host.add_body(**body, layout=LAYOUT_TILED)
        t2 = time.perf_counter()
        _say(f"  built on the device in {1e3 * (t1 - t0):.1f} ms, on the host in {1e3 * (t2 - t1):.1f} ms")
        for item in Context.DEBUG_ITEMS:
            a, b = dev.layout_bytes(bd, item), host.layout_bytes(bh, item)
            assert len(a) == len(b), item
            if a != b:
                x = np.frombuffer(a, np.uint8) != np.frombuffer(b, np.uint8)
                raise AssertionError(f"{item}: {int(x.sum())} bytes differ, first at {int(np.argmax(x))}")
        out = None
        if steps:
            dev.step(steps)
            host.step(steps)
            out = dev.download(bd)
            other = host.download(bh)
            np.testing.assert_array_equal(out["pos"], other["pos"])
            np.testing.assert_array_equal(out["vel"], other["vel"])
        return out
    finally:
        dev.close()
        host.close()


def test_lattice_same_bytes_and_steps(monkeypatch):
    sc = scenes.big_lattice(20, seed=3, dt=1e-5)
    b = sc.bodies[0]
    b["left"] = b["left"].copy()
    b["left"][::97] = -1  # springs without a mass do nothing and are not listed
    _say(f"case device build: lattice 20^3, {len(b['left'])} springs, some without an endpoint")
    got = _compare(monkeypatch, b, sc, steps=40)
    o = to_oracle(sc)
    o.run(40)
    np.testing.assert_array_equal(got["pos"], o.pos)
    np.testing.assert_array_equal(got["vel"], o.vel)


def test_knn_cloud_per_spring_constants_same_bytes(monkeypatch):
    rng = np.random.default_rng(5)
    b = scenes.knn_cloud(30000, box=3.0, seed=9)
    b["damping"] = rng.uniform(0.01, 0.1, len(b["left"]))
    sc = scenes.Scene(bodies=[b], planes=[(0.0, 0.0, 1.0, 0.0, 0.8, 1.0)], dt=1e-5)
    _say(f"case device build: kNN cloud 30000 masses, per-spring k and damping")
    _compare(monkeypatch, b, sc, steps=20)


def test_actuated_lattice_with_mass_fields_same_bytes(monkeypatch):
    rng = np.random.default_rng(8)
    b = scenes.lattice_body((0.0, 0.0, 0.6), (1.0, 1.0, 1.0), 25, damping=0.1)
    scenes.actuate(b, 0.3, rng)
    n = len(b["pos"])
    b["m"] = rng.uniform(0.05, 0.15, n)
    b["vel"] = rng.normal(0.0, 0.01, (n, 3))
    b["T"] = rng.uniform(0.0, 1e-3, n)
    ext = np.zeros((n, 3))
    ext[::50] = rng.normal(0.0, 1.0, (len(ext[::50]), 3))
    b["ext"] = ext
    b["drag"] = np.where(np.arange(n) % 37 == 0, 2, 0).astype(np.int32)
    b["fixed"] = (np.arange(n) % 101 == 0).astype(np.uint8)
    b["lc_ptr"] = np.arange(n + 1, dtype=np.int64)  # one shared local plane
    b["lc_type"] = np.full(n, 1, np.int32)
    b["lc_vec"] = np.tile([0.0, 0.0, 1.0], (n, 1))
    b["lc_num"] = np.zeros(n)
    sc = scenes.Scene(bodies=[b], planes=[(0.0, 0.0, 1.0, 0.0, 0.8, 1.0)], dt=1e-4)
    _say(f"case device build: actuated lattice 25^3, masses, velocities, clocks, extern force, drag, fixed, local plane")
    got = _compare(monkeypatch, b, sc, steps=50)
    o = to_oracle(sc)
    o.run(50)
    assert normwise_rel(got["pos"], o.pos) <= 1e-10
    assert normwise_rel(got["vel"], o.vel) <= 1e-10


def test_ties_flat_sheets_self_springs_and_fallback(monkeypatch):
    # positions on a coarse grid (many equal Morton codes: ties keep the
    # caller's order), a flat sheet (one axis of zero extent), self-springs
    rng = np.random.default_rng(2)
    b = scenes.knn_cloud(20000, box=1.0, seed=4)
    b["pos"] = np.round(b["pos"] * 8.0) / 8.0
    b["left"] = np.concatenate([b["left"], [7, 7, 123]]).astype(np.int32)
    b["right"] = np.concatenate([b["right"], [7, 9, 123]]).astype(np.int32)
    for key in ("k", "rest", "damping"):
        b[key] = np.concatenate([b[key], rng.uniform(0.5, 1.0, 3)])
    _say("case device build: coincident positions, self-springs")
    _compare(monkeypatch, b)
    sheet = scenes.lattice_body((0.0, 0.0, 1.0), (4.0, 4.0, 0.0), (150, 150, 1), damping=0.1)
    _say("case device build: a flat 150 x 150 sheet")
    _compare(monkeypatch, sheet, steps=10)
    # a row budget no four-chunk tile fits: the host's greedy decides
    _say("case device build: row budget below a four-chunk tile (host fallback)")
    _compare(monkeypatch, scenes.lattice_body((0.0, 0.0, 1.0), (1.0, 1.0, 1.0), 26), tile_rows=300)
