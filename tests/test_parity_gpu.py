"""Parity of the HIP step (through the C ABI) with the CPU oracle.

Bar: passive spring networks must be bit-identical to the oracle (the kernel
keeps the reference's FP64 operation order, is built with -ffp-contract=off
and sums forces in ascending spring index, as the oracle does).  Actuated
springs call the device sin, so they are held to a normwise relative error
of 1e-10 after 100 steps (the acceptance tolerance of the path).
"""
from __future__ import annotations

import sys

import numpy as np
import pytest

from titan_amd import LAYOUT_GATHER, LAYOUT_RESIDENT, LAYOUT_TILED, Context, scenes
from helpers import normwise_rel, to_oracle

pytestmark = pytest.mark.gpu
LAYOUTS = pytest.mark.parametrize("layout", [LAYOUT_GATHER, LAYOUT_TILED], ids=["gather", "tiled"])
# the small bodies of these cases also run in the resident layout
LAYOUTS3 = pytest.mark.parametrize("layout", [LAYOUT_GATHER, LAYOUT_TILED, LAYOUT_RESIDENT],
                                   ids=["gather", "tiled", "resident"])

TOL_ACT = 1e-10


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def run_engine(scene, n_steps, ctx=None, chunks=None, layout=0):
    own = ctx is None
    ctx = ctx or Context()
    ids = scene.load(ctx, layout=layout)
    if layout:
        assert all(ctx.body_layout(i) == layout for i in ids)
    for n in (chunks or [n_steps]):
        ctx.step(n)
    out = [ctx.download(i, pos=True, vel=True, acc=True, T=True) for i in ids]
    t = ctx.time()
    if own:
        ctx.close()
    cat = {k: np.concatenate([o[k] for o in out]) for k in out[0]}
    return cat, t


def run_oracle(scene, n_steps):
    o = to_oracle(scene)
    o.run(n_steps)
    return o


def assert_bitexact(eng, orc):
    np.testing.assert_array_equal(eng["pos"], orc.pos)
    np.testing.assert_array_equal(eng["vel"], orc.vel)


@LAYOUTS3
def test_readme_lattice_free_fall_bitexact(layout):
    sc = scenes.readme_lattice(10.0)
    _say(f"case readme lattice cz=10, {sc.n_masses} masses, {sc.n_springs} springs, 100 steps")
    eng, t = run_engine(sc, 100, layout=layout)
    orc = run_oracle(sc, 100)
    assert_bitexact(eng, orc)
    np.testing.assert_array_equal(eng["acc"], orc.acc)
    np.testing.assert_array_equal(eng["T"], orc.T)
    assert t == orc.t


@LAYOUTS3
def test_readme_lattice_contact_bitexact(layout):
    sc = scenes.readme_lattice(2.45)
    _say("case readme lattice cz=2.45 (contact + friction from step 0), 1000 steps")
    eng, _ = run_engine(sc, 1000, layout=layout)
    orc = run_oracle(sc, 1000)
    assert_bitexact(eng, orc)
    np.testing.assert_array_equal(eng["acc"], orc.acc)
    assert np.min(sc.bodies[0]["pos"][:, 2]) < 0  # in contact from the first step


@LAYOUTS3
def test_graph_path_matches_direct_launches(layout):
    sc = scenes.readme_lattice(2.45)
    _say("case graph replay vs single launches, 300 steps")
    a, _ = run_engine(sc, 300, layout=layout)                      # mostly hipGraph replays
    b, _ = run_engine(sc, 300, chunks=[1] * 300, layout=layout)    # one launch per call
    np.testing.assert_array_equal(a["pos"], b["pos"])
    np.testing.assert_array_equal(a["vel"], b["vel"])


@LAYOUTS3
def test_actuated_reference_kinds_and_general_law(layout):
    sc = scenes.readme_lattice(2.45)
    b = sc.bodies[0]
    s = len(b["left"])
    rng = np.random.default_rng(1)
    typ = rng.integers(0, 5, s).astype(np.int32)
    b["type"] = typ
    b["omega"] = rng.uniform(1, 20, s)
    b["act_a"] = rng.uniform(0, 0.2, s)
    b["act_phi"] = rng.uniform(0, 2 * np.pi, s)
    _say("case actuated springs (kinds 0-4), 100 steps")
    eng, _ = run_engine(sc, 100, layout=layout)
    orc = run_oracle(sc, 100)
    assert normwise_rel(eng["pos"], orc.pos) <= TOL_ACT
    assert normwise_rel(eng["vel"], orc.vel) <= TOL_ACT


@LAYOUTS3
def test_fixed_ext_drag_balls_per_spring_k(layout):
    sc = scenes.readme_lattice(2.45)
    b = sc.bodies[0]
    n, s = len(b["pos"]), len(b["left"])
    rng = np.random.default_rng(2)
    b["fixed"] = (np.arange(n) % 17 == 0).astype(np.uint8)
    b["ext"] = rng.normal(0, 0.5, (n, 3)) * (np.arange(n) % 3 == 0)[:, None]
    b["drag"] = (np.arange(n) % 5).astype(np.int32)
    b["k"] = rng.uniform(5e3, 2e4, s)
    b["damping"] = rng.uniform(0, 0.3, s)
    b["vel"] = rng.normal(0, 0.5, (n, 3))
    sc.balls = [(0.0, 0.0, 1.0, 1.2), (1.0, 1.0, 4.0, 0.7)]
    sc.planes.append((0.0, 0.6, 0.8, -1.0, 0.3, 0.5))
    _say("case fixed + extern + drag + balls + 2 planes + per-spring k/damping, 500 steps")
    eng, _ = run_engine(sc, 500, layout=layout)
    orc = run_oracle(sc, 500)
    assert_bitexact(eng, orc)
    np.testing.assert_array_equal(eng["T"], orc.T)


@LAYOUTS3
def test_local_constraints_shared_and_per_mass(layout):
    sc = scenes.readme_lattice(2.0)
    b = sc.bodies[0]
    n = len(b["pos"])
    # body 0: every mass gets the same local list (Container::addConstraint)
    kinds = [1, 2, 0, 3, 1]
    vecs = [(0, 0, 1), (0.5, 0.5, 2.0), (0.6, 0.0, 0.8), (1.0, 0.0, 0.0), (0.0, 0.6, 0.8)]
    nums = [0.3, 0.8, 0.2, 0.1, -0.5]
    b["lc_ptr"] = np.arange(n + 1, dtype=np.int64) * len(kinds)
    b["lc_type"] = np.tile(kinds, n).astype(np.int32)
    b["lc_vec"] = np.tile(np.asarray(vecs, float), (n, 1))
    b["lc_num"] = np.tile(nums, n)
    # body 1: per-mass lists of varying length
    b2 = scenes.lattice_body((3.0, 0.0, 1.5), (2.0, 2.0, 2.0), 4, damping=0.05)
    n2 = len(b2["pos"])
    rng = np.random.default_rng(3)
    cnt = rng.integers(0, 4, n2)
    b2["lc_ptr"] = np.concatenate([[0], np.cumsum(cnt)]).astype(np.int64)
    tot = int(cnt.sum())
    b2["lc_type"] = rng.integers(0, 4, tot).astype(np.int32)
    v = rng.normal(size=(tot, 3))
    v /= np.sqrt((v * v).sum(1))[:, None]
    b2["lc_vec"] = v
    b2["lc_num"] = rng.uniform(-0.2, 0.5, tot)
    b2["vel"] = rng.normal(0, 0.3, (n2, 3))
    sc.bodies.append(b2)
    _say("case local constraints (shared list + per-mass lists), 2 bodies, 400 steps")
    eng, _ = run_engine(sc, 400, layout=layout)
    orc = run_oracle(sc, 400)
    assert_bitexact(eng, orc)


@LAYOUTS3
def test_bodies_alone_equal_bodies_together(layout):
    s1 = scenes.readme_lattice(2.45)
    s2 = scenes.readme_lattice(3.0)
    s2.bodies[0]["k"] = np.full(len(s2.bodies[0]["left"]), 7000.0)
    both = scenes.Scene(bodies=s1.bodies + s2.bodies, planes=s1.planes, dt=s1.dt)
    _say("case two bodies alone vs together, 200 steps")
    a, _ = run_engine(s1, 200, layout=layout)
    b, _ = run_engine(s2, 200, layout=layout)
    ab, _ = run_engine(both, 200, layout=layout)
    np.testing.assert_array_equal(ab["pos"], np.concatenate([a["pos"], b["pos"]]))
    np.testing.assert_array_equal(ab["vel"], np.concatenate([a["vel"], b["vel"]]))


def test_step_until_breakpoint_rule_and_time():
    sc = scenes.readme_lattice(2.45)
    with Context() as ctx:
        ids = sc.load(ctx)
        k = ctx.step_until(0.0105)
        t = ctx.time()
        # first step boundary with T >= 0.0105 under T += 1e-4 accumulation
        T, n = 0.0, 0
        while not (0.0105 <= T):
            T += 1e-4
            n += 1
        assert k == n and t == T
        got = ctx.download(ids[0])
    orc = run_oracle(sc, n)
    np.testing.assert_array_equal(got["pos"], orc.pos)


def test_upload_download_roundtrip_and_resume():
    sc = scenes.readme_lattice(2.45)
    o = to_oracle(sc)
    with Context() as ctx:
        (bid,) = sc.load(ctx)
        ctx.step(50)
        st = ctx.download(bid, acc=True, T=True)
        # teleport the body as rotate_unittest does via get/set
        newpos = st["pos"] + np.array([0.1, -0.2, 0.05])
        ctx.upload(bid, pos=newpos, vel=st["vel"] * 0.5)
        ctx.step(50)
        got = ctx.download(bid)
    o.run(50)
    for c in range(3):
        o.p[c][:] = o.p[c] + [0.1, -0.2, 0.05][c]
        o.v[c][:] = o.v[c] * 0.5
    o.run(50)
    np.testing.assert_array_equal(got["pos"], o.pos)
    np.testing.assert_array_equal(got["vel"], o.vel)


@LAYOUTS3
def test_mass_upload_changes_uniform_masses(layout):
    # a body whose masses all share m passes it by value to the tiled step;
    # uploading unequal masses (and later equal ones again) must be seen
    sc = scenes.readme_lattice(2.45)
    n = len(sc.bodies[0]["pos"])
    o = to_oracle(sc)
    m1 = 0.1 + 0.01 * (np.arange(n) % 7)
    m2 = np.full(n, 0.25)
    with Context() as ctx:
        (bid,) = sc.load(ctx, layout=layout)
        ctx.step(40)
        ctx.upload(bid, m=m1)
        ctx.step(40)
        ctx.upload(bid, m=m2)
        ctx.step(40)
        got = ctx.download(bid, acc=True)
    o.run(40)
    o.m = np.array(m1, dtype=np.float64)
    o.run(40)
    o.m = np.array(m2, dtype=np.float64)
    o.run(40)
    np.testing.assert_array_equal(got["pos"], o.pos)
    np.testing.assert_array_equal(got["vel"], o.vel)
    np.testing.assert_array_equal(got["acc"], o.acc)


@LAYOUTS
def test_million_mass_lattice_bitexact_few_steps(layout):
    sc = scenes.big_lattice(100, seed=0)
    _say(f"case 100^3 lattice {sc.n_masses} masses {sc.n_springs} springs, 5 steps")
    eng, _ = run_engine(sc, 5, layout=layout)
    o = to_oracle(sc)
    o.run(5)
    assert_bitexact(eng, o)


def test_destroy_and_reuse_slots():
    sc = scenes.readme_lattice(2.45)
    with Context() as ctx:
        ids = sc.load(ctx)
        extra = scenes.lattice_body((5, 5, 3), (1, 1, 1), 3, damping=0.1)
        e = ctx.add_body(**extra)
        ctx.step(10)
        ctx.destroy_body(e)
        e2 = ctx.add_body(**extra)
        ctx.step(20)
        got = ctx.download(ids[0])
        got2 = ctx.download(e2)
    o = to_oracle(sc)
    o.run(30)
    np.testing.assert_array_equal(got["pos"], o.pos)
    o2 = to_oracle(scenes.Scene(bodies=[extra], planes=sc.planes, dt=sc.dt))
    o2.run(20)
    np.testing.assert_array_equal(got2["pos"], o2.pos)


@LAYOUTS
def test_multi_tile_lattice_with_contact(layout):
    # 12^3 masses: several 256-mass tiles along a Morton curve, with halos
    sc = scenes.Scene(bodies=[scenes.lattice_body((0.0, 0.0, 0.55), (1.2, 1.2, 1.2), 12, damping=0.1)],
                      planes=[(0.0, 0.0, 1.0, 0.0, 0.8, 1.0)], dt=1e-4)
    b = sc.bodies[0]
    rng = np.random.default_rng(9)
    b["pos"] = b["pos"] + rng.normal(0, 1e-3, b["pos"].shape)
    b["vel"] = rng.normal(0, 0.2, b["pos"].shape)
    _say(f"case 12^3 lattice layout {layout}, 400 steps")
    eng, _ = run_engine(sc, 400, layout=layout)
    orc = run_oracle(sc, 400)
    assert_bitexact(eng, orc)
    np.testing.assert_array_equal(eng["acc"], orc.acc)
    np.testing.assert_array_equal(eng["T"], orc.T)


def test_mixed_layouts_in_one_context():
    s1 = scenes.readme_lattice(2.45)
    s2 = scenes.Scene(bodies=[scenes.lattice_body((3.0, 0.0, 0.5), (1.0, 1.0, 1.0), 9, damping=0.1)],
                      planes=s1.planes, dt=s1.dt)
    _say("case gather body + tiled body in one context, 300 steps")
    with Context() as ctx:
        a = s1.load(ctx, layout=LAYOUT_GATHER)[0]
        ctx.clear_planes_and_balls()
        b = s2.load(ctx, layout=LAYOUT_TILED)[0]
        ctx.step(300)
        ga, gb = ctx.download(a), ctx.download(b)
        assert "k_step_gather" in ctx.kernel_name() and "k_step_tiles" in ctx.kernel_name()
    oa, ob = run_oracle(s1, 300), run_oracle(s2, 300)
    np.testing.assert_array_equal(ga["pos"], oa.pos)
    np.testing.assert_array_equal(gb["pos"], ob.pos)


def test_auto_layout_by_size():
    small = scenes.readme_lattice(2.45)
    big = scenes.lattice_body((0.0, 0.0, 0.55), (1.2, 1.2, 1.2), 11, damping=0.1)  # 1331 masses
    with Context() as ctx:
        (a,) = small.load(ctx)
        assert ctx.body_layout(a) == LAYOUT_RESIDENT  # the default: at most 256 masses resident
        assert ctx.body_layout(ctx.add_body(**scenes.lattice_body((0, 0, 1), (1, 1, 1), 7))) == LAYOUT_TILED
        ctx.set_schedule(0, 2304, 1024)  # bodies of at most 1024 masses resident
        assert ctx.body_layout(ctx.add_body(**small.bodies[0])) == LAYOUT_RESIDENT
        assert ctx.body_layout(ctx.add_body(**big)) == LAYOUT_TILED
        assert ctx.body_layout(ctx.add_body(**big, layout=LAYOUT_RESIDENT)) == LAYOUT_RESIDENT
        huge = scenes.lattice_body((0.0, 0.0, 0.55), (1.2, 1.2, 1.2), (33, 33, 31))  # 33759 masses
        with pytest.raises(RuntimeError):
            ctx.add_body(**huge, layout=LAYOUT_RESIDENT)  # more than 32767 masses


def test_resident_bodies_in_lds_and_in_the_arena_bitexact():
    # a launch with few bodies: the 125-mass body steps in LDS (split form),
    # the 343-mass one in LDS (walk), the 1331- and 2197-mass ones in their
    # arena state buffers; then the same bodies among 300 others, where
    # every body steps in the arena
    rng = np.random.default_rng(12)
    bodies = []
    for q, side in enumerate([5, 7, 11, 13]):
        b = scenes.lattice_body((4.0 * q, 0.0, 0.6), (1.2, 1.2, 1.2), side, damping=0.1)
        b["pos"] = b["pos"] + rng.normal(0, 1e-3, b["pos"].shape)
        b["vel"] = rng.normal(0, 0.2, b["pos"].shape)
        bodies.append(b)
    sc = scenes.Scene(bodies=bodies, planes=[(0.0, 0.0, 1.0, 0.0, 0.8, 1.0)], dt=1e-4)
    filler = [scenes.lattice_body((2.0 * (q % 20), 8.0 + 2.0 * (q // 20), 0.6), (1.0, 1.0, 1.0), 3, damping=0.1)
              for q in range(300)]
    calls = [1, 64, 85]
    orc = []
    for q in range(4):
        o = to_oracle(scenes.Scene(bodies=[bodies[q]], planes=sc.planes, dt=sc.dt))
        o.run(sum(calls))
        orc.append(o)
    for extra in (0, 300):
        _say(f"case resident bodies 125..2197 masses with {extra} other bodies, calls {calls}")
        with Context() as ctx:
            ids = sc.load(ctx, layout=LAYOUT_RESIDENT)
            for f in filler[:extra]:
                ctx.add_body(**f, layout=LAYOUT_RESIDENT)
            for n in calls:
                ctx.step(n)
            got = [ctx.download(i, acc=True, T=True) for i in ids]
        for q in range(4):
            np.testing.assert_array_equal(got[q]["pos"], orc[q].pos)
            np.testing.assert_array_equal(got[q]["vel"], orc[q].vel)
            np.testing.assert_array_equal(got[q]["acc"], orc[q].acc)
            np.testing.assert_array_equal(got[q]["T"], orc[q].T)


def test_resident_beside_gather_and_tiled_bodies():
    # one body per layout in one context, stepped in calls of odd, even and
    # graph-sized (>= 64) counts: each must match its own oracle run
    rng = np.random.default_rng(11)
    bodies = []
    for q, side in enumerate([10, 6, 11]):  # 1000 (resident), 216 (gather), 1331 (tiled)
        b = scenes.lattice_body((4.0 * q, 0.0, 0.58), (1.2, 1.2, 1.2), side, damping=0.1)
        b["pos"] = b["pos"] + rng.normal(0, 1e-3, b["pos"].shape)
        b["vel"] = rng.normal(0, 0.2, b["pos"].shape)
        bodies.append(b)
    sc = scenes.Scene(bodies=bodies, planes=[(0.0, 0.0, 1.0, 0.0, 0.8, 1.0)], dt=1e-4)
    lays = [LAYOUT_RESIDENT, LAYOUT_GATHER, LAYOUT_TILED]
    calls = [1, 7, 64, 130, 2]
    _say(f"case resident + gather + tiled bodies, calls {calls}")
    with Context() as ctx:
        ids = sc.load(ctx, layout=lays)
        assert [ctx.body_layout(i) for i in ids] == lays
        assert "k_rollout_resident" in ctx.kernel_name()
        for n in calls:
            ctx.step(n)
        got = [ctx.download(i, acc=True, T=True) for i in ids]
    for q in range(3):
        o = to_oracle(scenes.Scene(bodies=[bodies[q]], planes=sc.planes, dt=sc.dt))
        o.run(sum(calls))
        np.testing.assert_array_equal(got[q]["pos"], o.pos)
        np.testing.assert_array_equal(got[q]["vel"], o.vel)
        np.testing.assert_array_equal(got[q]["acc"], o.acc)
        np.testing.assert_array_equal(got[q]["T"], o.T)


def test_resident_many_balls_and_actuation():
    # the ball grid and actuated springs inside the resident kernel
    rng = np.random.default_rng(8)
    b = scenes.lattice_body((0.0, 0.0, 1.2), (2.0, 2.0, 2.0), 10, damping=0.1)
    b = scenes.actuate(b, 0.3, rng)
    balls = [(*rng.uniform(-1.5, 1.5, 2), rng.uniform(0.0, 2.5), rng.uniform(0.05, 0.35)) for _ in range(200)]
    sc = scenes.Scene(bodies=[b], planes=[(0.0, 0.0, 1.0, 0.0, 0.8, 1.0)], balls=balls, dt=1e-4)
    _say("case resident body, 200 balls, 30% actuated, 150 steps")
    eng, _ = run_engine(sc, 150, layout=LAYOUT_RESIDENT)
    ref, _ = run_engine(sc, 150, layout=LAYOUT_GATHER)
    np.testing.assert_array_equal(eng["pos"], ref["pos"])  # same device sin on both paths
    np.testing.assert_array_equal(eng["vel"], ref["vel"])
    orc = run_oracle(sc, 150)
    assert normwise_rel(eng["pos"], orc.pos) <= TOL_ACT
    assert normwise_rel(eng["vel"], orc.vel) <= TOL_ACT


@LAYOUTS3
def test_edge_case_bodies_bitexact(layout):
    # one free mass; masses without springs; springs with a missing endpoint
    # (-1: skipped, as the reference skips a null endpoint, src/sim.cu:1052)
    # and duplicate springs between one pair; a body with every mass fixed
    rng = np.random.default_rng(31)
    one = dict(pos=np.array([[0.0, 0.0, 0.3]]), m=np.array([0.2]), vel=np.array([[0.1, -0.2, 0.0]]))
    loose = dict(pos=rng.uniform(-1.0, 1.0, (10, 3)) + [6.0, 0.0, 1.5], m=np.full(10, 0.1),
                 vel=rng.normal(0, 0.5, (10, 3)))
    b = scenes.lattice_body((3.0, 0.0, 0.6), (1.0, 1.0, 1.0), 4, damping=0.1)
    s = len(b["left"])
    dup = np.arange(0, s, 17)
    left, right = b["left"].copy(), b["right"].copy()
    left[::11] = -1
    right[5::13] = -1
    b["left"] = np.concatenate([left, b["left"][dup]])
    b["right"] = np.concatenate([right, b["right"][dup]])
    b["k"] = np.concatenate([b["k"], rng.uniform(5e3, 2e4, len(dup))])
    b["rest"] = np.concatenate([b["rest"], b["rest"][dup] * 1.01])
    b["damping"] = np.concatenate([b["damping"], np.full(len(dup), 0.2)])
    b["vel"] = rng.normal(0, 0.3, b["pos"].shape)
    fixed = scenes.lattice_body((-3.0, 0.0, 0.6), (1.0, 1.0, 1.0), 3, damping=0.1)
    fixed["fixed"] = np.ones(len(fixed["pos"]), np.uint8)
    sc = scenes.Scene(bodies=[one, loose, b, fixed], planes=[(0.0, 0.0, 1.0, 0.0, 0.8, 1.0)], dt=1e-4)
    _say(f"case edge-case bodies (1 mass, no springs, -1 endpoints, duplicates, all fixed), layout {layout}")
    with Context() as ctx:
        ids = sc.load(ctx, layout=layout)
        ctx.step(0)
        for n in (1, 70, 129):
            ctx.step(n)
        got = [ctx.download(i, acc=True, T=True) for i in ids]
    for q, g in enumerate(got):
        o = to_oracle(scenes.Scene(bodies=[sc.bodies[q]], planes=sc.planes, dt=sc.dt))
        o.run(200)
        np.testing.assert_array_equal(g["pos"], o.pos)
        np.testing.assert_array_equal(g["vel"], o.vel)
        np.testing.assert_array_equal(g["acc"], o.acc)
        np.testing.assert_array_equal(g["T"], o.T)
    np.testing.assert_array_equal(got[3]["pos"], fixed["pos"])  # fixed masses never move


def _small_bodies(rng, act=False):
    # resident bodies of one to four chunks (27, 64, 125 and 216 masses and a
    # 200-mass kNN cloud of irregular degree), k and damping shared or per
    # spring in every combination, jittered and moving, some masses fixed
    bodies = []
    for q, (side, pk, pd) in enumerate([(3, False, False), (4, True, False), (5, False, True), (6, True, True)]):
        b = scenes.lattice_body((3.0 * q, 0.0, 0.9), (1.5, 1.5, 1.5), side, damping=0.1)
        s, n = len(b["left"]), len(b["pos"])
        b["pos"] = b["pos"] + rng.normal(0, 1e-3, b["pos"].shape)
        b["vel"] = rng.normal(0, 0.3, b["pos"].shape)
        b["k"] = rng.uniform(5e3, 2e4, s) if pk else np.full(s, 1e4 + 500.0 * q)
        b["damping"] = rng.uniform(0.0, 0.3, s) if pd else np.full(s, 0.05 * q)
        b["fixed"] = (np.arange(n) % 23 == 5).astype(np.uint8)
        bodies.append(scenes.actuate(b, 0.3, rng) if act else b)
    c = scenes.knn_cloud(200, box=1.0, seed=5, z0=0.05)
    bodies.append(scenes.actuate(c, 0.3, rng) if act else c)
    return scenes.Scene(bodies=bodies, planes=[(0.0, 0.0, 1.0, 0.0, 0.8, 1.0)], dt=1e-4)


def test_resident_small_bodies_bitexact():
    # few bodies of at most four chunks: the resident launch computes each
    # spring's term once over the whole workgroup, then sums per mass
    sc = _small_bodies(np.random.default_rng(21))
    _say("case resident small bodies (27..216 masses, kNN 200), per-spring k/damping combos, 300 steps")
    with Context() as ctx:
        ids = sc.load(ctx, layout=LAYOUT_RESIDENT)
        for n in (1, 64, 235):
            ctx.step(n)
        got = [ctx.download(i, acc=True, T=True) for i in ids]
    for q, g in enumerate(got):
        o = to_oracle(scenes.Scene(bodies=[sc.bodies[q]], planes=sc.planes, dt=sc.dt))
        o.run(300)
        np.testing.assert_array_equal(g["pos"], o.pos)
        np.testing.assert_array_equal(g["vel"], o.vel)
        np.testing.assert_array_equal(g["acc"], o.acc)
        np.testing.assert_array_equal(g["T"], o.T)


def test_resident_small_bodies_actuated():
    sc = _small_bodies(np.random.default_rng(22), act=True)
    _say("case resident small bodies, 30% actuated, 150 steps")
    eng, _ = run_engine(sc, 150, layout=LAYOUT_RESIDENT)
    ref, _ = run_engine(sc, 150, layout=LAYOUT_GATHER)
    np.testing.assert_array_equal(eng["pos"], ref["pos"])  # same device sin on both paths
    np.testing.assert_array_equal(eng["vel"], ref["vel"])
    orc = run_oracle(sc, 150)
    assert normwise_rel(eng["pos"], orc.pos) <= TOL_ACT
    assert normwise_rel(eng["vel"], orc.vel) <= TOL_ACT


def test_resident_more_bodies_than_compute_units():
    # 520 small bodies, more than the device's compute units (256 on MI355X):
    # the launch then walks each body's incidences per mass instead
    rng = np.random.default_rng(23)
    bodies = []
    for q in range(520):
        b = scenes.lattice_body((2.0 * (q % 26), 2.0 * (q // 26), 0.6), (1.0, 1.0, 1.0), 3 + q % 3, damping=0.1)
        b["pos"] = b["pos"] + rng.normal(0, 1e-3, b["pos"].shape)
        b["vel"] = rng.normal(0, 0.3, b["pos"].shape)
        bodies.append(b)
    sc = scenes.Scene(bodies=bodies, planes=[(0.0, 0.0, 1.0, 0.0, 0.8, 1.0)], dt=1e-4)
    _say("case 520 resident bodies (27..125 masses), 40 steps")
    with Context() as ctx:
        ids = sc.load(ctx, layout=LAYOUT_RESIDENT)
        ctx.step(40)
        got = [ctx.download(i) for i in ids]
    for q in range(0, 520, 7):
        o = to_oracle(scenes.Scene(bodies=[bodies[q]], planes=sc.planes, dt=sc.dt))
        o.run(40)
        np.testing.assert_array_equal(got[q]["pos"], o.pos)
        np.testing.assert_array_equal(got[q]["vel"], o.vel)


@LAYOUTS
def test_many_balls_culled_by_grid_bitexact(layout):
    # 300 balls (above the grid threshold) scattered through and around a
    # falling lattice: the engine tests only each mass's cell list, the oracle
    # every ball in creation order
    rng = np.random.default_rng(7)
    b = scenes.lattice_body((0.0, 0.0, 1.2), (2.0, 2.0, 2.0), 12, damping=0.1)
    balls = [(*rng.uniform(-1.5, 1.5, 2), rng.uniform(0.0, 2.5), rng.uniform(0.05, 0.35)) for _ in range(300)]
    sc = scenes.Scene(bodies=[b], planes=[(0.0, 0.0, 1.0, 0.0, 0.8, 1.0)], balls=balls, dt=1e-4)
    _say(f"case 300 balls culled by grid, {sc.n_masses} masses, 200 steps, layout {layout}")
    eng, _ = run_engine(sc, 200, layout=layout)
    orc = run_oracle(sc, 200)
    assert_bitexact(eng, orc)
    # the balls act: the run differs from the same scene without them
    free, _ = run_engine(scenes.Scene(bodies=[b], planes=sc.planes, dt=1e-4), 200, layout=layout)
    assert not np.array_equal(free["pos"], eng["pos"])


@LAYOUTS
def test_knn_cloud_with_balls_bitexact(layout):
    # irregular degrees, per-spring k, balls through the cloud and the ground plane
    sc = scenes.knn_scene(3000, n_balls=40, seed=2, dt=1e-5)
    _say(f"case kNN cloud {sc.n_masses} masses, {sc.n_springs} springs, 40 balls, 300 steps, layout {layout}")
    degs = np.bincount(np.concatenate([sc.bodies[0]["left"], sc.bodies[0]["right"]]))
    assert degs.min() < degs.max()
    eng, _ = run_engine(sc, 300, layout=layout)
    orc = run_oracle(sc, 300)
    assert_bitexact(eng, orc)


@LAYOUTS3
def test_per_body_constants_in_one_context_bitexact(layout):
    # four bodies: k and damping shared or per spring, in every combination
    rng = np.random.default_rng(4)
    bodies = []
    for q, (pk, pd) in enumerate([(False, False), (True, False), (False, True), (True, True)]):
        b = scenes.lattice_body((3.0 * q, 0.0, 2.4), (2.0, 2.0, 2.0), 7, damping=0.1)
        s = len(b["left"])
        b["k"] = rng.uniform(5e3, 2e4, s) if pk else np.full(s, 1e4 + 500.0 * q)
        b["damping"] = rng.uniform(0.0, 0.3, s) if pd else np.full(s, 0.05 * q)
        bodies.append(b)
    sc = scenes.Scene(bodies=bodies, planes=[(0.0, 0.0, 1.0, 0.0, 0.8, 1.0)], dt=1e-4)
    _say(f"case per-body constants, 4 bodies, 300 steps, layout {layout}")
    eng, _ = run_engine(sc, 300, layout=layout)
    orc = run_oracle(sc, 300)
    assert_bitexact(eng, orc)


@LAYOUTS
def test_extern_force_uploaded_after_graph_capture(layout):
    # the body starts without extern forces; after graph-launched steps an
    # upload gives some masses one (the engine then creates its extern-force
    # array, so launches captured before must not be replayed)
    sc = scenes.readme_lattice(2.45)
    n = len(sc.bodies[0]["pos"])
    rng = np.random.default_rng(12)
    ext = rng.normal(0, 0.5, (n, 3)) * (np.arange(n) % 4 == 0)[:, None]
    _say(f"case extern force uploaded after 128 graph-launched steps, then 128 more, layout {layout}")
    with Context() as ctx:
        (bid,) = sc.load(ctx, layout=layout)
        ctx.step(128)
        ctx.upload(bid, ext=ext)
        ctx.step(128)
        got = ctx.download(bid)
    o = to_oracle(sc)
    o.run(128)
    o.e = [np.ascontiguousarray(ext[:, c]) for c in range(3)]
    o.run(128)
    np.testing.assert_array_equal(got["pos"], o.pos)
    np.testing.assert_array_equal(got["vel"], o.vel)


def test_split_form_lean_and_general_mass_pass():
    # The split-form launch takes the lean mass pass (Euler, no balls, no
    # extern force, drag or local constraints) when every mass allows it.
    # Uploading extern forces mid-run must switch a graph-launched body to
    # the general pass and back; drag, a local plane, a ball or another
    # integrator keep it general.  Each against the oracle, bit for bit.
    sc = scenes.readme_lattice(2.45)
    n = sc.n_masses
    ext = np.zeros((n, 3))
    ext[::9] = np.random.default_rng(4).normal(0.0, 0.5, (len(ext[::9]), 3))
    _say("case split form: extern force uploaded mid-run and removed again (graph launches)")
    o = to_oracle(sc)
    with Context() as ctx:
        (bid,) = sc.load(ctx)
        assert ctx.body_layout(bid) == LAYOUT_RESIDENT
        assert "k_rollout_split" in ctx.kernel_name()
        ctx.step(70)
        o.run(70)
        ctx.upload(bid, ext=ext)
        o.e = [np.ascontiguousarray(ext[:, c]) for c in range(3)]
        ctx.step(70)
        o.run(70)
        got = ctx.download(bid)
        np.testing.assert_array_equal(got["pos"], o.pos)
        np.testing.assert_array_equal(got["vel"], o.vel)
        ctx.upload(bid, ext=np.zeros((n, 3)))
        o.e = None
        ctx.step(70)
        o.run(70)
        got = ctx.download(bid)
        np.testing.assert_array_equal(got["pos"], o.pos)
        np.testing.assert_array_equal(got["vel"], o.vel)
    b = dict(sc.bodies[0])
    b["drag"] = (np.arange(n) % 11 == 0).astype(np.int32)
    b["lc_ptr"] = np.arange(n + 1, dtype=np.int64)
    b["lc_type"] = np.full(n, 1, np.int32)
    b["lc_vec"] = np.tile([0.0, 0.0, 1.0], (n, 1))
    b["lc_num"] = np.full(n, -0.01)
    cases = [("drag and a local plane", scenes.Scene(bodies=[b], planes=sc.planes, dt=sc.dt), [], 0),
             ("a ball", sc, [((0.0, 0.0, 2.0), 0.3)], 0),
             ("RK2", sc, [], 2)]
    for name, s2, balls, integ in cases:
        _say(f"case split form with {name}: 150 steps")
        o = to_oracle(s2, integrator=integ)
        for c, r in balls:
            o.balls = np.ascontiguousarray(np.array([[*c, r]], dtype=np.float64))
        with Context() as ctx:
            if integ:
                ctx.set_integrator(integ)
            for c, r in balls:
                ctx.add_ball(c, r)
            (bid,) = s2.load(ctx)
            assert "k_rollout_split" in ctx.kernel_name()
            ctx.step(150)
            o.run(150)
            got = ctx.download(bid)
        np.testing.assert_array_equal(got["pos"], o.pos)
        np.testing.assert_array_equal(got["vel"], o.vel)


TILE_DTYPE = np.dtype([("body", "<i4"), ("chunk0", "<i4"), ("nmass", "<i4"), ("nhalo", "<i4"),
                       ("halo_off", "<i8"), ("first", "<i4"), ("nact", "<i4"), ("act_off", "<i8")])


def test_tile_actuator_rows_bound_tiles_and_match_gather():
    """A tile of an actuated body computes its actuators into LDS rows after
    its halo (six to a row), and those rows count against the tile row
    budget: under a 300-row budget a 60%-actuated 12^3 lattice is cut into
    more, smaller tiles than its passive copy.  Its steps equal those of the
    gather layout (which reads the actuated rest lengths that the previous
    launch wrote to a buffer) bit for bit, and the oracle's to 1e-10."""
    rng = np.random.default_rng(7)
    b = scenes.actuate(scenes.lattice_body((0.0, 0.0, 0.45), (1.0, 1.0, 1.0), 12, damping=0.1), 0.6, rng)
    sc = scenes.Scene(bodies=[b], planes=[(0.0, 0.0, 1.0, 0.0, 0.8, 1.0)], dt=1e-4)
    passive = dict(b, type=None, omega=None, act_a=None, act_phi=None)
    _say(f"case tile actuator rows: 12^3 lattice, {sc.n_actuated} of {sc.n_springs} springs actuated, "
         "300-row tiles, 200 steps")
    with Context() as ctx:
        ctx.set_schedule(tile_rows=300)
        [bid] = sc.load(ctx, layout=LAYOUT_TILED)
        pid = ctx.add_body(**passive, layout=LAYOUT_TILED)
        t = np.frombuffer(ctx.layout_bytes(bid, "tiles"), TILE_DTYPE)
        tp = np.frombuffer(ctx.layout_bytes(pid, "tiles"), TILE_DTYPE)
        rec = np.frombuffer(ctx.layout_bytes(bid, "tact"), "<f8").reshape(-1, 4)
        assert np.all(tp["nact"] == 0) and np.all(t["nact"] > 0)
        rows = t["nmass"] + t["nhalo"] + (t["nact"] + 5) // 6
        assert np.all((rows <= 300) | (t["nmass"] <= 64))
        assert len(t) > len(tp)  # the actuator rows took room from masses and halo
        assert np.array_equal(t["act_off"], np.concatenate([[0], np.cumsum(t["nact"])[:-1]]))
        assert len(rec) == t["nact"].sum()
        # each record is an actuated spring's (a, omega, phi, rest)
        act = np.asarray(b["type"]) == 4
        want = {tuple(r) for r in np.stack([b["act_a"][act], b["omega"][act], b["act_phi"][act],
                                            np.asarray(b["rest"])[act]], axis=1)}
        assert {tuple(r) for r in rec} == want
    eng, _ = run_engine(sc, 200, layout=LAYOUT_TILED)
    ref, _ = run_engine(sc, 200, layout=LAYOUT_GATHER)
    np.testing.assert_array_equal(eng["pos"], ref["pos"])
    np.testing.assert_array_equal(eng["vel"], ref["vel"])
    orc = run_oracle(sc, 200)
    assert normwise_rel(eng["pos"], orc.pos) <= TOL_ACT
    assert normwise_rel(eng["vel"], orc.vel) <= TOL_ACT
