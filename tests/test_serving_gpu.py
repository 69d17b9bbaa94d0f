"""Serving (titan_serve_*): continuous batching of rollout requests.

Each request's result must be the same whether it is served alone or in a
mix that arrives while others run (bit for bit: a body's arithmetic never
depends on the others), and the same as the oracle run of that body for its
horizon (bit for bit for passive bodies, 1e-10 normwise relative for
actuated ones, whose sin is the device's).  Admission respects the request
limit and the serving HBM budget, cancellation frees the slot, and the
statistics count what was served."""
from __future__ import annotations

import numpy as np
import pytest

from titan_amd import (INJECT_DESTROY, INJECT_DOWNLOAD, REQ_CANCELLED, REQ_DONE, REQ_QUEUED, REQ_RUNNING,
                       TITAN_E_INJECTED, Context, TitanError, scenes)
from helpers import normwise_rel, to_oracle

pytestmark = pytest.mark.gpu


def _requests():
    robots = scenes.robot_population(5, side_lo=3, side_hi=9, seed=11)
    lat = scenes.readme_lattice(2.45)
    bodies = robots.bodies[:4] + [lat.bodies[0], robots.bodies[4]]
    horizons = [150, 37, 400, 64, 129, 250]
    return bodies, horizons, lat.planes


def _serve(bodies, horizons, planes, *, order=None, quantum=64, max_running=4096, hbm=0, arrive_every=0):
    """Serve the requests; with arrive_every > 0, request q is submitted before
    pump number q * arrive_every (later arrivals join a running batch)."""
    order = list(range(len(bodies))) if order is None else order
    with Context() as ctx:
        ctx.set_dt(1e-4)
        for p in planes:
            ctx.add_plane(p[:3], p[3], p[4], p[5])
        ctx.serve_config(max_running=max_running, hbm_bytes=hbm, quantum=quantum)
        tickets, pumps, peak = {}, 0, 0
        pending = list(order)
        while pending or any(ctx.serve_poll(t)[0] in (REQ_QUEUED, REQ_RUNNING) for t in tickets.values()):
            while pending and (arrive_every == 0 or pumps >= arrive_every * len(tickets)):
                q = pending.pop(0)
                tickets[q] = ctx.serve_submit(bodies[q], horizons[q])
            running, _ = ctx.serve_pump()
            peak = max(peak, running)
            pumps += 1
            assert pumps < 10_000
        res = {q: ctx.serve_result(t) for q, t in tickets.items()}
        for q, t in tickets.items():
            assert ctx.serve_poll(t) == (REQ_DONE, horizons[q])
        stats = ctx.serve_stats()
    return res, peak, stats


def test_each_result_is_the_same_alone_and_in_the_mix():
    bodies, horizons, planes = _requests()
    print(f"case serving alone vs mix: {len(bodies)} requests, horizons {horizons}", flush=True)
    alone = {}
    for q in range(len(bodies)):
        r, _, _ = _serve([bodies[q]], [horizons[q]], planes)
        alone[q] = r[0]
    mixes = [dict(quantum=64), dict(quantum=16, arrive_every=3, order=[2, 0, 5, 1, 4, 3]),
             dict(quantum=33, max_running=2, order=[5, 4, 3, 2, 1, 0])]
    for kw in mixes:
        mix, _, stats = _serve(bodies, horizons, planes, **kw)
        assert stats["completed"] == len(bodies)
        for q in range(len(bodies)):
            np.testing.assert_array_equal(mix[q][0], alone[q][0])
            np.testing.assert_array_equal(mix[q][1], alone[q][1])
    # and against the oracle
    for q in range(len(bodies)):
        o = to_oracle(scenes.Scene(bodies=[bodies[q]], planes=planes, dt=1e-4))
        o.run(horizons[q])
        if bodies[q].get("type") is not None and np.any(np.asarray(bodies[q]["type"]) >= 2):
            assert normwise_rel(alone[q][0], o.pos) <= 1e-10
            assert normwise_rel(alone[q][1], o.vel) <= 1e-10
        else:
            np.testing.assert_array_equal(alone[q][0], o.pos)
            np.testing.assert_array_equal(alone[q][1], o.vel)


def _est(b):
    # the serving layer's documented per-body estimate (titan_amd/csrc/serve.cpp)
    return 256 * len(b["pos"]) + 96 * len(b["left"]) + (1 << 16)


def _replay(bodies, horizons, budget, max_running, quantum):
    """Peak running count, as titan_serve_pump reports it (after retiring), of
    its admission rule: in arrival order, within max_running and the byte
    budget (a request is always admitted when nothing else is admitted or
    running).  A pump admits first only when nothing runs; otherwise it steps
    the running set by up to `quantum` steps each, admits arrivals for the
    next pump, then retires the finished."""
    queue, running, peak = list(range(len(bodies))), [], 0
    left = list(horizons)

    def admit(fresh):
        while queue and len(running) + len(fresh) < max_running:
            busy = running + fresh
            if budget and busy and sum(_est(bodies[r]) for r in busy) + _est(bodies[queue[0]]) > budget:
                break
            fresh.append(queue.pop(0))

    while queue or running:
        fresh = []
        if not running:
            admit(fresh)
            running, fresh = fresh, []
        for r in running:
            left[r] -= min(quantum, left[r])
        admit(fresh)
        running = [r for r in running if left[r] > 0] + fresh
        peak = max(peak, len(running))
    return peak


def test_admission_respects_request_limit_and_budget(monkeypatch):
    monkeypatch.setenv("TITAN_SERVE_ADMIT", "wait")  # admission by pump, as _replay models it
    bodies, horizons, planes = _requests()
    _, peak, stats = _serve(bodies, horizons, planes, max_running=2, quantum=32)
    assert peak == 2 and stats["completed"] == len(bodies)
    big = max(_est(b) for b in bodies)
    for budget in (big, big + 2 * min(_est(b) for b in bodies)):
        want = _replay(bodies, horizons, budget, 4096, 32)
        _, peak, stats = _serve(bodies, horizons, planes, hbm=budget, quantum=32)
        assert peak == want and stats["completed"] == len(bodies)
    assert _replay(bodies, horizons, big, 4096, 32) < len(bodies)


def test_cancel_and_stats(monkeypatch):
    monkeypatch.setenv("TITAN_SERVE_ADMIT", "wait")  # the first pump admits the first two
    bodies, horizons, planes = _requests()
    with Context() as ctx:
        ctx.set_dt(1e-4)
        ctx.serve_config(max_running=2, quantum=16)
        t = [ctx.serve_submit(b, h) for b, h in zip(bodies, horizons)]
        ctx.serve_pump()
        assert ctx.serve_poll(t[0])[0] == REQ_RUNNING and ctx.serve_poll(t[5])[0] == REQ_QUEUED
        ctx.serve_cancel(t[0])  # running
        ctx.serve_cancel(t[5])  # queued
        assert ctx.serve_poll(t[0])[0] == REQ_CANCELLED and ctx.serve_poll(t[5])[0] == REQ_CANCELLED
        while any(ctx.serve_poll(x)[0] in (REQ_QUEUED, REQ_RUNNING) for x in t):
            ctx.serve_pump()
        s = ctx.serve_stats()
        assert (s["submitted"], s["completed"], s["cancelled"], s["running"], s["queued"]) == (6, 4, 2, 0, 0)
        want = sum(len(bodies[q]["pos"]) * horizons[q] for q in (1, 2, 3, 4))
        assert s["mass_steps"] == want
        assert 0 < s["p50_latency_s"] <= s["p99_latency_s"] <= s["max_latency_s"]
        ctx.serve_reset_stats()
        assert ctx.serve_stats()["completed"] == 0


def test_requests_without_results_complete_after_their_steps():
    # a request whose result nobody reads is released at once but completes
    # only when the step stream has run its last steps; cancelling it in
    # between loses nothing
    bodies, horizons, planes = _requests()
    with Context() as ctx:
        ctx.set_dt(1e-4)
        ctx.serve_config(quantum=16)
        t = [ctx.serve_submit(b, h, want_result=False) for b, h in zip(bodies, horizons)]
        while ctx.serve_pump() != (0, 0):
            pass
        s = ctx.serve_stats()
        assert (s["completed"], s["running"], s["queued"]) == (len(t), 0, 0)
        assert all(ctx.serve_poll(x)[0] == REQ_DONE for x in t)
        assert s["mass_steps"] == sum(len(b["pos"]) * h for b, h in zip(bodies, horizons))
        late = ctx.serve_submit(bodies[0], 5, want_result=False)
        ctx.serve_pump()  # its five steps are enqueued and it is retired
        assert ctx.serve_poll(late)[0] in (REQ_RUNNING, REQ_DONE)
        ctx.serve_cancel(late)
        assert ctx.serve_poll(late)[0] in (REQ_DONE, REQ_CANCELLED)
        while ctx.serve_pump() != (0, 0):
            pass
        s = ctx.serve_stats()
        assert s["completed"] + s["cancelled"] == len(t) + 1 and s["running"] == 0


@pytest.mark.parametrize("site", [INJECT_DOWNLOAD, INJECT_DESTROY])
def test_failed_retirement_loses_no_request(site, monkeypatch):
    # The pump that retires request 1 (37 steps, quantum 16: its third pump)
    # also admits the four requests that arrived meanwhile; its retirement
    # fails once (the result's download, or the body's release).  The pump
    # reports the failure, the new arrivals still run, request 1 is retired by
    # a later pump, cancelling works, and every result is the one served
    # without a failure.
    bodies, horizons, planes = _requests()
    want, _, _ = _serve(bodies, horizons, planes, quantum=16)
    print(f"case failed retirement at site {site}", flush=True)
    monkeypatch.setenv("TITAN_SERVE_ADMIT", "wait")  # the failing pump admits the new arrivals
    with Context() as ctx:
        ctx.set_dt(1e-4)
        for p in planes:
            ctx.add_plane(p[:3], p[3], p[4], p[5])
        ctx.serve_config(quantum=16)
        t = {q: ctx.serve_submit(bodies[q], horizons[q]) for q in (0, 1)}
        ctx.serve_pump()
        ctx.serve_pump()
        for q in range(2, len(bodies)):
            t[q] = ctx.serve_submit(bodies[q], horizons[q])
        ctx.inject_failure(0, site)
        with pytest.raises(TitanError) as err:
            ctx.serve_pump()
        assert err.value.code == TITAN_E_INJECTED
        assert ctx.serve_poll(t[1])[0] == REQ_RUNNING
        assert all(ctx.serve_poll(t[q])[0] == REQ_RUNNING for q in range(2, len(bodies)))
        ctx.serve_cancel(t[3])
        ctx.serve_cancel(t[3])  # a second cancel of a finished request changes nothing
        pumps = 0
        while any(ctx.serve_poll(x)[0] in (REQ_QUEUED, REQ_RUNNING) for x in t.values()):
            ctx.serve_pump()
            pumps += 1
            assert pumps < 2000
        assert ctx.serve_poll(t[3])[0] == REQ_CANCELLED
        for q in t:
            if q == 3:
                continue
            assert ctx.serve_poll(t[q]) == (REQ_DONE, horizons[q])
            got = ctx.serve_result(t[q])
            np.testing.assert_array_equal(got[0], want[q][0])
            np.testing.assert_array_equal(got[1], want[q][1])
        s = ctx.serve_stats()
        assert (s["completed"], s["cancelled"], s["running"], s["queued"]) == (len(bodies) - 1, 1, 0, 0)


def test_knn_cloud_requests_alone_and_mixed():
    # the served workload's other half: actuated kNN-12 clouds (here 1500 and
    # 6000 masses, as scenes.serving_requests builds them), served alone and
    # mixed with a robot and the README lattice, against the oracle
    rng = np.random.default_rng(21)
    clouds = [scenes.actuate(scenes.knn_cloud(n, box=1.0, seed=s, z0=0.1), f, rng)
              for n, s, f in ((1500, 5, 0.2), (6000, 6, 0.3))]
    robots = scenes.robot_population(1, side_lo=8, side_hi=8, seed=13)
    lat = scenes.readme_lattice(2.45)
    bodies = clouds + [robots.bodies[0], lat.bodies[0]]
    horizons = [170, 300, 90, 233]
    planes = lat.planes
    print(f"case kNN requests {[len(b['pos']) for b in bodies]} horizons {horizons}", flush=True)
    alone = {q: _serve([bodies[q]], [horizons[q]], planes)[0][0] for q in range(len(bodies))}
    for kw in (dict(quantum=64), dict(quantum=16, arrive_every=2, order=[3, 1, 2, 0])):
        mix, _, stats = _serve(bodies, horizons, planes, **kw)
        assert stats["completed"] == len(bodies)
        for q in range(len(bodies)):
            np.testing.assert_array_equal(mix[q][0], alone[q][0])
            np.testing.assert_array_equal(mix[q][1], alone[q][1])
    for q in range(len(bodies)):
        o = to_oracle(scenes.Scene(bodies=[bodies[q]], planes=planes, dt=1e-4))
        o.run(horizons[q])
        assert normwise_rel(alone[q][0], o.pos) <= 1e-10
        assert normwise_rel(alone[q][1], o.vel) <= 1e-10


def _body_with_every_array(side, seed):
    """A passive lattice with every optional array of titan_body_desc set:
    velocities, local times, extern forces, fixed masses, drag, and local
    constraints (constraint planes, directions and contact planes)."""
    rng = np.random.default_rng(seed)
    b = scenes.lattice_body((0.0, 0.0, 0.58), (1.0, 1.0, 1.0), side, damping=0.1)
    n = len(b["pos"])
    cnt = rng.integers(0, 3, n)
    b["lc_ptr"] = np.concatenate([[0], np.cumsum(cnt)]).astype(np.int64)
    tot = int(cnt.sum())
    b["lc_type"] = rng.choice([0, 3, 1], size=tot).astype(np.int32)
    v = rng.normal(size=(tot, 3))
    b["lc_vec"] = v / np.sqrt((v * v).sum(1))[:, None]
    b["lc_num"] = rng.uniform(0.0, 0.4, tot)
    b["vel"] = rng.normal(0, 0.3, (n, 3))
    b["T"] = rng.uniform(0.0, 1e-2, n)
    b["fixed"] = (np.arange(n) % 19 == 0).astype(np.uint8)
    b["drag"] = (np.arange(n) % 4).astype(np.int32)
    b["ext"] = rng.normal(0, 0.2, (n, 3)) * (np.arange(n) % 2 == 0)[:, None]
    return b


def test_every_array_of_a_request_reaches_the_device():
    # submission copies a request's arrays into one block of its own, a large
    # body (4 MB or more) by several threads: a small body and a 40^3 lattice
    # (about 30 MB of arrays), served together, match the oracle bit for bit
    planes = [(0.0, 0.0, 1.0, 0.0, 0.8, 1.0)]
    bodies = [_body_with_every_array(6, 1), _body_with_every_array(40, 2)]
    horizons = [90, 40]
    print(f"case serving every array: masses {[len(b['pos']) for b in bodies]}, "
          f"springs {[len(b['left']) for b in bodies]}, horizons {horizons}", flush=True)
    res, _, stats = _serve(bodies, horizons, planes)
    assert stats["completed"] == 2
    for q in range(2):
        o = to_oracle(scenes.Scene(bodies=[bodies[q]], planes=planes, dt=1e-4))
        o.run(horizons[q])
        np.testing.assert_array_equal(res[q][0], o.pos)
        np.testing.assert_array_equal(res[q][1], o.vel)
