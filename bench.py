#!/usr/bin/env python3
"""Benchmark of the MI355X FP64 mass-spring step.

    python bench.py [--gpus 1] [--steps 1000] [--warmup 50] [--no-cpu]

Workload (one GPU): a 100x100x100 lattice (1,000,000 masses, 12,731,796
springs, 26-neighbourhood, k = 1e4, damping 0.1, m = 0.1 kg) spanning 5 m
around (0, 0, 2.5), positions jittered by N(0, 1e-4) m (seed 0), rest lengths =
jittered distances, contact plane z = 0 with mu_k = 0.8 / mu_s = 1.0,
g = 9.81, dt = 1e-5, FP64.  Inputs are resident in HBM before the timed
region; a step is one launch of the fused step kernel over the whole lattice.

Prints one JSON line: mass-steps/s, the roofline fraction of the step kernel
(algorithmic bytes per step / average step time from HIP events on the
engine's stream, against 8 TB/s), and the CPU oracle (OpenMP) timed on a
bounded sample of the same workload.  Unless --no-extra, the same line also
carries "batched": a 4096-body robot population stepped together against the
same GPU stepping one body per launch, and "serving": Poisson streams of
rollout requests served with continuous batching (p50/p99 latency and
mass-steps/s per arrival rate) against serving them one at a time, and
"irregular": a 2M-mass kNN-12 cloud with 1000 contact balls.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "mass-steps/s FP64 + % HBM roofline; p50/p99 rollout latency under load"
HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def cpu_threads() -> int:
    # the GPU host gives one GPU's share of cores (16); respect OMP_NUM_THREADS
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit():
        return int(env)
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_baseline(scene, budget_s: float = 15.0) -> dict:
    """The oracle on the host's cores over a bounded number of steps of the
    same workload (state copied fresh from the scene)."""
    from oracle import oracle as orc  # checker / CPU baseline only
    from tests.helpers import to_oracle

    threads = cpu_threads()
    o = to_oracle(scene, threads=threads)
    o.run(5)  # warm-up (page-in, thread pool)
    k = 10    # calibration: the per-step time after warm-up sizes the timed sample
    t0 = time.perf_counter()
    o.run(k)
    per = (time.perf_counter() - t0) / k
    # three timed runs sharing the budget; the median run's rate is reported
    n = max(3, int(budget_s / 3 / max(per, 1e-6)))
    runs = []
    for _ in range(3):
        t0 = time.perf_counter()
        o.run(n)
        runs.append(time.perf_counter() - t0)
    dt = sorted(runs)[1]
    value = scene.n_masses * n / dt
    return {"value": value, "unit": "mass-steps/s", "cores": threads, "kind": "port",
            "sample": f"oracle (C++17/OpenMP, {threads} threads, {cpu_model()}) on the same "
                      f"1M-mass lattice, median of 3 runs of {n} steps ({', '.join(f'{r:.2f}' for r in runs)} s) "
                      f"after {5 + k} untimed ones",
            "ms_per_step": 1e3 * dt / n}


def cpu_population(sc, rollout: int, sub_bodies: int) -> dict:
    """The oracle on the robot population: the whole population for one
    rollout on all of the host's threads, and its first `sub_bodies` bodies
    for one rollout on one thread (mass-steps/s; bodies are independent, so
    the rate of a sample stands for the population's)."""
    from oracle import oracle as orc  # CPU baseline only
    from tests.helpers import to_oracle
    from titan_amd import scenes

    threads = cpu_threads()
    out = {}
    for name, scene, th in (("all_cores", sc, threads),
                            ("one_core", scenes.Scene(bodies=sc.bodies[:sub_bodies], planes=sc.planes, dt=sc.dt), 1)):
        o = to_oracle(scene, threads=th)
        o.run(2)  # warm-up
        t0 = time.perf_counter()
        o.run(rollout)
        dt = time.perf_counter() - t0
        out[name] = {"value": scene.n_masses * rollout / dt, "unit": "mass-steps/s", "cores": th, "kind": "port",
                     "sample": f"oracle (C++17/OpenMP, {cpu_model()}), {len(scene.bodies)} bodies "
                               f"({scene.n_masses} masses), one {rollout}-step rollout, {dt:.1f} s"}
        log(f"batched CPU {name}: {out[name]['value']:.3e} mass-steps/s ({th} threads, {dt:.1f} s)")
    return out


def bench_batched(n_bodies: int, seed: int, rollout: int, sample: int, schedule=None, cpu: bool = True) -> dict:
    """Batched uneven bodies: a robot population (lattices of side U{3..12},
    30% of springs actuated, each body with its own contact plane, dt 1e-4)
    stepped in one launch per step (launches this large are issued directly,
    smaller ones replayed from captured 64-step graphs), against the
    same GPU stepping one body per launch.  The schedule (which bodies are
    tiled, the tile row budget) is picked by the library's autotuner on a
    sample of the population, unless `schedule` gives it.  The
    one-body-per-launch figure is measured on the first `sample` bodies (none
    when 0); both runs are checked to give identical states."""
    import numpy as np
    from titan_amd import Context, scenes

    t0 = time.perf_counter()
    sc = scenes.robot_population(n_bodies, seed=seed)
    n_m, n_s = sc.n_masses, sc.n_springs
    log(f"batched: {n_bodies} bodies, {n_m} masses, {n_s} springs, {sc.n_actuated} actuated "
        f"({time.perf_counter() - t0:.1f} s)")
    sub = scenes.Scene(bodies=sc.bodies[:sample], planes=sc.planes, dt=sc.dt)
    tune_set = sc.bodies[:64]
    with Context() as ctx:
        ctx.set_dt(sc.dt)
        ctx.set_global_acc(sc.g)
        if schedule is None:
            t0 = time.perf_counter()
            best, rates = ctx.autotune(tune_set, replicas=8, steps=20)
            log(f"batched: autotune {time.perf_counter() - t0:.1f} s -> tiled_min_masses {best[0]}, "
                f"tile_rows {best[1]}, resident_max_masses {best[2]}")
        else:
            best, rates = tuple(schedule), []
            ctx.set_schedule(*best)
            log(f"batched: given schedule tiled_min_masses {best[0]}, tile_rows {best[1]}, "
                f"resident_max_masses {best[2]}")
        t0 = time.perf_counter()
        ids = [ctx.add_body(**b) for b in sc.bodies]
        ctx.sync()
        log(f"batched: uploaded in {time.perf_counter() - t0:.1f} s, device bytes {ctx.mem_used()}")
        ctx.step(rollout)  # warm-up rollout
        ctx.sync()
        ctx.timer_start()
        ctx.step(rollout)
        ms_batched = ctx.timer_stop()
        del ids
    rate_batched = n_m * rollout / (ms_batched * 1e-3)
    log(f"batched: {rollout}-step rollout {ms_batched:.2f} ms, {rate_batched:.3e} mass-steps/s")
    out = {
        "workload": f"{n_bodies} lattice robots, side U{{3..12}} (seed {seed}), {n_m} masses, {n_s} springs, "
                    f"{sc.n_actuated} actuated (a~U(0,0.2), omega~U(1,20), phi~U(0,2pi)), own contact plane "
                    f"each, dt=1e-4, FP64, {rollout}-step rollouts, one launch per step, issued directly (graph "
                    f"capture is kept for launches under 1024 workgroups; this population's ~11,000 measured "
                    f"equal either way) (the "
                    f"resident layout, one launch per rollout, is among the autotuner's candidates and is "
                    f"slower here: one workgroup per body walks its chunks in turn every step)",
        "value": rate_batched, "unit": "mass-steps/s", "ms_per_rollout": ms_batched,
        "schedule": {"tiled_min_masses": best[0], "tile_rows": best[1], "resident_max_masses": best[2],
                     "autotune": [{"tiled_min_masses": c[0], "tile_rows": c[1], "resident_max_masses": c[2],
                                   "mass_steps_per_s": r}
                                  for c, r in rates]},
    }
    if cpu:
        out["cpu_baseline"] = cpu_population(sc, rollout, 256)
    if sample <= 0:
        return out

    # the same schedule, the sample bodies: batched vs one body per launch, same states
    def run(one_per_launch: bool):
        with Context() as c:
            c.set_dt(sc.dt)
            c.set_global_acc(sc.g)
            c.set_schedule(*best)
            bid = [c.add_body(**b) for b in sub.bodies]
            c.sync()
            c.timer_start()
            if one_per_launch:
                for b in bid:
                    c.step_bodies([b], [rollout])
            else:
                c.step(rollout)
            ms = c.timer_stop()
            st = [c.download(b) for b in bid]
        return ms, st
    run(True)  # warm-up
    ms_one, st_one = run(True)
    _, st_bat = run(False)
    same = all(np.array_equal(a["pos"], b["pos"]) and np.array_equal(a["vel"], b["vel"])
               for a, b in zip(st_one, st_bat))
    sub_m = sub.n_masses
    rate_one = sub_m * rollout / (ms_one * 1e-3)
    log(f"batched: one body per launch on {sample} bodies ({sub_m} masses): {ms_one:.1f} ms, "
        f"{rate_one:.3e} mass-steps/s; identical states: {same}")
    return {
        **out,
        "one_body_per_launch": {"value": rate_one, "unit": "mass-steps/s",
                                "sample": f"first {sample} bodies ({sub_m} masses), {rollout} steps each, "
                                          f"one launch per body per step over its own tiles or chunks"},
        "speedup_vs_one_body_per_launch": rate_batched / rate_one,
        "identical_to_one_body_per_launch": same,
    }


def bench_serving(seed: int, n_req: int, lams, quantum: int, n_base: int) -> dict:
    """Serving: rollout requests (lattices of side U{3..40} or kNN-12 clouds of
    1e2..1e5 masses, 10-30% of springs actuated, horizons U{100..10000} steps,
    a ground plane with friction, dt 1e-4) arriving as a Poisson process at
    each rate lambda, admitted and stepped together by the library's serving
    loop (continuous batching, `quantum` steps per pump).  Reports p50/p99
    completion latency and mass-steps/s per lambda, against the same requests
    served one at a time with one launch per step."""
    import numpy as np
    from titan_amd import Context, scenes

    t0 = time.perf_counter()
    reqs = scenes.serving_requests(n_req, seed=seed)
    total = sum(len(b["pos"]) * h for b, h in reqs)
    log(f"serving: {n_req} requests, {sum(len(b['pos']) for b, _ in reqs)} masses, {total / 1e9:.2f}e9 "
        f"mass-steps ({time.perf_counter() - t0:.1f} s to build)")
    rng = np.random.default_rng(seed + 1)

    def setup(ctx, sched=None):
        ctx.set_dt(1e-4)
        ctx.add_plane((0.0, 0.0, 1.0), 0.0, 0.8, 1.0)
        if sched is not None:
            ctx.set_schedule(*sched)

    with Context() as ctx:  # the schedule, by the autotuner, on a sample of the requests
        setup(ctx)
        t0 = time.perf_counter()
        sched, _ = ctx.autotune([b for b, _ in reqs[:6]], replicas=2, steps=16)
        log(f"serving: autotune {time.perf_counter() - t0:.1f} s -> tiled_min_masses {sched[0]}, "
            f"tile_rows {sched[1]}, resident_max_masses {sched[2]}")

    with Context() as ctx:  # warm-up: load the kernels, build a graph
        setup(ctx, sched)
        ctx.serve_config(quantum=quantum)
        ctx.serve_submit(reqs[0][0], 2 * quantum, want_result=False)
        while ctx.serve_pump() != (0, 0):
            pass
    # a server reserves its HBM at startup: the arenas are grown once for the
    # whole request set, so that no growth (which waits for the steps in
    # flight) happens under load
    reserve = (sum(len(b["pos"]) for b, _ in reqs), sum(len(b["left"]) for b, _ in reqs),
               sum(int(np.count_nonzero(np.asarray(b["type"]) >= 2)) for b, _ in reqs
                   if b.get("type") is not None), n_req)
    per_lam = []
    for lam in lams:
        arr = np.cumsum(rng.exponential(1.0 / lam, n_req))
        with Context() as ctx:
            setup(ctx, sched)
            ctx.serve_config(quantum=quantum)
            ctx.reserve(*reserve)
            t0 = time.perf_counter()
            i = pumps = 0
            submit_s = pump_s = 0.0  # host time inside serve_submit and serve_pump
            while True:
                now = time.perf_counter() - t0
                while i < n_req and arr[i] <= now:
                    ts = time.perf_counter()
                    ctx.serve_submit(reqs[i][0], reqs[i][1], want_result=False)
                    submit_s += time.perf_counter() - ts
                    i += 1
                ts = time.perf_counter()
                running, queued = ctx.serve_pump()
                pump_s += time.perf_counter() - ts
                pumps += 1
                if running == 0 and queued == 0:
                    if i == n_req:
                        break
                    time.sleep(max(0.0, arr[i] - (time.perf_counter() - t0)))
            wall = time.perf_counter() - t0
            st = ctx.serve_stats()
        per_lam.append({"lambda_req_per_s": lam, "requests": n_req, "completed": st["completed"],
                        "p50_latency_s": st["p50_latency_s"], "p99_latency_s": st["p99_latency_s"],
                        "max_latency_s": st["max_latency_s"], "mass_steps_per_s": st["mass_steps_per_s"],
                        "wall_s": wall, "pumps": pumps, "host_submit_s": submit_s, "host_pump_s": pump_s,
                        "device_idle_s": st["device_idle_s"]})
        log(f"serving: lambda {lam}/s: p50 {st['p50_latency_s']:.3f} s, p99 {st['p99_latency_s']:.3f} s, "
            f"{st['mass_steps_per_s']:.3e} mass-steps/s, {pumps} pumps, {wall:.1f} s "
            f"(host: {submit_s:.2f} s in submit, {pump_s:.2f} s in pump; device idle between quanta "
            f"{st['device_idle_s']:.2f} s)")
    # the same requests served one at a time, one launch per step
    base = reqs[:n_base]
    with Context() as ctx:
        setup(ctx, sched)
        ids = [ctx.add_body(**b) for b, _ in base]
        ctx.sync()
        ctx.timer_start()
        for bid, (_, h) in zip(ids, base):
            ctx.step_bodies([bid], [h])
        ms = ctx.timer_stop()
    base_rate = sum(len(b["pos"]) * h for b, h in base) / (ms * 1e-3)
    best = max(x["mass_steps_per_s"] for x in per_lam)
    log(f"serving: one request at a time, one launch per step: {base_rate:.3e} mass-steps/s "
        f"({n_base} requests, {ms:.0f} ms)")
    # the same requests all present from the start, stepped together to their
    # horizons (no arrivals, admissions or pumps): the batched rate of this mix
    with Context() as ctx:
        setup(ctx, sched)
        ids = [ctx.add_body(**b) for b, _ in reqs]
        ctx.sync()
        ctx.timer_start()
        ctx.step_bodies(ids, [h for _, h in reqs])
        ms_all = ctx.timer_stop()
    batched_rate = total / (ms_all * 1e-3)
    log(f"serving: the same requests batched from the start: {batched_rate:.3e} mass-steps/s ({ms_all:.0f} ms)")
    return {
        "workload": f"{n_req} rollout requests per lambda (seed {seed}): lattices side U{{3..40}} or kNN-12 "
                    f"clouds of 1e2..1e5 masses, 10-30% actuated, horizon U{{100..10000}} steps, ground "
                    f"plane with friction, dt=1e-4, FP64; Poisson arrivals; continuous batching with "
                    f"{quantum}-step pumps; the arenas reserved for the request set before the arrivals start",
        "per_lambda": per_lam,
        "one_request_at_a_time": {"value": base_rate, "unit": "mass-steps/s",
                                  "sample": f"first {n_base} requests, one launch per step each"},
        "speedup_vs_one_request_at_a_time": best / base_rate,
        "same_requests_batched": {"value": batched_rate, "unit": "mass-steps/s",
                                  "sample": f"all {n_req} requests present at once, stepped to their horizons"},
        "served_over_batched": best / batched_rate,
        "schedule": {"tiled_min_masses": sched[0], "tile_rows": sched[1], "resident_max_masses": sched[2]},
    }


def bench_readme(steps: int, layouts=("gather", "tiled", "resident"), cpu: bool = True) -> dict:
    """The README scene (125 masses, 1036 springs, damping 0.1, ground plane
    mu_k 0.8 / mu_s 1.0, dt 1e-4), both starting heights: latency-bound, so
    reported as time per step for each layout (graph-launched, the call's
    steps in one launch for the resident layout), beside the oracle on one
    host thread."""
    from oracle import oracle as orc  # CPU baseline only
    from tests.helpers import to_oracle
    from titan_amd import LAYOUT_GATHER, LAYOUT_RESIDENT, LAYOUT_TILED, Context, scenes

    out = {}
    for cz in (10.0, 2.45):
        sc = scenes.readme_lattice(cz)
        per = {}
        for name, lay in (("gather", LAYOUT_GATHER), ("tiled", LAYOUT_TILED), ("resident", LAYOUT_RESIDENT)):
            if name not in layouts:
                continue
            with Context() as ctx:
                sc.load(ctx, layout=lay)
                ctx.step(steps // 10)
                ctx.sync()
                ctx.timer_start()
                ctx.step(steps)
                per[name] = 1e3 * ctx.timer_stop() / steps
        cpu_us = None
        if cpu:
            o = to_oracle(sc, threads=1)
            o.run(steps // 10)
            t0 = time.perf_counter()
            o.run(steps)
            cpu_us = 1e6 * (time.perf_counter() - t0) / steps
        log(f"readme cz={cz}: us/step " + ", ".join(f"{k} {v:.2f}" for k, v in per.items()) +
            f"; CPU oracle (1 thread) {cpu_us}")
        out[f"cz{cz:g}"] = {"us_per_step": per, "cpu_us_per_step_1_thread": cpu_us}
    return {"workload": f"README lattice 5x5x5 (125 masses, 1036 springs), cz 10 (free fall) and 2.45 "
                        f"(in contact from step 0), dt=1e-4, FP64, {steps} timed steps after {steps // 10}",
            **out}


def bench_irregular(n: int, n_balls: int, seed: int, steps: int, warmup: int) -> dict:
    """Irregular spring graph: a 2M-mass kNN-12 cloud (variable degree,
    per-spring k), 1000 contact balls and a ground plane, dt 1e-5, FP64; one
    launch per step over the whole cloud."""
    import numpy as np
    from titan_amd import Context, scenes

    t0 = time.perf_counter()
    sc = scenes.knn_scene(n, n_balls=n_balls, seed=seed)
    log(f"irregular: {sc.n_masses} masses, {sc.n_springs} springs, {n_balls} balls "
        f"({time.perf_counter() - t0:.1f} s to build)")
    with Context() as ctx:
        sc.load_globals(ctx)
        t0 = time.perf_counter()
        bid = ctx.add_body(**sc.bodies[0])
        ctx.sync()
        setup_s = time.perf_counter() - t0
        log(f"irregular: uploaded in {setup_s * 1e3:.1f} ms, device bytes {ctx.mem_used()}, "
            f"kernel {ctx.kernel_name()}")
        ctx.step(warmup)
        ctx.sync()
        ctx.timer_start()
        ctx.step(steps)
        ms = ctx.timer_stop()
        st = ctx.download(bid)
        kernel = ctx.kernel_name()
    finite = bool(np.isfinite(st["pos"]).all() and np.isfinite(st["vel"]).all())
    step_s = ms * 1e-3 / steps
    bytes_step = sc.algorithmic_bytes_per_step()
    achieved = bytes_step / step_s / 1e9
    log(f"irregular: {steps} steps {ms:.1f} ms, {step_s * 1e6:.1f} us/step, "
        f"{sc.n_masses * steps / (ms * 1e-3):.3e} mass-steps/s, {achieved / HBM_PEAK_GBS:.3f} of HBM peak")
    return {
        "workload": f"kNN-12 cloud of {sc.n_masses} masses (seed {seed}) in a 10 m box, {sc.n_springs} springs "
                    f"after dedup, k~LogUniform(1e3,1e5), damping 0.05, {n_balls} contact balls r~U(0.1,0.5), "
                    f"ground plane with friction, dt=1e-5, FP64, {steps} steps after {warmup}",
        "value": sc.n_masses * steps / (ms * 1e-3), "unit": "mass-steps/s", "us_per_step": step_s * 1e6,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "bytes_per_launch": bytes_step, "kernel": kernel},
        "finite_after_run": finite, "body_create_s": setup_s,
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--side", type=int, default=100)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--layout", choices=["auto", "gather", "tiled"], default="auto",
                    help="device layout of the body (results are identical; speed differs)")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the README, batched, serving and irregular-graph workloads")
    ap.add_argument("--only", choices=["lattice", "readme", "batched", "serving", "irregular"], default=None,
                    help="run one workload (its figures go in the same JSON line)")
    ap.add_argument("--bodies", type=int, default=4096, help="batched workload: robot count")
    ap.add_argument("--schedule", default=None, metavar="MIN,ROWS,RES",
                    help="batched workload: this schedule (tiled_min_masses,tile_rows,resident_max_masses) "
                         "instead of the autotuner's")
    ap.add_argument("--per-body-sample", type=int, default=512,
                    help="batched workload: bodies timed one per launch for the comparison (0: none)")
    ap.add_argument("--requests", type=int, default=256, help="serving workload: requests per arrival rate")
    ap.add_argument("--lambdas", default="50,200,1000,2000",
                    help="serving workload: Poisson arrival rates (requests/s), comma-separated")
    ap.add_argument("--copy-calib", type=int, default=0, metavar="MIB",
                    help="only run the library's streaming copy and read kernels (k_stream_copy, "
                         "k_stream_read) over MIB MiB, 3 timed launches each after a warm-up: the "
                         "known-size kernels for scaling the FETCH_SIZE / WRITE_SIZE counters")
    ap.add_argument("--readme-layouts", default="gather,tiled,resident",
                    help="README workload: the layouts to time (comma-separated)")
    args = ap.parse_args()

    if args.copy_calib:
        from titan_amd import measure_hbm

        print(json.dumps(measure_hbm(0, args.copy_calib << 20, 3)), flush=True)
        return 0

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "this engine is measured on one GPU; --gpus must be 1 "
                                   "(bodies shard across GPUs with no per-step exchange; a single "
                                   "large body would need a halo exchange that is not built)"}))
        return 2

    extra = {}
    if args.only == "readme" or (args.only is None and not args.no_extra):
        extra["readme"] = bench_readme(10000, layouts=tuple(args.readme_layouts.split(",")), cpu=not args.no_cpu)
    if args.only == "batched" or (args.only is None and not args.no_extra):
        sched = tuple(int(x) for x in args.schedule.split(",")) if args.schedule else None
        extra["batched"] = bench_batched(args.bodies, args.seed, rollout=100,
                                         sample=min(args.per_body_sample, args.bodies), schedule=sched,
                                         cpu=not args.no_cpu)
    if args.only == "serving" or (args.only is None and not args.no_extra):
        lams = tuple(float(x) for x in args.lambdas.split(","))
        extra["serving"] = bench_serving(args.seed, n_req=args.requests, lams=lams,
                                         quantum=256, n_base=16)
    if args.only == "irregular" or (args.only is None and not args.no_extra):
        extra["irregular"] = bench_irregular(2_000_000, 1000, args.seed, steps=500, warmup=20)
    if args.only in ("readme", "batched", "serving", "irregular"):
        print(json.dumps(extra), flush=True)
        return 0
    out = bench_lattice(args)
    out.update(extra)
    print(json.dumps(out), flush=True)
    return 0


def bench_lattice(args) -> dict:
    """The 1M-mass lattice: the headline metric, its roofline and the CPU baseline."""
    from titan_amd import Context, scenes

    log(f"building scene lattice {args.side}^3 seed {args.seed}")
    t0 = time.perf_counter()
    scene = scenes.big_lattice(args.side, seed=args.seed, dt=1e-5)
    log(f"scene: {scene.n_masses} masses, {scene.n_springs} springs ({time.perf_counter() - t0:.1f} s)")

    ctx = Context()
    scene.load_globals(ctx)
    t0 = time.perf_counter()
    bid = ctx.add_body(**scene.bodies[0], layout={"auto": 0, "gather": 1, "tiled": 2}[args.layout])
    ctx.sync()
    setup_s = time.perf_counter() - t0
    log(f"uploaded in {setup_s * 1e3:.1f} ms (titan_body_create: validation, layout built on the device, copies); "
        f"device bytes {ctx.mem_used()}; kernel {ctx.kernel_name()}")

    ctx.step(args.warmup)
    ctx.sync()
    log(f"warm-up {args.warmup} steps done")

    ctx.sync()
    ctx.timer_start()
    w0 = time.perf_counter()
    ctx.step(args.steps)
    ev_ms = ctx.timer_stop()  # records the end event and waits for it
    ctx.sync()
    wall = time.perf_counter() - w0
    log(f"timed {args.steps} steps: wall {wall * 1e3:.2f} ms, events {ev_ms:.2f} ms")

    n = scene.n_masses
    value = n * args.steps / wall
    bytes_step = scene.algorithmic_bytes_per_step()
    step_s = ev_ms * 1e-3 / args.steps
    achieved = bytes_step / step_s / 1e9
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                "kernel": ctx.kernel_name(), "bytes_per_launch": bytes_step,
                "avg_launch_us": step_s * 1e6}
    traffic_file = ROOT / "profiles" / "traffic_lattice100.json"
    if traffic_file.exists() and args.side == 100:
        try:
            tj = json.loads(traffic_file.read_text())
            roofline["traffic"] = tj["hbm_bytes_per_launch"]
            roofline["traffic_source"] = tj.get("source", str(traffic_file.relative_to(ROOT)))
        except Exception:
            pass
    # what this device's HBM sustains for a streaming kernel of the library's
    # own, measured in the same run: the achievable ceiling under the 8 TB/s peak
    from titan_amd import measure_hbm

    cal = measure_hbm(0, 1 << 30, 5)
    roofline["stream_copy_GBps"] = cal["copy_GBps"]
    roofline["stream_read_GBps"] = cal["read_GBps"]
    log(f"streaming copy {cal['copy_GBps']:.0f} GB/s, read {cal['read_GBps']:.0f} GB/s (1 GiB buffers)")

    # sanity: the lattice must still be finite after the run
    st = ctx.download(bid)
    import numpy as np
    finite = bool(np.isfinite(st["pos"]).all() and np.isfinite(st["vel"]).all())
    ctx.close()

    cpu = None
    if not args.no_cpu:
        log("CPU baseline (oracle, OpenMP)")
        cpu = cpu_baseline(scene)
        log(f"CPU: {cpu['value']:.3e} mass-steps/s, {cpu['ms_per_step']:.1f} ms/step")

    out = {
        "metric": METRIC, "value": value, "unit": "mass-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": f"synthetic (seeded lattice, seed {args.seed})",
        "config": {"workload": f"lattice {args.side}^3: {n} masses, {scene.n_springs} springs, "
                               "26-neighbourhood, k=1e4, damping 0.1, contact plane z=0 (mu_k 0.8, mu_s 1.0), "
                               "dt=1e-5, FP64, one launch per step",
                   "masses": n, "springs": scene.n_springs, "dt": 1e-5, "finite_after_run": finite,
                   "body_create_s": setup_s},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    return out


if __name__ == "__main__":
    sys.exit(main())
