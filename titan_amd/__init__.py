"""titan_amd — Python view of the MI355X FP64 mass-spring step engine.

The engine is the native library ``libtitan_mi355x.so`` (HIP kernels for
gfx950 behind the C ABI in ``include/titan_mi355x.h``).  This module binds
that ABI with ctypes for the tests and the benchmark; the C++ drop-in for the
reference's ``titan::Simulation`` lives in ``titan_amd/cpp``.  There is no
fallback: if the library is not built, importing :func:`lib` raises.
"""
from __future__ import annotations

import ctypes
from ctypes import POINTER, c_char_p, c_double, c_float, c_int, c_int32, c_int64, c_uint8, c_uint64, c_void_p
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libtitan_mi355x.so"

TITAN_OK = 0
TITAN_E_INVALID = -1
TITAN_E_HIP = -2
TITAN_E_NOMEM = -3
TITAN_E_NOTFOUND = -4
TITAN_E_INJECTED = -5

PASSIVE_SOFT, PASSIVE_STIFF, ACTIVE_CONTRACT_THEN_EXPAND, ACTIVE_EXPAND_THEN_CONTRACT, SPRING_GENERAL = range(5)
CONSTRAINT_PLANE, CONTACT_PLANE, BALL, DIRECTION = range(4)

# every symbol declared in include/titan_mi355x.h
EXPORTS = (
    "titan_ctx_create", "titan_ctx_destroy", "titan_last_error", "titan_set_dt", "titan_get_dt",
    "titan_set_global_acc", "titan_add_plane", "titan_add_ball", "titan_clear_planes_and_balls",
    "titan_body_create", "titan_body_destroy", "titan_body_count", "titan_step", "titan_step_until",
    "titan_time", "titan_sync", "titan_body_download", "titan_body_upload", "titan_timer_start",
    "titan_timer_stop", "titan_mem_used", "titan_kernel_name", "titan_lattice_spring_count",
    "titan_lattice", "titan_debug_inject_failure", "titan_body_layout", "titan_step_bodies",
    "titan_body_time", "titan_set_schedule", "titan_get_schedule", "titan_autotune",
    "titan_serve_config", "titan_serve_submit", "titan_serve_pump", "titan_serve_poll",
    "titan_serve_cancel", "titan_serve_stats", "titan_serve_reset_stats", "titan_set_integrator",
    "titan_get_integrator", "titan_debug_inject_failure_at", "titan_step_noalloc",
    "titan_measure_hbm", "titan_body_remove", "titan_debug_body_bytes", "titan_reserve",
)

REQ_QUEUED, REQ_RUNNING, REQ_DONE, REQ_CANCELLED, REQ_FAILED = range(5)

LAYOUT_AUTO, LAYOUT_GATHER, LAYOUT_TILED, LAYOUT_RESIDENT = 0, 1, 2, 3
RESIDENT_LDS_MAX = 1024   # resident bodies held in LDS (launches with few bodies)
RESIDENT_MAX = 32767      # the resident layout's largest body
RESIDENT_SPLIT_MAX = 256  # resident bodies of at most four chunks have the split form (the default schedule's bound)
# integrators: the reference's default Euler branch and its VERLET / RK2 build options
INTEGRATOR_EULER, INTEGRATOR_VERLET, INTEGRATOR_RK2 = 0, 1, 2
# call sites of titan_debug_inject_failure_at
INJECT_STEP, INJECT_DOWNLOAD, INJECT_DESTROY = 0, 1, 2


class TitanError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"titan error {code}: {msg}")
        self.code = code


class BodyDesc(ctypes.Structure):
    _fields_ = [
        ("n_masses", c_int64), ("n_springs", c_int64),
        ("pos", POINTER(c_double)), ("vel", POINTER(c_double)), ("m", POINTER(c_double)),
        ("T", POINTER(c_double)), ("ext", POINTER(c_double)), ("fixed", POINTER(c_uint8)),
        ("drag", POINTER(c_int32)), ("left", POINTER(c_int32)), ("right", POINTER(c_int32)),
        ("k", POINTER(c_double)), ("rest", POINTER(c_double)), ("damping", POINTER(c_double)),
        ("type", POINTER(c_int32)), ("omega", POINTER(c_double)), ("act_a", POINTER(c_double)),
        ("act_phi", POINTER(c_double)), ("lc_ptr", POINTER(c_int64)), ("lc_type", POINTER(c_int32)),
        ("lc_vec", POINTER(c_double)), ("lc_num", POINTER(c_double)), ("t0", c_double),
        ("layout", c_int32),
    ]


_lib = None


def lib() -> ctypes.CDLL:
    """Load libtitan_mi355x.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(f"{LIB_PATH} is not built; run `python -m titan_amd.build`")
    L = ctypes.CDLL(str(LIB_PATH))
    vp = c_void_p
    sig = {
        "titan_ctx_create": (c_int, [c_int, c_uint64, POINTER(vp)]),
        "titan_ctx_destroy": (c_int, [vp]),
        "titan_last_error": (c_char_p, [vp]),
        "titan_set_dt": (c_int, [vp, c_double]),
        "titan_get_dt": (c_int, [vp, POINTER(c_double)]),
        "titan_set_global_acc": (c_int, [vp, c_double, c_double, c_double]),
        "titan_add_plane": (c_int, [vp] + [c_double] * 6),
        "titan_add_ball": (c_int, [vp] + [c_double] * 4),
        "titan_clear_planes_and_balls": (c_int, [vp]),
        "titan_body_create": (c_int, [vp, POINTER(BodyDesc), POINTER(c_int32)]),
        "titan_body_destroy": (c_int, [vp, c_int32]),
        "titan_body_remove": (c_int, [vp, c_int32, vp, c_int64, vp, c_int64]),
        "titan_debug_body_bytes": (c_int, [vp, c_int32, c_int32, vp, c_int64, POINTER(c_int64)]),
        "titan_body_count": (c_int, [vp, POINTER(c_int32)]),
        "titan_step": (c_int, [vp, c_int64]),
        "titan_step_noalloc": (c_int, [vp, c_int64]),
        "titan_measure_hbm": (c_int, [c_int, c_uint64, c_int32, POINTER(c_double), POINTER(c_double)]),
        "titan_step_until": (c_int, [vp, c_double, POINTER(c_int64)]),
        "titan_time": (c_int, [vp, POINTER(c_double)]),
        "titan_sync": (c_int, [vp]),
        "titan_body_download": (c_int, [vp, c_int32] + [POINTER(c_double)] * 4),
        "titan_body_upload": (c_int, [vp, c_int32] + [POINTER(c_double)] * 5),
        "titan_timer_start": (c_int, [vp]),
        "titan_timer_stop": (c_int, [vp, POINTER(c_float)]),
        "titan_mem_used": (c_int, [vp, POINTER(c_uint64)]),
        "titan_kernel_name": (c_int, [vp, c_char_p, c_int64]),
        "titan_lattice_spring_count": (c_int64, [c_int, c_int, c_int]),
        "titan_lattice": (c_int, [c_double] * 6 + [c_int] * 3 + [POINTER(c_double), POINTER(c_int32),
                                                                   POINTER(c_int32), POINTER(c_double)]),
        "titan_debug_inject_failure": (c_int, [vp, c_int32]),
        "titan_debug_inject_failure_at": (c_int, [vp, c_int32, c_int32]),
        "titan_body_layout": (c_int, [vp, c_int32, POINTER(c_int32)]),
        "titan_step_bodies": (c_int, [vp, vp, vp, c_int32]),
        "titan_body_time": (c_int, [vp, c_int32, POINTER(c_double), POINTER(c_int64)]),
        "titan_set_schedule": (c_int, [vp, vp]),
        "titan_reserve": (c_int, [vp, c_int64, c_int64, c_int64, c_int32]),
        "titan_set_integrator": (c_int, [vp, c_int32]),
        "titan_get_integrator": (c_int, [vp, POINTER(c_int32)]),
        "titan_get_schedule": (c_int, [vp, vp]),
        "titan_autotune": (c_int, [vp, vp, c_int32, c_int32, c_int32, vp, vp]),
        "titan_serve_config": (c_int, [vp, c_int32, c_uint64, c_int64]),
        "titan_serve_submit": (c_int, [vp, vp, POINTER(c_int64)]),
        "titan_serve_pump": (c_int, [vp, POINTER(c_int32), POINTER(c_int32)]),
        "titan_serve_poll": (c_int, [vp, c_int64, POINTER(c_int32), POINTER(c_int64)]),
        "titan_serve_cancel": (c_int, [vp, c_int64]),
        "titan_serve_stats": (c_int, [vp, vp]),
        "titan_serve_reset_stats": (c_int, [vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def measure_hbm(device: int = 0, nbytes: int = 1 << 30, iters: int = 5) -> dict:
    """The library's streaming copy and read kernels on `device`: GB/s of a
    copy (bytes read + written) and of a read, over buffers of `nbytes`."""
    L = lib()
    cp, rd = c_double(), c_double()
    rc = L.titan_measure_hbm(int(device), int(nbytes), int(iters), ctypes.byref(cp), ctypes.byref(rd))
    if rc != 0:
        raise TitanError(rc, L.titan_last_error(None).decode())
    return {"copy_GBps": cp.value, "read_GBps": rd.value, "bytes": int(nbytes), "iters": int(iters)}


def _ptr(a, ctype):
    if a is None:
        return None
    return a.ctypes.data_as(POINTER(ctype))


def _f64(a, shape=None):
    if a is None:
        return None
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None:
        a = a.reshape(shape)
    return a


def _i32(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.int32)


def lattice(center, dims, nx: int, ny: int, nz: int):
    """Lattice::Lattice positions and springs (reference order): pos (N,3), left, right, rest."""
    L = lib()
    n = nx * ny * nz
    s = L.titan_lattice_spring_count(nx, ny, nz)
    pos = np.empty((n, 3))
    left = np.empty(s, np.int32)
    right = np.empty(s, np.int32)
    rest = np.empty(s)
    rc = L.titan_lattice(*map(float, center), *map(float, dims), nx, ny, nz, _ptr(pos, c_double),
                         _ptr(left, c_int32), _ptr(right, c_int32), _ptr(rest, c_double))
    if rc != 0:
        raise TitanError(rc, "titan_lattice failed")
    return pos, left, right, rest


def body_desc(pos, m, *, vel=None, T=None, ext=None, fixed=None, drag=None, left=None, right=None,
              k=None, rest=None, damping=None, type=None, omega=None, act_a=None, act_phi=None,
              lc_ptr=None, lc_type=None, lc_vec=None, lc_num=None, t0: float = 0.0,
              layout: int = LAYOUT_AUTO):
    """A titan_body_desc over contiguous copies of the given arrays.  Returns
    (desc, arrays): keep `arrays` alive while the descriptor is in use."""
    pos = _f64(pos, (-1, 3))
    n = pos.shape[0]
    m = _f64(np.broadcast_to(np.asarray(m, dtype=np.float64), (n,)))
    keep = dict(pos=pos, m=m, vel=_f64(vel, (-1, 3)), T=_f64(T), ext=_f64(ext, (-1, 3)),
                fixed=None if fixed is None else np.ascontiguousarray(fixed, dtype=np.uint8),
                drag=_i32(drag), left=_i32(left), right=_i32(right))
    s = 0 if keep["left"] is None else keep["left"].shape[0]
    for name, v in (("k", k), ("rest", rest), ("damping", damping), ("omega", omega),
                    ("act_a", act_a), ("act_phi", act_phi)):
        keep[name] = None if v is None else _f64(np.broadcast_to(np.asarray(v, dtype=np.float64), (s,)))
    keep["type"] = None if type is None else _i32(np.broadcast_to(np.asarray(type), (s,)))
    keep["lc_ptr"] = None if lc_ptr is None else np.ascontiguousarray(lc_ptr, dtype=np.int64)
    keep["lc_type"] = _i32(lc_type)
    keep["lc_vec"] = _f64(lc_vec, (-1, 3)) if lc_vec is not None else None
    keep["lc_num"] = _f64(lc_num)
    d = BodyDesc()
    d.n_masses = n
    d.n_springs = s
    d.t0 = float(t0)
    d.layout = int(layout)
    ctypes_of = dict(fixed=c_uint8, drag=c_int32, left=c_int32, right=c_int32, type=c_int32,
                     lc_ptr=c_int64, lc_type=c_int32)
    for name, arr in keep.items():
        setattr(d, name, _ptr(arr, ctypes_of.get(name, c_double)))
    return d, keep


class Request(ctypes.Structure):
    _fields_ = [("body", BodyDesc), ("horizon", c_int64), ("out_pos", POINTER(c_double)),
                ("out_vel", POINTER(c_double))]


class ServeReport(ctypes.Structure):
    _fields_ = [("submitted", c_int64), ("completed", c_int64), ("cancelled", c_int64), ("failed", c_int64),
                ("running", c_int64), ("queued", c_int64), ("p50_latency_s", c_double),
                ("p99_latency_s", c_double), ("max_latency_s", c_double), ("mass_steps", c_double),
                ("mass_steps_per_s", c_double), ("device_idle_s", c_double)]


class Schedule(ctypes.Structure):
    _fields_ = [("tiled_min_masses", c_int64), ("tile_rows", c_int32), ("resident_max_masses", c_int32)]


# titan_autotune's candidates (tiled_min_masses, tile_rows, resident_max_masses),
# in the order of its rate_out (min -1: never tiled)
AUTOTUNE_CANDIDATES = [(mn, rw, rs) for rs in (0, RESIDENT_SPLIT_MAX, RESIDENT_LDS_MAX, RESIDENT_MAX)
                       for mn in (0, 128, 256, 512, 1024, -1) for rw in (682, 2304)]


class Context:
    """One engine context on one GPU (ctypes wrapper of titan_ctx)."""

    def __init__(self, device: int = 0, hbm_budget: int = 0):
        self._L = lib()
        h = c_void_p()
        rc = self._L.titan_ctx_create(device, hbm_budget, ctypes.byref(h))
        if rc != 0:
            raise TitanError(rc, self._L.titan_last_error(None).decode())
        self._h = h
        self._bodies: dict[int, int] = {}
        self._serve_out: dict[int, tuple] = {}

    # -- plumbing --
    def _chk(self, rc: int):
        if rc != 0:
            raise TitanError(rc, self._L.titan_last_error(self._h).decode())

    def close(self):
        if getattr(self, "_h", None):
            self._L.titan_ctx_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- globals --
    def set_dt(self, dt: float):
        self._chk(self._L.titan_set_dt(self._h, dt))

    def set_integrator(self, kind: int):
        """INTEGRATOR_EULER, _VERLET or _RK2; before any body is created."""
        self._chk(self._L.titan_set_integrator(self._h, int(kind)))

    def integrator(self) -> int:
        v = c_int32()
        self._chk(self._L.titan_get_integrator(self._h, ctypes.byref(v)))
        return v.value

    def get_dt(self) -> float:
        v = c_double()
        self._chk(self._L.titan_get_dt(self._h, ctypes.byref(v)))
        return v.value

    def set_global_acc(self, g):
        self._chk(self._L.titan_set_global_acc(self._h, *map(float, g)))

    def add_plane(self, normal, offset, mu_k=0.0, mu_s=0.0):
        self._chk(self._L.titan_add_plane(self._h, *map(float, normal), float(offset), float(mu_k), float(mu_s)))

    def add_ball(self, center, r):
        self._chk(self._L.titan_add_ball(self._h, *map(float, center), float(r)))

    def clear_planes_and_balls(self):
        self._chk(self._L.titan_clear_planes_and_balls(self._h))

    # -- bodies --
    def add_body(self, pos, m, *, layout: int = LAYOUT_AUTO, **fields) -> int:
        """Create a body from titan_body_desc fields (see body_desc)."""
        d, keep = body_desc(pos, m, layout=layout, **fields)
        bid = c_int32()
        self._chk(self._L.titan_body_create(self._h, ctypes.byref(d), ctypes.byref(bid)))
        del keep
        self._bodies[bid.value] = d.n_masses
        return bid.value

    # -- schedule --
    def set_schedule(self, tiled_min_masses: int | None = None, tile_rows: int | None = None,
                     resident_max_masses: int | None = None):
        """Set the schedule for bodies created afterwards; an omitted field
        keeps the context's current value."""
        cur = self.schedule()
        vals = [cur[q] if v is None else int(v) for q, v in enumerate((tiled_min_masses, tile_rows, resident_max_masses))]
        s = Schedule(*vals)
        self._chk(self._L.titan_set_schedule(self._h, ctypes.byref(s)))

    def schedule(self) -> tuple[int, int, int]:
        """(tiled_min_masses, tile_rows, resident_max_masses)"""
        s = Schedule()
        self._chk(self._L.titan_get_schedule(self._h, ctypes.byref(s)))
        return s.tiled_min_masses, s.tile_rows, s.resident_max_masses

    def autotune(self, samples, replicas: int, steps: int):
        """Time the candidate schedules on `replicas` copies of the sample
        bodies (dicts of body fields); the fastest becomes this context's
        schedule.  Returns the winner (tiled_min_masses, tile_rows,
        resident_max_masses) and [(candidate, rate)] in AUTOTUNE_CANDIDATES
        order, with min -1 for "never tiled"."""
        built = [body_desc(**b) for b in samples]
        arr = (BodyDesc * len(built))(*[d for d, _ in built])
        best = Schedule()
        rates = (c_double * len(AUTOTUNE_CANDIDATES))()
        self._chk(self._L.titan_autotune(self._h, arr, len(built), int(replicas), int(steps),
                                         ctypes.byref(best), rates))
        del built
        return ((best.tiled_min_masses, best.tile_rows, best.resident_max_masses),
                list(zip(AUTOTUNE_CANDIDATES, list(rates))))

    # -- serving --
    def serve_config(self, max_running: int = 4096, hbm_bytes: int = 0, quantum: int = 256):
        self._chk(self._L.titan_serve_config(self._h, int(max_running), int(hbm_bytes), int(quantum)))

    def serve_submit(self, body: dict, horizon: int, want_result: bool = True) -> int:
        """Queue a rollout request (the body's arrays are copied); returns its ticket.
        With want_result, the final pos/vel land in arrays read by serve_result."""
        d, keep = body_desc(**body)
        req = Request()
        req.body = d
        req.horizon = int(horizon)
        out = None
        if want_result:
            out = (np.zeros((d.n_masses, 3)), np.zeros((d.n_masses, 3)))
            req.out_pos = _ptr(out[0], c_double)
            req.out_vel = _ptr(out[1], c_double)
        t = c_int64()
        self._chk(self._L.titan_serve_submit(self._h, ctypes.byref(req), ctypes.byref(t)))
        del keep
        self._serve_out[t.value] = out
        return t.value

    def serve_pump(self) -> tuple[int, int]:
        """One admit / step / retire round; returns (running, queued)."""
        r, q = c_int32(), c_int32()
        self._chk(self._L.titan_serve_pump(self._h, ctypes.byref(r), ctypes.byref(q)))
        return r.value, q.value

    def serve_poll(self, ticket: int) -> tuple[int, int]:
        st, n = c_int32(), c_int64()
        self._chk(self._L.titan_serve_poll(self._h, int(ticket), ctypes.byref(st), ctypes.byref(n)))
        return st.value, n.value

    def serve_result(self, ticket: int):
        """(pos, vel) of a completed request."""
        st, _ = self.serve_poll(ticket)
        if st != REQ_DONE:
            raise RuntimeError(f"request {ticket} is not done (state {st})")
        return self._serve_out[ticket]

    def serve_cancel(self, ticket: int):
        self._chk(self._L.titan_serve_cancel(self._h, int(ticket)))

    def serve_stats(self) -> dict:
        r = ServeReport()
        self._chk(self._L.titan_serve_stats(self._h, ctypes.byref(r)))
        return {name: getattr(r, name) for name, _ in ServeReport._fields_}

    def serve_reset_stats(self):
        self._chk(self._L.titan_serve_reset_stats(self._h))

    def destroy_body(self, bid: int):
        self._chk(self._L.titan_body_destroy(self._h, bid))
        self._bodies.pop(bid, None)

    def remove(self, bid: int, masses=(), springs=()):
        """Remove springs and masses of a live body (indices in its creation order)."""
        m = np.ascontiguousarray(masses, dtype=np.int64)
        s = np.ascontiguousarray(springs, dtype=np.int64)
        self._chk(self._L.titan_body_remove(self._h, int(bid), m.ctypes.data_as(c_void_p), int(m.size),
                                            s.ctypes.data_as(c_void_p), int(s.size)))

    def body_count(self) -> int:
        v = c_int32()
        self._chk(self._L.titan_body_count(self._h, ctypes.byref(v)))
        return v.value

    # -- stepping --
    def step(self, n: int):
        self._chk(self._L.titan_step(self._h, int(n)))

    def step_noalloc(self, n: int):
        """titan_step that allocates nothing (scratch from the library's static
        device array; no graph capture, no ball-grid rebuild)."""
        self._chk(self._L.titan_step_noalloc(self._h, int(n)))

    def step_bodies(self, ids, steps):
        """Advance each body in `ids` by its own number of steps (the others do not move)."""
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        steps = np.ascontiguousarray(np.broadcast_to(steps, ids.shape), dtype=np.int64)
        self._chk(self._L.titan_step_bodies(self._h, ids.ctypes.data_as(ctypes.c_void_p),
                                             steps.ctypes.data_as(ctypes.c_void_p), int(ids.size)))

    def body_time(self, body: int):
        """(T, steps) of one body's own clock."""
        t, n = c_double(), c_int64()
        self._chk(self._L.titan_body_time(self._h, int(body), ctypes.byref(t), ctypes.byref(n)))
        return t.value, n.value

    def step_until(self, t: float) -> int:
        k = c_int64()
        self._chk(self._L.titan_step_until(self._h, float(t), ctypes.byref(k)))
        return k.value

    def time(self) -> float:
        v = c_double()
        self._chk(self._L.titan_time(self._h, ctypes.byref(v)))
        return v.value

    def sync(self):
        self._chk(self._L.titan_sync(self._h))

    # -- transfer --
    def download(self, bid: int, pos=True, vel=True, acc=False, T=False) -> dict:
        n = self._bodies[bid]
        out = {}
        bufs = []
        for name, want, shape in (("pos", pos, (n, 3)), ("vel", vel, (n, 3)), ("acc", acc, (n, 3)), ("T", T, (n,))):
            a = np.empty(shape) if want else None
            if want:
                out[name] = a
            bufs.append(_ptr(a, c_double))
        self._chk(self._L.titan_body_download(self._h, bid, *bufs))
        return out

    def upload(self, bid: int, pos=None, vel=None, m=None, T=None, ext=None):
        arrs = [_f64(pos, (-1, 3)), _f64(vel, (-1, 3)), _f64(m), _f64(T), _f64(ext, (-1, 3))]
        self._chk(self._L.titan_body_upload(self._h, bid, *[_ptr(a, c_double) for a in arrs]))

    # -- timing / introspection --
    def timer_start(self):
        self._chk(self._L.titan_timer_start(self._h))

    def timer_stop(self) -> float:
        v = c_float()
        self._chk(self._L.titan_timer_stop(self._h, ctypes.byref(v)))
        return v.value

    def reserve(self, masses: int, springs: int, actuated: int = 0, bodies: int = 1):
        """Grow the arenas once to hold `masses` masses and `springs` springs
        (`actuated` of them actuated) in `bodies` bodies (titan_reserve)."""
        self._chk(self._L.titan_reserve(self._h, int(masses), int(springs), int(actuated), int(bodies)))

    def mem_used(self) -> int:
        v = c_uint64()
        self._chk(self._L.titan_mem_used(self._h, ctypes.byref(v)))
        return v.value

    def body_layout(self, bid: int) -> int:
        v = c_int32()
        self._chk(self._L.titan_body_layout(self._h, bid, ctypes.byref(v)))
        return v.value

    def kernel_name(self) -> str:
        buf = ctypes.create_string_buffer(256)
        self._chk(self._L.titan_kernel_name(self._h, buf, 256))
        return buf.value.decode()

    DEBUG_ITEMS = ("names", "rest", "k", "damping", "deg", "mflags", "state", "mass", "tiles", "halo", "chunks", "perm",
                   "tact")

    def layout_bytes(self, bid: int, item: str) -> bytes:
        """One item of a body's device layout (titan_debug_body_bytes), as raw bytes."""
        what = self.DEBUG_ITEMS.index(item)
        n = c_int64()
        self._chk(self._L.titan_debug_body_bytes(self._h, int(bid), what, None, 0, ctypes.byref(n)))
        buf = ctypes.create_string_buffer(max(1, n.value))
        self._chk(self._L.titan_debug_body_bytes(self._h, int(bid), what, buf, n.value, ctypes.byref(n)))
        return buf.raw[:n.value]

    def inject_failure(self, after: int, site: int = INJECT_STEP):
        """Make a call at `site` fail with TITAN_E_INJECTED after `after` successful ones."""
        self._chk(self._L.titan_debug_inject_failure_at(self._h, int(site), int(after)))
