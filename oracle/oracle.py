"""ctypes view of the CPU oracle (oracle/liboracle.so) — TEST INFRASTRUCTURE ONLY.

Imported only by tests/, __graft_entry__.smoke() and the cpu_baseline leg of
bench.py, always as what the HIP path's results are compared with or as the
timed CPU baseline, never as the path under test.  The scene is described with the same arrays the engine's
titan_body_desc takes (positions and velocities as (N, 3), springs as
parallel arrays), so one scene can be handed to both.
"""
from __future__ import annotations

import ctypes
import subprocess
from ctypes import POINTER, c_double, c_int, c_int8, c_int32, c_int64, c_uint8
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"

PASSIVE_SOFT, PASSIVE_STIFF, CONTRACT_THEN_EXPAND, EXPAND_THEN_CONTRACT, GENERAL = range(5)
CONSTRAINT_PLANE, CONTACT_PLANE, BALL, DIRECTION = range(4)
EULER, VERLET, RK2 = range(3)


class _Scene(ctypes.Structure):
    _fields_ = [
        ("n_masses", c_int64), ("n_springs", c_int64),
        ("px", POINTER(c_double)), ("py", POINTER(c_double)), ("pz", POINTER(c_double)),
        ("vx", POINTER(c_double)), ("vy", POINTER(c_double)), ("vz", POINTER(c_double)),
        ("ax", POINTER(c_double)), ("ay", POINTER(c_double)), ("az", POINTER(c_double)),
        ("T", POINTER(c_double)), ("m", POINTER(c_double)),
        ("ex", POINTER(c_double)), ("ey", POINTER(c_double)), ("ez", POINTER(c_double)),
        ("fixed", POINTER(c_uint8)), ("drag", POINTER(c_int32)),
        ("left", POINTER(c_int32)), ("right", POINTER(c_int32)),
        ("k", POINTER(c_double)), ("rest", POINTER(c_double)), ("damping", POINTER(c_double)),
        ("type", POINTER(c_int32)), ("omega", POINTER(c_double)),
        ("act_a", POINTER(c_double)), ("act_phi", POINTER(c_double)),
        ("n_planes", c_int32), ("plane", POINTER(c_double)),
        ("n_balls", c_int32), ("ball", POINTER(c_double)),
        ("lc_ptr", POINTER(c_int64)), ("lc_type", POINTER(c_int32)),
        ("lc_vec", POINTER(c_double)), ("lc_num", POINTER(c_double)),
        ("gx", c_double), ("gy", c_double), ("gz", c_double), ("dt", c_double),
        ("n_threads", c_int32), ("integrator", c_int32), ("rk2_backup", POINTER(c_double)),
    ]


_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = ctypes.CDLL(str(LIB_PATH))
        P = POINTER(_Scene)
        L.orc_run.argtypes = [P, c_double, c_int64, POINTER(c_double)]
        L.orc_run.restype = c_int
        L.orc_scatter_step.argtypes = [P, c_double]
        L.orc_scatter_step.restype = c_int
        L.orc_spring_forces.argtypes = [P, c_double, POINTER(c_double)]
        L.orc_spring_forces.restype = c_int
        L.orc_build_incidence.argtypes = [P, POINTER(c_int64), POINTER(c_int32), POINTER(c_int8)]
        L.orc_build_incidence.restype = c_int
        L.orc_lattice_springs.argtypes = [c_int, c_int, c_int]
        L.orc_lattice_springs.restype = c_int64
        L.orc_lattice.argtypes = [c_double] * 6 + [c_int] * 3 + [POINTER(c_double), POINTER(c_int32),
                                                                  POINTER(c_int32), POINTER(c_double)]
        L.orc_lattice.restype = c_int
        L.orc_energy.argtypes = [P, c_double]
        L.orc_energy.restype = c_double
        _lib = L
    return _lib


def _p(a, t):
    return None if a is None else a.ctypes.data_as(POINTER(t))


class Scene:
    """A mutable scene the oracle steps in place (state held as SoA numpy arrays)."""

    def __init__(self, pos, m, *, vel=None, T=None, ext=None, fixed=None, drag=None,
                 left=None, right=None, k=None, rest=None, damping=None, type=None, omega=None,
                 act_a=None, act_phi=None, planes=(), balls=(), lc_ptr=None, lc_type=None,
                 lc_vec=None, lc_num=None, g=(0.0, 0.0, -9.81), dt=1e-4, threads=0,
                 integrator=EULER):
        pos = np.asarray(pos, dtype=np.float64).reshape(-1, 3)
        n = pos.shape[0]
        self.n = n
        self.p = [np.ascontiguousarray(pos[:, c]) for c in range(3)]
        v = np.zeros((n, 3)) if vel is None else np.asarray(vel, dtype=np.float64).reshape(-1, 3)
        self.v = [np.ascontiguousarray(v[:, c]) for c in range(3)]
        self.a = [np.zeros(n) for _ in range(3)]
        self.T = np.zeros(n) if T is None else np.array(T, dtype=np.float64)
        self.m = np.ascontiguousarray(np.broadcast_to(np.asarray(m, dtype=np.float64), (n,)))
        self.e = None if ext is None else [np.ascontiguousarray(np.asarray(ext, dtype=np.float64).reshape(-1, 3)[:, c]) for c in range(3)]
        self.fixed = None if fixed is None else np.ascontiguousarray(fixed, dtype=np.uint8)
        self.drag = None if drag is None else np.ascontiguousarray(drag, dtype=np.int32)
        s = 0 if left is None else len(left)
        self.s = s

        def f(x):
            return None if x is None else np.ascontiguousarray(np.broadcast_to(np.asarray(x, dtype=np.float64), (s,)))

        self.left = np.ascontiguousarray(left if left is not None else np.zeros(0), dtype=np.int32)
        self.right = np.ascontiguousarray(right if right is not None else np.zeros(0), dtype=np.int32)
        self.k, self.rest = f(k if k is not None else np.zeros(s)), f(rest if rest is not None else np.zeros(s))
        self.damping = f(damping if damping is not None else 0.0)
        self.type = None if type is None else np.ascontiguousarray(np.broadcast_to(np.asarray(type), (s,)), dtype=np.int32)
        self.omega, self.act_a, self.act_phi = f(omega), f(act_a), f(act_phi)
        self.planes = np.ascontiguousarray(np.asarray(planes, dtype=np.float64).reshape(-1, 6))
        self.balls = np.ascontiguousarray(np.asarray(balls, dtype=np.float64).reshape(-1, 4))
        self.lc_ptr = None if lc_ptr is None else np.ascontiguousarray(lc_ptr, dtype=np.int64)
        self.lc_type = None if lc_type is None else np.ascontiguousarray(lc_type, dtype=np.int32)
        self.lc_vec = None if lc_vec is None else np.ascontiguousarray(np.asarray(lc_vec, dtype=np.float64).reshape(-1))
        self.lc_num = None if lc_num is None else np.ascontiguousarray(lc_num, dtype=np.float64)
        self.g = tuple(map(float, g))
        self.dt = float(dt)
        self.threads = int(threads)
        self.integrator = int(integrator)
        self.rk2_backup = np.zeros(6 * n + 6)
        self.t = 0.0

    def _struct(self) -> _Scene:
        S = _Scene()
        S.n_masses, S.n_springs = self.n, self.s
        S.px, S.py, S.pz = (_p(x, c_double) for x in self.p)
        S.vx, S.vy, S.vz = (_p(x, c_double) for x in self.v)
        S.ax, S.ay, S.az = (_p(x, c_double) for x in self.a)
        S.T, S.m = _p(self.T, c_double), _p(self.m, c_double)
        if self.e is not None:
            S.ex, S.ey, S.ez = (_p(x, c_double) for x in self.e)
        S.fixed, S.drag = _p(self.fixed, c_uint8), _p(self.drag, c_int32)
        S.left, S.right = _p(self.left, c_int32), _p(self.right, c_int32)
        S.k, S.rest, S.damping = _p(self.k, c_double), _p(self.rest, c_double), _p(self.damping, c_double)
        S.type, S.omega = _p(self.type, c_int32), _p(self.omega, c_double)
        S.act_a, S.act_phi = _p(self.act_a, c_double), _p(self.act_phi, c_double)
        S.n_planes, S.plane = self.planes.shape[0], _p(self.planes, c_double)
        S.n_balls, S.ball = self.balls.shape[0], _p(self.balls, c_double)
        S.lc_ptr, S.lc_type = _p(self.lc_ptr, c_int64), _p(self.lc_type, c_int32)
        S.lc_vec, S.lc_num = _p(self.lc_vec, c_double), _p(self.lc_num, c_double)
        S.gx, S.gy, S.gz = self.g
        S.dt = self.dt
        S.n_threads = self.threads
        S.integrator = self.integrator
        S.rk2_backup = _p(self.rk2_backup, c_double)
        return S

    def run(self, n_steps: int) -> float:
        """n steps with T accumulated by += dt (RK2: += 0.5 dt per stage) from
        the scene's current time."""
        S = self._struct()
        t = c_double()
        rc = lib().orc_run(ctypes.byref(S), self.t, int(n_steps), ctypes.byref(t))
        if rc != 0:
            raise RuntimeError("orc_run failed")
        self.t = t.value
        return self.t

    def scatter_step(self):
        S = self._struct()
        if lib().orc_scatter_step(ctypes.byref(S), self.t) != 0:
            raise RuntimeError("orc_scatter_step failed")
        self.t += self.dt

    def spring_forces(self, t=None) -> np.ndarray:
        S = self._struct()
        f = np.empty((self.s, 3))
        lib().orc_spring_forces(ctypes.byref(S), self.t if t is None else t, _p(f, c_double))
        return f

    def energy(self, g_abs: float) -> float:
        S = self._struct()
        return lib().orc_energy(ctypes.byref(S), g_abs)

    @property
    def pos(self) -> np.ndarray:
        return np.stack(self.p, axis=1)

    @property
    def vel(self) -> np.ndarray:
        return np.stack(self.v, axis=1)

    @property
    def acc(self) -> np.ndarray:
        return np.stack(self.a, axis=1)


def lattice(center, dims, nx, ny, nz):
    """Oracle restatement of Lattice::Lattice: pos (N,3), left, right, rest."""
    L = lib()
    n = nx * ny * nz
    s = L.orc_lattice_springs(nx, ny, nz)
    pos = np.empty((n, 3))
    left = np.empty(s, np.int32)
    right = np.empty(s, np.int32)
    rest = np.empty(s)
    L.orc_lattice(*map(float, center), *map(float, dims), nx, ny, nz, _p(pos, c_double),
                  _p(left, c_int32), _p(right, c_int32), _p(rest, c_double))
    return pos, left, right, rest


def lattice_spring_count(nx, ny, nz) -> int:
    return lib().orc_lattice_springs(nx, ny, nz)
