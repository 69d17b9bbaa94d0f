"""Seeded synthetic scenes for the tests and the benchmark.

Each builder returns a ``Scene`` holding one or more bodies as plain arrays
(the fields of ``titan_body_desc``) plus the global settings, so the same
scene can be handed to the engine (:meth:`Scene.load`) and, in the tests, to
the CPU oracle.  Lattices come from the engine's own generator, which keeps
the reference's mass and spring order (src/object.cu:232-293).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import Context, lattice

# Titan defaults: Mass(pos) has m = 0.1 (include/Titan/mass.h:18), Spring has
# k = 10000 and damping 0 (include/Titan/spring.h:21-25), dt = 1e-4 and
# g = (0, 0, -9.81) (src/sim.cu:77-86).
DEFAULT_MASS = 0.1
DEFAULT_K = 10000.0


@dataclass
class Scene:
    bodies: list = field(default_factory=list)      # dicts of titan_body_desc fields
    planes: list = field(default_factory=list)      # (nx, ny, nz, d, mu_k, mu_s)
    balls: list = field(default_factory=list)       # (cx, cy, cz, r)
    g: tuple = (0.0, 0.0, -9.81)
    dt: float = 1e-4
    name: str = ""

    @property
    def n_masses(self) -> int:
        return sum(len(b["pos"]) for b in self.bodies)

    @property
    def n_springs(self) -> int:
        return sum(len(b.get("left", ())) for b in self.bodies)

    @property
    def n_actuated(self) -> int:
        n = 0
        for b in self.bodies:
            t = b.get("type")
            if t is not None:
                n += int(np.count_nonzero(np.asarray(t) >= 2))
        return n

    def load_globals(self, ctx: Context):
        """Set the scene's time step, gravity, planes and balls on ctx."""
        ctx.set_dt(self.dt)
        ctx.set_global_acc(self.g)
        for p in self.planes:
            ctx.add_plane(p[:3], p[3], p[4], p[5])
        for b in self.balls:
            ctx.add_ball(b[:3], b[3])

    def load(self, ctx: Context, layout=0) -> list[int]:
        """Set the scene's globals on ctx and create its bodies; `layout` is one
        layout for every body or a list with one per body."""
        self.load_globals(ctx)
        lays = list(layout) if isinstance(layout, (list, tuple)) else [layout] * len(self.bodies)
        return [ctx.add_body(**b, layout=l) for b, l in zip(self.bodies, lays)]

    def algorithmic_bytes_per_step(self) -> int:
        """104 B per mass + 32 B per spring + 24 B per actuated spring (+24 B per
        mass when any extern force is set): read pos, vel, m; write pos, vel;
        read both endpoints and k, rest, damping once per spring."""
        ext = any(b.get("ext") is not None for b in self.bodies)
        return 104 * self.n_masses + 32 * self.n_springs + 24 * self.n_actuated + (24 * self.n_masses if ext else 0)


def rest_lengths(pos, left, right):
    """|left - right| per spring (Spring::defaultLength, src/spring.cu:14)."""
    d = pos[left] - pos[right]
    return np.sqrt(d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2])


def lattice_body(center, dims, n, *, m=DEFAULT_MASS, k=DEFAULT_K, damping=0.0):
    nx, ny, nz = (n, n, n) if np.isscalar(n) else n
    pos, left, right, rest = lattice(center, dims, nx, ny, nz)
    return dict(pos=pos, m=np.full(len(pos), m), left=left, right=right, k=np.full(len(left), k),
                rest=rest, damping=np.full(len(left), damping))


def readme_lattice(center_z: float = 10.0) -> Scene:
    """The README scene: a 5x5x5 lattice spanning a 5 m cube centred at (0, 0, 10),
    m = 0.1, k = 1e4, damping 0.1, contact plane z = 0 with mu_k = 0.8 and
    mu_s = 1.0, g = 9.81, dt = 1e-4.  ``center_z = 2.45`` puts the bottom face
    5 cm into the plane so contact and friction act from the first step."""
    b = lattice_body((0.0, 0.0, center_z), (5.0, 5.0, 5.0), 5, damping=0.1)
    return Scene(bodies=[b], planes=[(0.0, 0.0, 1.0, 0.0, 0.8, 1.0)], dt=1e-4,
                 name=f"lattice5_cz{center_z:g}")


def big_lattice(n: int = 100, seed: int = 0, dt: float = 1e-5) -> Scene:
    """The 1M-mass benchmark scene: an n^3 lattice spanning 5 m centred at
    (0, 0, 2.5) (bottom layer at z = 0), positions jittered by N(0, 1e-4) m from
    numpy's default_rng(seed), rest lengths = jittered distances, k = 1e4,
    damping 0.1, contact plane z = 0 with mu_k = 0.8, mu_s = 1.0, dt = 1e-5."""
    pos, left, right, _ = lattice((0.0, 0.0, 2.5), (5.0, 5.0, 5.0), n, n, n)
    rng = np.random.default_rng(seed)
    pos = pos + rng.normal(0.0, 1e-4, size=pos.shape)
    rest = rest_lengths(pos, left, right)
    b = dict(pos=pos, m=np.full(len(pos), DEFAULT_MASS), left=left, right=right,
             k=np.full(len(left), DEFAULT_K), rest=rest, damping=np.full(len(left), 0.1))
    return Scene(bodies=[b], planes=[(0.0, 0.0, 1.0, 0.0, 0.8, 1.0)], dt=dt, name=f"lattice{n}_seed{seed}")


def robot_population(n_bodies: int = 4096, side_lo: int = 3, side_hi: int = 12, frac_act: float = 0.3,
                     seed: int = 0, dt: float = 1e-4, spacing: float = 2.0) -> Scene:
    """Uneven soft-robot population: lattice bodies with side ~ U{side_lo..side_hi},
    1 m cubes, each with its own (frictionless, as the reference's local contact
    planes are) ground plane z = 0; a fraction of springs actuated with
    rest * (1 + a sin(omega t + phi)), a ~ U(0, 0.2), omega ~ U(1, 20) rad/s,
    phi ~ U(0, 2 pi); k = 1e4, damping 0.1."""
    rng = np.random.default_rng(seed)
    sides = rng.integers(side_lo, side_hi + 1, size=n_bodies)
    grid = int(np.ceil(np.sqrt(n_bodies)))
    bodies = []
    for i, s in enumerate(sides):
        cx, cy = spacing * (i % grid), spacing * (i // grid)
        b = lattice_body((cx, cy, 0.6), (1.0, 1.0, 1.0), int(s), damping=0.1)
        ns = len(b["left"])
        act = rng.random(ns) < frac_act
        b["type"] = np.where(act, 4, 0).astype(np.int32)
        b["omega"] = np.where(act, rng.uniform(1.0, 20.0, ns), 0.0)
        b["act_a"] = np.where(act, rng.uniform(0.0, 0.2, ns), 0.0)
        b["act_phi"] = np.where(act, rng.uniform(0.0, 2 * np.pi, ns), 0.0)
        n = len(b["pos"])
        b["lc_ptr"] = np.arange(n + 1, dtype=np.int64)
        b["lc_type"] = np.full(n, 1, np.int32)                   # CONTACT_PLANE
        b["lc_vec"] = np.tile([0.0, 0.0, 1.0], (n, 1))
        b["lc_num"] = np.zeros(n)
        bodies.append(b)
    return Scene(bodies=bodies, dt=dt, name=f"robots{n_bodies}_seed{seed}")


def knn_cloud(n: int, *, k_nn: int = 12, box: float = 10.0, seed: int = 0, k_lo: float = 1e3,
              k_hi: float = 1e5, damping: float = 0.05, m: float = DEFAULT_MASS, z0: float = 0.0):
    """A seeded uniform point cloud of n masses in a box (side `box`, bottom at
    z0) joined by springs to each point's k_nn nearest neighbours, deduplicated
    (pair i < j once, in ascending (i, j) order), so degrees vary; rest =
    initial distance, k ~ LogUniform(k_lo, k_hi), damping constant.  Returns
    the body dict."""
    from scipy.spatial import cKDTree

    rng = np.random.default_rng(seed)
    pos = rng.uniform(0.0, box, size=(n, 3))
    pos[:, 0] -= box / 2
    pos[:, 1] -= box / 2
    pos[:, 2] += z0
    kk = min(k_nn, n - 1)
    _, nbr = cKDTree(pos).query(pos, k=kk + 1, workers=-1)
    i = np.repeat(np.arange(n), kk)
    j = nbr[:, 1:].reshape(-1)
    a, b = np.minimum(i, j), np.maximum(i, j)
    pairs = np.unique(a.astype(np.int64) * n + b)
    left = (pairs // n).astype(np.int32)
    right = (pairs % n).astype(np.int32)
    rest = rest_lengths(pos, left, right)
    ks = np.exp(rng.uniform(np.log(k_lo), np.log(k_hi), size=len(left)))
    return dict(pos=pos, m=np.full(n, m), left=left, right=right, k=ks, rest=rest,
                damping=np.full(len(left), damping))


def actuate(body: dict, frac: float, rng) -> dict:
    """Make a fraction of a body's springs actuated with the general law
    rest * (1 + a sin(omega t + phi)), a ~ U(0, 0.2), omega ~ U(1, 20), phi ~ U(0, 2 pi)."""
    ns = len(body["left"])
    act = rng.random(ns) < frac
    body["type"] = np.where(act, 4, 0).astype(np.int32)
    body["omega"] = np.where(act, rng.uniform(1.0, 20.0, ns), 0.0)
    body["act_a"] = np.where(act, rng.uniform(0.0, 0.2, ns), 0.0)
    body["act_phi"] = np.where(act, rng.uniform(0.0, 2 * np.pi, ns), 0.0)
    return body


def serving_requests(n: int, seed: int = 0, side_lo: int = 3, side_hi: int = 40, knn_lo: float = 1e2,
                     knn_hi: float = 1e5, h_lo: int = 100, h_hi: int = 10000):
    """Rollout requests: each a seeded body, half lattices of side U{side_lo..side_hi}
    (1 m cubes, damping 0.1) and half kNN-12 clouds of LogUniform(knn_lo, knn_hi)
    masses (1 m box), with 10-30% of springs actuated and a horizon of
    U{h_lo..h_hi} steps.  Returns [(body, horizon)]."""
    rng = np.random.default_rng(seed)
    out = []
    for q in range(n):
        if rng.random() < 0.5:
            s = int(rng.integers(side_lo, side_hi + 1))
            b = lattice_body((0.0, 0.0, 0.6), (1.0, 1.0, 1.0), s, damping=0.1)
        else:
            nm = int(np.exp(rng.uniform(np.log(knn_lo), np.log(knn_hi))))
            b = knn_cloud(nm, box=1.0, seed=int(rng.integers(1 << 30)), z0=0.1)
        actuate(b, rng.uniform(0.1, 0.3), rng)
        out.append((b, int(rng.integers(h_lo, h_hi + 1))))
    return out


def knn_scene(n: int = 2_000_000, n_balls: int = 1000, seed: int = 0, dt: float = 1e-5) -> Scene:
    """The irregular-graph workload: a seeded uniform cloud of n masses in a
    10 m box (bottom at z = 0), kNN-12 springs deduplicated (variable
    degree), k ~ LogUniform(1e3, 1e5), damping 0.05, n_balls rigid contact
    balls (centres uniform in the box, radius U(0.1, 0.5) m), a ground plane
    z = 0 with friction, dt = 1e-5."""
    b = knn_cloud(n, box=10.0, seed=seed)
    rng = np.random.default_rng(seed + 7)
    c = rng.uniform(0.0, 10.0, size=(n_balls, 3))
    c[:, :2] -= 5.0
    balls = [(float(x), float(y), float(z), float(r))
             for (x, y, z), r in zip(c, rng.uniform(0.1, 0.5, n_balls))]
    return Scene(bodies=[b], planes=[(0.0, 0.0, 1.0, 0.0, 0.8, 1.0)], balls=balls, dt=dt,
                 name=f"knn{n}_balls{n_balls}_seed{seed}")
