"""A second, independent restatement of the reference's step in plain Python
floats: TEST INFRASTRUCTURE ONLY (it writes the golden fixtures of this
directory and checks the C++ oracle; nothing in the product imports it).

It is written from the reference's source text, not from oracle/oracle.cpp:
the spring kernel (src/sim.cu:1046-1080) runs serially over the springs in
index order, adding each spring's force to its right endpoint and its
negation to its left one (a serial run of the atomic scatter), and the mass
kernel (src/sim.cu:1090-1169) then applies gravity, the external force, the
global planes and balls, the local constraints and drag, and the integrator.
Python's float is an IEEE double, its +, -, *, / and math.sqrt are correctly
rounded, and every expression below keeps the reference's C++ evaluation
order (operator precedence, left to right; Vec * double multiplies each
component by the scalar, double * Vec likewise; pow(x, 2) is the correctly
rounded x * x).  Pure-Python loops: for scenes of a few hundred masses."""
from __future__ import annotations

import math

NORMAL = 20000.0  # src/object.cu:26

PASSIVE_SOFT, PASSIVE_STIFF, CONTRACT_THEN_EXPAND, EXPAND_THEN_CONTRACT, GENERAL = range(5)
CONSTRAINT_PLANE, CONTACT_PLANE, BALL, DIRECTION = range(4)
EULER, VERLET, RK2 = range(3)


# include/Titan/vec.h:109-159, src/vec.cu:39-41
def add(a, b):
    return (a[0] + b[0], a[1] + b[1], a[2] + b[2])


def sub(a, b):
    return (a[0] - b[0], a[1] - b[1], a[2] - b[2])


def neg(a):
    return (-a[0], -a[1], -a[2])


def mul(v, x):  # Vec * double and double * Vec: Vec(v[i] * x)
    return (v[0] * x, v[1] * x, v[2] * x)


def div(v, x):
    return (v[0] / x, v[1] / x, v[2] / x)


def norm(v):  # sqrt(pow(x, 2) + pow(y, 2) + pow(z, 2))
    return math.sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2])


def dot(a, b):  # (a * b).sum()
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]


def lattice(center, dims, nx, ny, nz):
    """Lattice::Lattice (src/object.cu:232-293): positions, spring endpoints
    (left, right) in creation order, rest lengths |right - left|."""
    def idx(i, j, k):
        return k + j * nz + i * ny * nz

    pos = [None] * (nx * ny * nz)
    for i in range(nx):
        for j in range(ny):
            for k in range(nz):
                f = ((i / (nx - 1.0) - 0.5) if nx > 1 else 0.0,
                     (j / (ny - 1.0) - 0.5) if ny > 1 else 0.0,
                     (k / (nz - 1.0) - 0.5) if nz > 1 else 0.0)
                # Vec(...) * dims is element-wise, then + center
                pos[idx(i, j, k)] = add((f[0] * dims[0], f[1] * dims[1], f[2] * dims[2]), center)
    springs = []
    for i in range(nx):
        for j in range(ny):
            for k in range(nz):
                for l in range(2 if i != nx - 1 else 1):
                    for m in range(2 if j != ny - 1 else 1):
                        for n in range(2 if k != nz - 1 else 1):
                            if l or m or n:
                                springs.append((idx(i, j, k), idx(i + l, j + m, k + n)))
                if k != nz - 1:
                    if j != ny - 1:
                        springs.append((idx(i, j, k + 1), idx(i, j + 1, k)))
                    if i != nx - 1:
                        springs.append((idx(i, j, k + 1), idx(i + 1, j, k)))
                    if j != ny - 1 and i != nx - 1:
                        springs.append((idx(i, j, k + 1), idx(i + 1, j + 1, k)))
                        springs.append((idx(i + 1, j, k + 1), idx(i, j + 1, k)))
                        springs.append((idx(i, j + 1, k + 1), idx(i + 1, j, k)))
                if j != ny - 1 and i != nx - 1:
                    springs.append((idx(i, j + 1, k), idx(i + 1, j, k)))
    left = [s[0] for s in springs]
    right = [s[1] for s in springs]
    rest = [norm(sub(pos[r], pos[l])) for l, r in springs]
    return pos, left, right, rest


class World:
    """Masses, springs and constraints as the reference's CUDA_MASS /
    CUDA_SPRING records hold them, stepped as Simulation::execute does
    (src/sim.cu:1569-1590)."""

    def __init__(self, pos, m, left, right, k, rest, damping, *, vel=None, g=(0.0, 0.0, -9.81), dt=1e-4,
                 planes=(), balls=(), fixed=None, drag=None, ext=None, types=None, omega=None, act_a=None,
                 act_phi=None, local=None, integrator=EULER):
        n = len(pos)
        self.pos = [tuple(map(float, p)) for p in pos]
        self.vel = [(0.0, 0.0, 0.0)] * n if vel is None else [tuple(map(float, v)) for v in vel]
        self.acc = [(0.0, 0.0, 0.0)] * n
        self.m = [float(x) for x in m]
        self.T = [0.0] * n
        self.fixed = [False] * n if fixed is None else [bool(x) for x in fixed]
        self.drag = [0] * n if drag is None else [int(x) for x in drag]  # int (include/Titan/object.h:213)
        self.ext = [(0.0, 0.0, 0.0)] * n if ext is None else [tuple(map(float, e)) for e in ext]
        # per mass: list of (kind, vec, number) in insertion order
        self.local = [[] for _ in range(n)] if local is None else local
        s = len(left)
        self.left, self.right = list(left), list(right)
        self.k, self.rest, self.damping = [float(x) for x in k], [float(x) for x in rest], [float(x) for x in damping]
        self.types = [PASSIVE_SOFT] * s if types is None else list(types)
        self.omega = [0.0] * s if omega is None else [float(x) for x in omega]
        self.act_a = [0.0] * s if act_a is None else [float(x) for x in act_a]
        self.act_phi = [0.0] * s if act_phi is None else [float(x) for x in act_phi]
        self.g = tuple(map(float, g))
        self.dt = float(dt)
        self.planes = [tuple(map(float, p)) for p in planes]  # (nx, ny, nz, offset, mu_k, mu_s), normal unit
        self.balls = [tuple(map(float, b)) for b in balls]    # (cx, cy, cz, r)
        self.integrator = integrator
        self.t = 0.0
        self.backup_pos = list(self.pos)
        self.backup_vel = list(self.vel)

    # computeSpringForces, src/sim.cu:1046-1080, one spring after another
    def spring_forces(self, t):
        force = [(0.0, 0.0, 0.0)] * len(self.pos)
        for i in range(len(self.left)):
            L, R = self.left[i], self.right[i]
            if L < 0 or R < 0:
                continue
            temp = sub(self.pos[R], self.pos[L])
            scale = 1.0
            typ = self.types[i]
            if typ == CONTRACT_THEN_EXPAND:
                scale = 1 - 0.2 * math.sin(self.omega[i] * t)
            elif typ == EXPAND_THEN_CONTRACT:
                scale = 1 + 0.2 * math.sin(self.omega[i] * t)
            elif typ == GENERAL:  # this build's general law; a = -/+0.2, phi = 0 gives the two kinds above
                scale = 1 + self.act_a[i] * math.sin(self.omega[i] * t + self.act_phi[i])
            f = mul(div(temp, norm(temp)), self.k[i] * (self.rest[i] * scale - norm(temp)))
            f = add(f, mul(div(temp, norm(temp)), dot(sub(self.vel[L], self.vel[R]), div(temp, norm(temp))) *
                           self.damping[i]))
            if not self.fixed[R]:
                force[R] = add(force[R], f)
            if not self.fixed[L]:
                force[L] = add(force[L], neg(f))
        return force

    # CudaContactPlane::applyForce, src/object.cu:73-106
    @staticmethod
    def contact_plane(pos, vel, force, n, offset, mu_k, mu_s):
        disp = dot(pos, n) - offset
        f_normal = mul(n, dot(force, n))
        if disp < 0 and (mu_s > 0 or mu_k > 0):
            v_perp = sub(vel, mul(n, dot(vel, n)))
            v_norm = norm(v_perp)
            if v_norm > 1e-16:
                friction_mag = mu_k * norm(f_normal)
                force = sub(force, div(mul(v_perp, friction_mag), v_norm))
            else:
                f_perp = sub(force, f_normal)
                if mu_s * norm(f_normal) > norm(f_perp):
                    force = sub(force, f_perp)
        contact = mul(n, -disp * NORMAL) if disp < 0 else mul(n, 0)
        return add(force, contact)

    # CudaBall::applyForce, src/object.cu:53-56
    @staticmethod
    def ball(pos, force, c, r):
        dist = norm(sub(pos, c))
        return add(force, div(mul(sub(pos, c), NORMAL), dist) if dist <= r else (0.0, 0.0, 0.0))

    # CudaConstraintPlane::applyForce, src/object.cu:115-124
    @staticmethod
    def constraint_plane(vel, force, n, friction):
        normal_force = dot(force, n)
        force = add(force, mul(neg(n), normal_force))
        v_norm = norm(vel)
        if v_norm >= 1e-16:
            vel = add(vel, mul(neg(n), dot(vel, n)))
            force = add(force, div(mul(vel, -friction * normal_force), v_norm))
        return vel, force

    # CudaDirection::applyForce, src/object.cu:133-141
    @staticmethod
    def direction(vel, force, tg, friction):
        normal_force = sub(force, mul(tg, dot(force, tg)))
        force = add(force, neg(normal_force))
        if norm(vel) >= 1e-16:
            vel = mul(tg, dot(vel, tg))
            force = add(force, mul(tg, -norm(normal_force) * friction))
        return vel, force

    # massForcesAndUpdate, src/sim.cu:1090-1169 (stage: the RK2 template flag, 0 = first half)
    def mass_update(self, i, force, stage):
        if self.fixed[i]:
            return
        pos, vel, m = self.pos[i], self.vel[i], self.m[i]
        force = add(force, mul(self.g, m))
        force = add(force, self.ext[i])
        for p in self.planes:
            force = self.contact_plane(pos, vel, force, p[:3], p[3], p[4], p[5])
        for b in self.balls:
            force = self.ball(pos, force, b[:3], b[3])
        for kind in (CONTACT_PLANE, BALL, CONSTRAINT_PLANE, DIRECTION):
            for kd, v, num in self.local[i]:
                if kd != kind:
                    continue
                if kind == CONTACT_PLANE:  # local planes carry no friction (src/object.cu:58-63)
                    force = self.contact_plane(pos, vel, force, v, num, 0.0, 0.0)
                elif kind == BALL:
                    force = self.ball(pos, force, v, num)
                elif kind == CONSTRAINT_PLANE:
                    vel, force = self.constraint_plane(vel, force, v, num)
                else:
                    vel, force = self.direction(vel, force, v, num)
        if norm(vel) != 0.0:
            nv = norm(vel)
            force = add(force, div(mul(vel, -self.drag[i] * (nv * nv)), nv))
        dt = self.dt
        if self.integrator == RK2:
            acc = div(force, m)
            if stage == 0:
                self.backup_vel[i] = vel
                self.backup_pos[i] = pos
                pos = add(pos, mul(mul(vel, 0.5), dt))
                vel = add(vel, mul(mul(acc, 0.5), dt))
            else:
                pos = add(self.backup_pos[i], mul(vel, dt))
                vel = add(self.backup_vel[i], mul(acc, dt))
            self.T[i] += 0.5 * dt
        elif self.integrator == VERLET:
            vel = add(vel, mul(mul(add(self.acc[i], div(force, m)), 0.5), dt))
            acc = div(force, m)
            pos = add(pos, add(mul(vel, dt), mul(mul(acc, 0.5), dt * dt)))
            self.T[i] += dt
        else:
            acc = div(force, m)
            vel = add(vel, mul(acc, dt))
            pos = add(pos, mul(vel, dt))
            self.T[i] += dt
        self.acc[i], self.vel[i], self.pos[i] = acc, vel, pos

    def launch(self, stage):
        force = self.spring_forces(self.t)
        for i in range(len(self.pos)):
            self.mass_update(i, force[i], stage)

    def step(self, n=1):
        for _ in range(n):
            if self.integrator == RK2:
                self.launch(0)
                self.t += 0.5 * self.dt
                self.launch(1)
                self.t += 0.5 * self.dt
            else:
                self.launch(0)
                self.t += self.dt
