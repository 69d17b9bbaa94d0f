"""Pins of the CPU oracle (oracle/oracle.cpp) against what the mathematics
fixes: closed forms, hand-computed branch cases, brute force on small
graphs, invariants, and the physical checks of the reference's own tests
(test/physics/*.cpp).  Each check is chosen so that a plausible restatement
mistake (a dropped term, a flipped sign, a swapped operand, the wrong force
fed to friction) fails it.  CPU only."""
from __future__ import annotations

import math
from fractions import Fraction

import numpy as np
import pytest

from oracle import oracle as orc

NORMAL = 20000.0  # src/object.cu:26


def one_mass(pos=(0.0, 0.0, 1.0), vel=(0.0, 0.0, 0.0), m=0.1, **kw):
    return orc.Scene(np.array([pos]), [m], vel=np.array([vel]), **kw)


# ---------------------------------------------------------------- Vec / pow

def test_pow_two_is_the_rounded_square():
    # The oracle writes pow(x, 2) (include/Titan/vec.h:150) as x * x, the
    # correctly rounded square, which is what gcc and clang fold pow(x, 2.0)
    # into.  (glibc's pow itself is 1 ulp off for about 1 input in 1700, so
    # the library routine is not the definition used.)
    rng = np.random.default_rng(0)
    xs = np.concatenate([rng.normal(size=2000), rng.uniform(-1e-3, 1e-3, 2000), rng.uniform(-1e3, 1e3, 2000)])
    for x in xs:
        x = float(x)
        assert x * x == float(Fraction(x) ** 2)


# ----------------------------------------------------- Euler + gravity + extern

@pytest.mark.parametrize("ext", [False, True])
def test_free_fall_matches_discrete_closed_form(ext):
    # external_unittest.cpp:17-57 scenario; symplectic Euler from rest gives
    # v_n = n a dt and x_n = x0 + a dt^2 n (n + 1) / 2 exactly in real arithmetic
    m, dt, n = 0.1, 1e-4, 10000
    if ext:
        s = one_mass((1.0, 0.0, 1.0), m=m, ext=np.array([[0.0, 0.0, -9.8 * m]]), g=(0, 0, 0), dt=dt)
    else:
        s = one_mass((1.0, 0.0, 1.0), m=m, g=(0, 0, -9.8), dt=dt)
    s.run(n)
    a = Fraction(-98, 10)
    z = Fraction(1) + a * Fraction(dt) ** 2 * n * (n + 1) / 2
    vz = a * n * Fraction(dt)
    assert abs(s.pos[0, 2] - float(z)) < 1e-9
    assert abs(s.vel[0, 2] - float(vz)) < 1e-9
    assert s.pos[0, 0] == 1.0 and s.pos[0, 1] == 0.0
    assert s.T[0] == pytest.approx(1.0, abs=1e-12)


def test_reference_external_test_bounds():
    # the reference's own assertions: x, y within 1e-6, z within 0.1 of
    # 1 - 4.9 t^2, sampled every 0.1 s (external_unittest.cpp:13-14,24-33)
    s = one_mass((1.0, 0.0, 1.0), g=(0, 0, -9.8), dt=1e-4)
    for _ in range(10):
        t = s.run(1000)
        assert abs(s.pos[0, 0] - 1) < 1e-6 and abs(s.pos[0, 1]) < 1e-6
        assert abs(s.pos[0, 2] - (1 - 4.9 * t * t)) < 0.1


# ------------------------------------------------------------ spring + damping

def _matrix_power_closed_form(k, c, m, dt, y0, n):
    # relative coordinate y = (xR - xL) - L0 of two equal masses: symplectic
    # Euler is (y, v) -> M (y, v), M = [[1 - w dt^2 - g dt ... ]] in exact arithmetic
    w2 = Fraction(2) * Fraction(k) / Fraction(m)
    g = Fraction(2) * Fraction(c) / Fraction(m)
    h = Fraction(dt)
    M = [[1 - w2 * h * h, h * (1 - g * h)], [-w2 * h, 1 - g * h]]
    y, v = Fraction(y0), Fraction(0)
    for _ in range(n):
        y, v = M[0][0] * y + M[0][1] * v, M[1][0] * y + M[1][1] * v
    return float(y), float(v)


@pytest.mark.parametrize("damping", [0.0, 0.3])
def test_two_mass_oscillator_matches_matrix_power(damping):
    k, m, dt, L0, d0, n = 100.0, 0.5, 1e-3, 1.0, 0.05, 400
    s = orc.Scene(np.array([[0.0, 0, 0], [L0 + d0, 0, 0]]), [m, m], left=[0], right=[1], k=[k],
                  rest=[L0], damping=[damping], g=(0, 0, 0), dt=dt)
    s.run(n)
    y, v = _matrix_power_closed_form(k, damping, m, dt, d0, n)
    assert abs((s.pos[1, 0] - s.pos[0, 0] - L0) - y) < 1e-11
    assert abs((s.vel[1, 0] - s.vel[0, 0]) - v) < 1e-9
    # action = reaction, exactly: the pair's momentum stays zero
    assert s.vel[0, 0] + s.vel[1, 0] == 0.0


# --------------------------------------------------------------- integrators

@pytest.mark.parametrize("integ", [orc.VERLET, orc.RK2])
def test_free_fall_closed_forms_of_the_other_integrators(integ):
    # from rest under a constant acceleration a:
    # VERLET (src/sim.cu:1156-1160, acc starts at 0): v_n = (n - 1/2) a dt and
    #   x_n = x0 + a dt^2 n (n + 1) / 2;
    # RK2 (src/sim.cu:1141-1155, two stages, T += dt/2 after each): the midpoint
    #   rule is exact here, v_n = n a dt and x_n = x0 + a dt^2 n^2 / 2
    m, dt, n = 0.1, 1e-4, 2000
    s = one_mass((1.0, 0.0, 1.0), m=m, g=(0, 0, -9.8), dt=dt, integrator=integ)
    s.run(n)
    a, h = Fraction(-98, 10), Fraction(dt)
    if integ == orc.VERLET:
        z, vz = 1 + a * h * h * n * (n + 1) / 2, a * h * (n - Fraction(1, 2))
    else:
        z, vz = 1 + a * h * h * n * n / 2, a * h * n
    assert abs(s.pos[0, 2] - float(z)) < 1e-12
    assert abs(s.vel[0, 2] - float(vz)) < 1e-12
    assert s.acc[0, 2] == -9.8 and s.pos[0, 0] == 1.0
    # the clock: RK2 adds dt/2 twice per step, to the scene time and the mass time
    t = 0.0
    for _ in range(n):
        if integ == orc.RK2:
            t += 0.5 * dt
            t += 0.5 * dt
        else:
            t += dt
    assert s.t == t and s.T[0] == t


def _integrator_map(integ, k, c, m, dt, y0, n):
    # relative coordinate y = (xR - xL) - L0 of two equal masses, y'' = a(y, v)
    # = -w2 y - g v, iterated in exact arithmetic with each integrator's own
    # update rule (Verlet carries the previous step's acceleration, 0 at first)
    w2, g, h = Fraction(2) * Fraction(k) / Fraction(m), Fraction(2) * Fraction(c) / Fraction(m), Fraction(dt)
    acc = lambda y, v: -w2 * y - g * v
    y, v, ap = Fraction(y0), Fraction(0), Fraction(0)
    for _ in range(n):
        if integ == orc.RK2:
            a1 = acc(y, v)
            ym, vm = y + h / 2 * v, v + h / 2 * a1
            a2 = acc(ym, vm)
            y, v = y + h * vm, v + h * a2
        else:
            ac = acc(y, v)
            v = v + (ap + ac) / 2 * h
            ap = ac
            y = y + v * h + ap / 2 * h * h
    return float(y), float(v), float(ap)


@pytest.mark.parametrize("integ", [orc.VERLET, orc.RK2])
@pytest.mark.parametrize("damping", [0.0, 0.3])
def test_two_mass_oscillator_follows_each_integrator(integ, damping):
    k, m, dt, L0, d0, n = 100.0, 0.5, 1e-3, 1.0, 0.05, 200
    s = orc.Scene(np.array([[0.0, 0, 0], [L0 + d0, 0, 0]]), [m, m], left=[0], right=[1], k=[k],
                  rest=[L0], damping=[damping], g=(0, 0, 0), dt=dt, integrator=integ)
    s.run(n)
    y, v, ap = _integrator_map(integ, k, damping, m, dt, d0, n)
    assert abs((s.pos[1, 0] - s.pos[0, 0] - L0) - y) < 1e-11
    assert abs((s.vel[1, 0] - s.vel[0, 0]) - v) < 1e-9
    if integ == orc.VERLET:
        assert abs((s.acc[1, 0] - s.acc[0, 0]) - ap) < 1e-7
    assert s.vel[0, 0] + s.vel[1, 0] == 0.0


def test_rk2_is_second_order_and_euler_first():
    # undamped oscillator against y0 cos(w t) at t = 0.5 s: halving dt divides
    # the error by about 4 for RK2 and about 2 for Euler (and for this Verlet
    # variant, whose first step starts from a zero acceleration)
    import math
    k, m, L0, d0 = 100.0, 0.5, 1.0, 0.05
    w = math.sqrt(2 * k / m)

    def err(integ, dt):
        n = int(round(0.5 / dt))
        s = orc.Scene(np.array([[0.0, 0, 0], [L0 + d0, 0, 0]]), [m, m], left=[0], right=[1], k=[k],
                      rest=[L0], g=(0, 0, 0), dt=dt, integrator=integ)
        s.run(n)
        return abs((s.pos[1, 0] - s.pos[0, 0] - L0) - d0 * math.cos(w * n * dt))

    for integ, lo, hi in [(orc.RK2, 3.6, 4.4), (orc.EULER, 1.8, 2.2), (orc.VERLET, 1.8, 2.2)]:
        r = err(integ, 1e-3) / err(integ, 5e-4)
        assert lo < r < hi, (integ, r)


def test_spring_force_direction_and_sign():
    # stretched spring pulls the right endpoint back toward the left one
    s = orc.Scene(np.array([[0.0, 0, 0], [0.0, 2.0, 0]]), [1.0, 1.0], left=[0], right=[1], k=[10.0],
                  rest=[1.5], g=(0, 0, 0), dt=1e-3)
    F = s.spring_forces()
    assert F[0, 1] == 10.0 * (1.5 - 2.0)  # k (rest - |temp|) * unit(temp), temp = R - L = +y
    assert F[0, 0] == 0.0 and F[0, 2] == 0.0
    # damping term: dot(vL - vR, u) * damping * u
    s2 = orc.Scene(np.array([[0.0, 0, 0], [0.0, 1.5, 0]]), [1.0, 1.0], vel=np.array([[0, 1.0, 0], [0, 3.0, 0]]),
                   left=[0], right=[1], k=[10.0], rest=[1.5], damping=[0.25], g=(0, 0, 0), dt=1e-3)
    F2 = s2.spring_forces()
    assert F2[0, 1] == (1.0 - 3.0) * 0.25


def test_momentum_invariant_random_graph():
    rng = np.random.default_rng(5)
    n, s = 60, 300
    pos = rng.uniform(-1, 1, (n, 3))
    L, R = rng.integers(0, n, s), rng.integers(0, n, s)
    keep = L != R
    L, R = L[keep], R[keep]
    d = pos[R] - pos[L]
    rest = np.sqrt((d * d).sum(1)) * rng.uniform(0.8, 1.2, len(L))
    sc = orc.Scene(pos, np.full(n, 0.2), vel=rng.normal(0, 0.1, (n, 3)), left=L, right=R,
                   k=rng.uniform(10, 100, len(L)), rest=rest, damping=rng.uniform(0, 0.2, len(L)),
                   g=(0, 0, 0), dt=1e-3)
    p0 = (sc.vel * 0.2).sum(0)
    sc.run(200)
    p1 = (sc.vel * 0.2).sum(0)
    assert np.max(np.abs(p1 - p0)) < 1e-12



def test_norm_and_dot_on_a_general_direction():
    # the Vec ops under the spring law (include/Titan/vec.h:109-159): norm of
    # (3, 4, 12) is 13 exactly, the unit vector has three unequal components
    # and dot(vL - vR, u) = (3 - 8 + 6) / 13 mixes all three, so a dropped or
    # swapped component, dot(a, a) for dot(a, b) or a reversed difference
    # changes the force
    k, rest, c = 10.0, 11.0, 0.5
    s = orc.Scene(np.array([[0.0, 0, 0], [3.0, 4.0, 12.0]]), [1.0, 1.0],
                  vel=np.array([[1.0, -2.0, 0.5], [0.0, 0.0, 0.0]]), left=[0], right=[1], k=[k],
                  rest=[rest], damping=[c], g=(0, 0, 0), dt=1e-3)
    F = s.spring_forces()[0]
    u = [Fraction(3, 13), Fraction(4, 13), Fraction(12, 13)]
    dotv = Fraction(1, 13)
    exact = [Fraction(k) * (Fraction(rest) - 13) * ui + dotv * Fraction(c) * ui for ui in u]
    for got, want in zip(F, exact):
        assert got == pytest.approx(float(want), rel=2e-16, abs=0)
    # the damping term alone (at rest length) and its sign
    s2 = orc.Scene(np.array([[0.0, 0, 0], [3.0, 4.0, 12.0]]), [1.0, 1.0],
                   vel=np.array([[1.0, -2.0, 0.5], [0.0, 0.0, 0.0]]), left=[0], right=[1], k=[k],
                   rest=[13.0], damping=[c], g=(0, 0, 0), dt=1e-3)
    F2 = s2.spring_forces()[0]
    for got, ui in zip(F2, u):
        assert got == pytest.approx(float(dotv * Fraction(c) * ui), rel=2e-16, abs=0)
        assert got > 0

# ------------------------------------------------------ scatter == gather order

def test_gather_equals_serial_scatter_bit_for_bit():
    rng = np.random.default_rng(6)
    n, s = 80, 500
    pos = rng.uniform(-1, 1, (n, 3))
    L, R = rng.integers(0, n, s), rng.integers(0, n, s)
    keep = L != R
    L, R = L[keep], R[keep]
    L[:5] = R[:5] + 0  # duplicate pairs are ordinary springs
    L[:5] = (R[:5] + 1) % n
    d = pos[R] - pos[L]
    rest = np.sqrt((d * d).sum(1)) * rng.uniform(0.8, 1.2, len(L))
    fixed = (np.arange(n) % 7 == 0).astype(np.uint8)
    kw = dict(vel=rng.normal(0, 0.1, (n, 3)), left=L, right=R, k=rng.uniform(10, 100, len(L)),
              rest=rest, damping=rng.uniform(0, 0.2, len(L)), fixed=fixed,
              planes=[(0.0, 0.0, 1.0, -0.5, 0.4, 0.6)], dt=1e-3)
    a = orc.Scene(pos, np.full(n, 0.2), **kw)
    b = orc.Scene(pos, np.full(n, 0.2), **kw)
    for _ in range(50):
        a.scatter_step()
    b.run(50)
    np.testing.assert_array_equal(a.pos, b.pos)
    np.testing.assert_array_equal(a.vel, b.vel)
    # fixed masses never move (src/sim.cu:1100-1101)
    np.testing.assert_array_equal(b.pos[fixed == 1], pos[fixed == 1])


# --------------------------------------------------------------- contact plane

def _step_acc(sc):
    sc.run(1)
    return sc.acc[0]


def test_plane_penalty_and_static_hold_at_rest():
    m, d0 = 0.1, -0.002
    sc = one_mass((0.0, 0.0, d0), m=m, planes=[(0, 0, 1, 0.0, 0.8, 1.0)], dt=1e-4)
    acc = _step_acc(sc)
    assert acc[2] == (m * -9.81 + (-d0 * NORMAL)) / m
    assert acc[0] == 0.0 and acc[1] == 0.0


def test_plane_no_force_above():
    sc = one_mass((0.0, 0.0, 0.001), planes=[(0, 0, 1, 0.0, 0.8, 1.0)], dt=1e-4)
    assert _step_acc(sc)[2] == -9.81 * 0.1 / 0.1


def test_plane_kinetic_friction_uses_pre_penalty_normal_force():
    m, d0, mu_k = 0.1, -0.003, 0.5
    sc = one_mass((0.0, 0.0, d0), vel=(2.0, 0.0, 0.0), m=m, planes=[(0, 0, 1, 0.0, mu_k, 1.0)], dt=1e-4)
    acc = _step_acc(sc)
    # |f_n| = m g, not m g + penalty (src/object.cu:77-84)
    assert acc[0] == pytest.approx(-mu_k * 9.81, rel=1e-14)
    assert acc[2] == pytest.approx((m * -9.81 - d0 * NORMAL) / m, rel=1e-14)


@pytest.mark.parametrize("mu_s,held", [(1.0, True), (0.2, False)])
def test_plane_static_branch(mu_s, held):
    m = 0.1
    sc = one_mass((0.0, 0.0, -0.001), m=m, ext=np.array([[0.3, 0.0, 0.0]]),
                  planes=[(0, 0, 1, 0.0, 0.4, mu_s)], dt=1e-4)
    acc = _step_acc(sc)
    assert acc[0] == (0.0 if held else 0.3 / m)


def test_plane_frictionless_skips_friction_branch():
    sc = one_mass((0.0, 0.0, -0.001), vel=(1.0, 0, 0), planes=[(0, 0, 1, 0.0, 0.0, 0.0)], dt=1e-4)
    assert _step_acc(sc)[0] == 0.0


def test_reference_simple_test_energy_within_one_percent():
    # simple_unittest.cpp:15-38: a mass dropped from 1 m on a frictionless
    # plane, g = 9.8, dt = 1e-4; EMA(0.9) of the energy within 1% over 5 s,
    # sampled every 0.1 s.  The reference's energy() (utils.h:12-30) leaves out
    # the penalty spring's elastic energy, so a sample that lands inside a
    # contact would dip by up to 35%; here that term (NORMAL disp^2 / 2) is
    # added, and the total is held to the reference's 1%.
    sc = one_mass((1.0, 0.0, 1.0), m=0.1, planes=[(0, 0, 1, 0.0, 0.0, 0.0)], g=(0, 0, -9.8), dt=1e-4)

    def total():
        z = sc.pos[0, 2]
        return sc.energy(9.8) + (0.5 * NORMAL * z * z if z < 0 else 0.0)

    e0 = total()
    avg = e0
    contact_samples = 0
    for _ in range(50):
        sc.run(1000)
        contact_samples += sc.pos[0, 2] < 0
        avg = 0.1 * total() + 0.9 * avg
        assert abs(avg - e0) <= 0.01 * e0
    assert contact_samples >= 1  # the check does cover a sample inside a contact



def test_energy_two_springs_hand_computed():
    # orc_energy (test/testutil/utils.h:12-30): sum of g z m, m |v|^2 / 2 and
    # k (|R - L| - rest)^2 / 2 over live springs; every term below is exact
    # in binary, so the sum is compared for equality.  Spring 0: |(3, 4, 12)|
    # = 13, rest 10, k 8 -> 36; spring 1: length 2, rest 1, k 6 -> 3; spring 2
    # has a missing endpoint and adds nothing.  Kinetic: 2 * 9 / 2 + 1 * 49 / 2
    # = 33.5; gravity (g = 10): 10 (1 * 2 + 13 * 0.5 + 15 * 1) = 235.
    pos = np.array([[0.0, 0.0, 1.0], [3.0, 4.0, 13.0], [3.0, 4.0, 15.0]])
    vel = np.array([[1.0, 2.0, 2.0], [0.0, 0.0, 0.0], [2.0, 3.0, 6.0]])
    s = orc.Scene(pos, [2.0, 0.5, 1.0], vel=vel, left=[0, 1, -1], right=[1, 2, 0], k=[8.0, 6.0, 100.0],
                  rest=[10.0, 1.0, 0.5], g=(0, 0, 0), dt=1e-3)
    assert s.energy(10.0) == 36.0 + 3.0 + 33.5 + 235.0
    # swapping a spring's endpoints leaves its energy unchanged; k enters once
    s2 = orc.Scene(pos, [2.0, 0.5, 1.0], vel=vel, left=[1, 2, -1], right=[0, 1, 0], k=[16.0, 6.0, 100.0],
                   rest=[10.0, 1.0, 0.5], g=(0, 0, 0), dt=1e-3)
    assert s2.energy(10.0) == s.energy(10.0) + 36.0

# ----------------------------------------------------------------------- ball

def test_ball_constant_magnitude_inside_zero_outside():
    sc = one_mass((0.0, 0.3, 0.4), m=2.0, g=(0, 0, 0), balls=[(0.0, 0.0, 0.0, 1.0)], dt=1e-4)
    acc = _step_acc(sc)
    assert acc[1] == (0.3 * NORMAL / 0.5) / 2.0 and acc[2] == (0.4 * NORMAL / 0.5) / 2.0
    sc2 = one_mass((0.0, 0.0, 1.5), m=2.0, g=(0, 0, 0), balls=[(0.0, 0.0, 0.0, 1.0)], dt=1e-4)
    assert np.all(_step_acc(sc2) == 0.0)


# ---------------------------------------------- local constraints and drag

def test_constraint_plane_hand_computed():
    sc = one_mass((0, 0, 0), vel=(1.0, 0.0, 0.5), m=1.0, ext=np.array([[1.0, 0.0, -2.0]]), g=(0, 0, 0),
                  lc_ptr=[0, 1], lc_type=[orc.CONSTRAINT_PLANE], lc_vec=[(0.0, 0.0, 1.0)], lc_num=[0.5],
                  dt=1e-3)
    sc.run(1)
    # normal force -2 removed; vel projected to (1, 0, 0); friction adds
    # (-0.5 * -2) * vel / |vel_before| (src/object.cu:115-124)
    fx = 1.0 + (1.0 * (-0.5 * -2.0)) / math.sqrt(1.25)
    assert sc.acc[0, 0] == fx and sc.acc[0, 2] == 0.0
    assert sc.vel[0, 0] == 1.0 + fx * 1e-3 and sc.vel[0, 2] == 0.0


def test_direction_hand_computed():
    sc = one_mass((0, 0, 0), vel=(1.0, 2.0, 0.0), m=1.0, ext=np.array([[3.0, 4.0, 0.0]]), g=(0, 0, 0),
                  lc_ptr=[0, 1], lc_type=[orc.DIRECTION], lc_vec=[(1.0, 0.0, 0.0)], lc_num=[0.25], dt=1e-3)
    sc.run(1)
    # force off the tangent (0, 4, 0) removed, vel -> (1, 0, 0), friction
    # -|(0,4,0)| * 0.25 along t (src/object.cu:133-141)
    assert sc.acc[0, 0] == 3.0 - 4.0 * 0.25 and sc.acc[0, 1] == 0.0
    assert sc.vel[0, 1] == 0.0


def test_local_contact_plane_is_frictionless_and_kinds_apply_in_reference_order():
    base = dict(vel=(1.0, 0.0, 0.0), m=0.1)
    a = one_mass((0, 0, -0.01), **base, lc_ptr=[0, 1], lc_type=[orc.CONTACT_PLANE],
                 lc_vec=[(0.0, 0.0, 1.0)], lc_num=[0.0], dt=1e-4)
    a.run(1)
    assert a.acc[0, 0] == 0.0  # no friction on a local contact plane (object.cu:58-63)
    # insertion order [direction, contact plane] must apply as [contact plane, direction]
    kw = dict(lc_vec=[(0.0, 0.0, 1.0), (1.0, 0.0, 0.0)], lc_num=[0.0, 0.1], dt=1e-4)
    b = one_mass((0, 0, -0.01), **base, lc_ptr=[0, 2], lc_type=[orc.CONTACT_PLANE, orc.DIRECTION], **kw)
    kw2 = dict(lc_vec=[(1.0, 0.0, 0.0), (0.0, 0.0, 1.0)], lc_num=[0.1, 0.0], dt=1e-4)
    c = one_mass((0, 0, -0.01), **base, lc_ptr=[0, 2], lc_type=[orc.DIRECTION, orc.CONTACT_PLANE], **kw2)
    b.run(1)
    c.run(1)
    np.testing.assert_array_equal(b.acc, c.acc)
    assert b.acc[0, 2] == 0.0  # the direction constraint removed the contact force


def test_integer_drag():
    sc = one_mass((0, 0, 0), vel=(3.0, 4.0, 0.0), m=1.0, g=(0, 0, 0), drag=[2], dt=1e-4)
    sc.run(1)
    assert sc.acc[0, 0] == (3.0 * (-2.0 * 25.0)) / 5.0 and sc.acc[0, 1] == (4.0 * (-2.0 * 25.0)) / 5.0


# ---------------------------------------------------------------- actuation

def test_actuation_scale_and_general_law_mapping():
    w, t = 2.0, math.pi / 4  # w t = pi / 2
    rest = 1.0
    s = math.sin(w * t)
    for typ, scale in ((orc.CONTRACT_THEN_EXPAND, 1 - 0.2 * s), (orc.EXPAND_THEN_CONTRACT, 1 + 0.2 * s)):
        sc = orc.Scene(np.array([[0.0, 0, 0], [0.0, 0, 2.0]]), [1.0, 1.0], left=[0], right=[1], k=[10.0],
                       rest=[rest], type=[typ], omega=[w], g=(0, 0, 0), dt=1e-3)
        F = sc.spring_forces(t)
        assert F[0, 2] == 10.0 * (rest * scale - 2.0)
    # the general law with a = -/+0.2, phi = 0 is bit-identical to kinds 2 / 3
    rng = np.random.default_rng(7)
    n = 40
    pos = rng.uniform(-1, 1, (n, 3))
    L, R = np.arange(n - 1), np.arange(1, n)
    om = rng.uniform(1, 20, n - 1)
    ref_t = rng.integers(2, 4, n - 1)
    common = dict(vel=rng.normal(0, 0.1, (n, 3)), left=L, right=R, k=50.0, rest=0.3, damping=0.05, omega=om,
                  g=(0, 0, -9.81), dt=1e-3)
    a = orc.Scene(pos, np.full(n, 0.1), type=ref_t, **common)
    b = orc.Scene(pos, np.full(n, 0.1), type=np.full(n - 1, orc.GENERAL), act_a=np.where(ref_t == 2, -0.2, 0.2),
                  act_phi=np.zeros(n - 1), **common)
    a.run(300)
    b.run(300)
    np.testing.assert_array_equal(a.pos, b.pos)


# ----------------------------------------------------------------- lattice

@pytest.mark.parametrize("n,expected", [(5, 1036), (10, 10476), (20, 93556), (100, 12731796)])
def test_lattice_spring_counts(n, expected):
    # 13 half-directions of the 26-neighbourhood: closed form
    # 3 n^2 (n-1) + 6 n (n-1)^2 + 4 (n-1)^3
    assert orc.lattice_spring_count(n, n, n) == expected
    assert expected == 3 * n * n * (n - 1) + 6 * n * (n - 1) ** 2 + 4 * (n - 1) ** 3


@pytest.mark.parametrize("dims", [(4, 5, 6), (3, 3, 3), (1, 4, 5), (2, 1, 3)])
def test_lattice_topology_is_the_26_neighbourhood(dims):
    nx, ny, nz = dims
    pos, L, R, rest = orc.lattice((1.0, -2.0, 3.0), (2.0, 3.0, 4.0), nx, ny, nz)
    idx = {}
    for i in range(nx):
        for j in range(ny):
            for k in range(nz):
                idx[k + j * nz + i * ny * nz] = (i, j, k)
    pairs = set()
    for l, r in zip(L, R):
        a, b = idx[int(l)], idx[int(r)]
        assert max(abs(a[c] - b[c]) for c in range(3)) == 1
        key = frozenset((int(l), int(r)))
        assert key not in pairs
        pairs.add(key)
    expect = sum(1 for a in idx for b in idx if a < b and max(abs(idx[a][c] - idx[b][c]) for c in range(3)) == 1)
    assert len(pairs) == expect
    # positions: (i/(nx-1) - 0.5) * dims + center, 0 on a degenerate axis
    for a, (i, j, k) in idx.items():
        fx = i / (nx - 1) - 0.5 if nx > 1 else 0.0
        assert pos[a, 0] == fx * 2.0 + 1.0
    d = pos[R] - pos[L]
    np.testing.assert_array_equal(rest, np.sqrt(d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2]))


def test_breakpoint_step_count_rule():
    # the driver stops at the first iteration with breakpoint <= T
    # (src/sim.cu:1505); T accumulates by += dt, so 0.0105 is reached after
    # 106 steps of 1e-4, not 105 (105 additions give 0.010499999...)
    T, k = 0.0, 0
    while not (0.0105 <= T):
        T += 1e-4
        k += 1
    assert k == 106
    sc = one_mass(dt=1e-4)
    assert sc.run(k - 1) < 0.0105
    assert sc.run(1) == T
