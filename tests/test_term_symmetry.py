"""The identity the resident kernel's split form rests on, checked in IEEE
double arithmetic on the CPU: a spring's term formed from endpoint b's side
(axis pa - pb, velocity difference vb - va) is exactly the negation of the
term formed from a's side, except for the sign of an exact zero, and
x - (-G) is x + G bit for bit.  The term is formed as the step forms it
(src/sim.cu:1055-1065): u = t / |t|, G = u k (rest - |t|) + u ((dv . u) d),
with |t| = sqrt((tx tx + ty ty) + tz tz)."""
from __future__ import annotations

import numpy as np


def _term(t, dv, k, rest, damp):
    n = np.sqrt(t[:, 0] * t[:, 0] + t[:, 1] * t[:, 1] + t[:, 2] * t[:, 2])
    u = t / n[:, None]
    mag = k * (rest - n)
    dd = (dv[:, 0] * u[:, 0] + dv[:, 1] * u[:, 1] + dv[:, 2] * u[:, 2]) * damp
    return u * mag[:, None] + u * dd[:, None]


def _same_up_to_zero_sign(x, y):
    return np.array_equal(x, y)  # == treats +0 and -0 as equal, and nothing else


def test_term_from_the_other_side_is_the_negation():
    rng = np.random.default_rng(0)
    n = 200_000
    pa = rng.normal(0, 1, (n, 3))
    pb = pa + rng.normal(0, 0.05, (n, 3))
    pb[::7, 0] = pa[::7, 0]           # axis components that are exactly zero
    pb[::11, 2] = pa[::11, 2]
    va, vb = rng.normal(0, 1, (n, 3)), rng.normal(0, 1, (n, 3))
    vb[::5, 1] = va[::5, 1]           # and velocity differences that are
    k = np.exp(rng.uniform(np.log(1e3), np.log(1e5), n))
    rest = np.sqrt(((pa - pb) ** 2).sum(1)) * rng.uniform(0.9, 1.1, n)
    damp = rng.uniform(0, 0.3, n)
    g_a = _term(pb - pa, va - vb, k, rest, damp)
    g_b = _term(pa - pb, vb - va, k, rest, damp)
    assert _same_up_to_zero_sign(g_b, -g_a)
    # and the sums a mass forms agree: f - G at a's side is f + (-G); at b's
    # side the walk subtracts its own term, which is -G
    f = rng.normal(0, 10, (n, 3))
    assert _same_up_to_zero_sign(f - g_b, f + g_a)


def test_a_plausible_slip_breaks_it():
    # the check is sharp: forming b's axis as pb - pa (the a side's axis)
    # while keeping b's velocity difference gives a different term
    rng = np.random.default_rng(1)
    n = 1000
    pa, pb = rng.normal(0, 1, (n, 3)), rng.normal(0, 1, (n, 3))
    va, vb = rng.normal(0, 1, (n, 3)), rng.normal(0, 1, (n, 3))
    k, rest, damp = np.full(n, 1e4), np.ones(n), np.full(n, 0.1)
    g_a = _term(pb - pa, va - vb, k, rest, damp)
    wrong = _term(pb - pa, vb - va, k, rest, damp)
    assert not _same_up_to_zero_sign(wrong, -g_a)
