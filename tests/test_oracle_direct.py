"""Pins the oracle's restatement of the reference's device direct sum (grav_cuda.cu:65-141): the
float-narrowed default form and the periodic box, against accel_brute (itself bit for bit with the
reference, tests/test_oracle.py) and against what the geometry fixes. CPU only."""
from __future__ import annotations

import numpy as np

from oracle.bindings import HostParticles
from tests import tree_cases


def _rel(a, b):
    a = np.stack(a, 1)
    b = np.stack(b, 1)
    return np.linalg.norm(a - b, axis=1) / np.linalg.norm(b, axis=1)


def test_fp64_open_box_is_accel_brute(oracle):
    x, y, z, m, soft = tree_cases.plummer_small(1500, seed=5, eps=0.02)
    hp = HostParticles(x, y, z, m, soft)
    tg = np.arange(0, 1500, 7, dtype=np.uint32)
    a = oracle.direct_sum(hp, tg, G=2.0, fp64=True, box=0.0)
    b = oracle.accel_brute(hp, tg, G=2.0)
    for u, v in zip(a, b):
        assert np.array_equal(u, v)


def test_float_form_within_float_rounding_of_fp64(oracle):
    x, y, z, m, soft = tree_cases.plummer_small(3000, seed=6, eps=0.01)
    soft = soft.copy()
    soft[::5] = 0.0  # some unsoftened particles: the larger softening of the pair applies
    hp = HostParticles(x, y, z, m, soft)
    tg = np.arange(0, 3000, 3, dtype=np.uint32)
    e = _rel(oracle.direct_sum(hp, tg, fp64=False), oracle.accel_brute(hp, tg))
    # a float sum of 3000 terms: well under 1e-4, yet far above zero (the arithmetic IS float)
    assert e.max() < 1e-4
    assert np.median(e) > 1e-9


def test_float_form_two_body(oracle):
    """Two particles: the float form equals the FP64 spline (pinned to the reference) to float
    rounding in each branch, with the larger softening applying to both."""
    for d in (0.01, 0.03, 0.07, 0.2, 1.5):  # u < 1/2, 1/2 <= u < 1, and Newtonian, for h = 0.1
        x = np.array([0.3, 0.3 + d])
        y = np.array([-0.2, -0.2])
        z = np.array([0.1, 0.1])
        m = np.array([1.0, 0.5])
        soft = np.array([0.1, 0.04])
        hp = HostParticles(x, y, z, m, soft)
        ax, ay, az = oracle.direct_sum(hp, np.array([0, 1], np.uint32), G=1.0, fp64=False)
        g = oracle.spline_force_over_r(d, 0.1)
        np.testing.assert_allclose(ax, [0.5 * g * d, -1.0 * g * d], rtol=2e-6)
        assert ay[0] == 0.0 and az[1] == 0.0


def test_float_form_skips_self_and_coincident(oracle):
    x = np.array([0.0, 0.0, 1.0])
    y = np.zeros(3)
    z = np.zeros(3)
    m = np.array([1.0, 2.0, 3.0])
    soft = np.array([0.1, 0.1, 0.1])
    hp = HostParticles(x, y, z, m, soft)
    ax, ay, az = oracle.direct_sum(hp, np.array([0, 2], np.uint32), fp64=False)
    # particle 0 sees only particle 2 (1 is on top of it); 2 sees 0 and 1 at r = 1 (Newtonian)
    np.testing.assert_allclose(ax, [3.0, -3.0], rtol=1e-6)


def _explicit_images(x, y, z, t, L):
    """Sources moved to their nearest image around target t (numpy arithmetic)."""
    out = []
    for c in (x, y, z):
        d = c - c[t]
        d = np.where(d > 0.5 * L, d - L, np.where(d < -0.5 * L, d + L, d))
        out.append(c[t] + d)
    return out


def test_periodic_box_equals_explicit_images(oracle):
    rng = np.random.default_rng(8)
    n, L = 600, 2.0
    x, y, z = (rng.uniform(0, L, n) for _ in range(3))
    m = rng.uniform(0.5, 1.5, n) / n
    soft = np.full(n, 0.05)
    hp = HostParticles(x, y, z, m, soft)
    tg = np.array([0, 17, 123, 599], np.uint32)
    got64 = oracle.direct_sum(hp, tg, fp64=True, box=L)
    got32 = oracle.direct_sum(hp, tg, fp64=False, box=L)
    for k, t in enumerate(tg):
        xi, yi, zi = _explicit_images(x, y, z, int(t), L)
        want = oracle.accel_brute(HostParticles(xi, yi, zi, m, soft), np.array([t], np.uint32))
        w = np.array([v[0] for v in want])
        g64 = np.array([v[k] for v in got64])
        g32 = np.array([v[k] for v in got32])
        assert np.linalg.norm(g64 - w) <= 1e-12 * np.linalg.norm(w)
        assert np.linalg.norm(g32 - w) <= 1e-4 * np.linalg.norm(w)
    # and the box matters: the open sum differs
    op = oracle.direct_sum(hp, tg, fp64=True, box=0.0)
    assert not np.allclose(op[0], got64[0])
