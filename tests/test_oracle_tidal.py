"""Tidal tensor and jerk of the grouped walk (tree.cc:811-871) in the oracle: bit for bit against
the reference, and against finite differences of the direct-sum acceleration — in the target's
position (the tidal tensor is the gradient of the acceleration, without G) and along ballistic
motion of every particle (the jerk is its time derivative)."""
from __future__ import annotations

import numpy as np
import pytest

from oracle.bindings import HostParticles
from tests import tree_cases
from tests.test_oracle_reuse import _scenario


@pytest.fixture()
def one_thread(reference):
    prev = reference.max_threads()
    reference.set_threads(1)
    yield
    reference.set_threads(prev)


def test_tidal_and_jerk_match_reference(oracle, reference, one_thread):
    pos, v0, kicks, t3 = _scenario(n=3000, seed=41)
    P = HostParticles(*pos)
    rp = reference.particles(P)
    to = oracle.build_vel(P, v0)
    tr = reference.build_vel(rp, v0)
    try:
        vlast_o = tuple(a.copy() for a in v0)
        vlast_r = tuple(a.copy() for a in v0)
        tk, act, vk = kicks[0]
        oracle.kick_nodes(to, act, vk, pos[3], vlast_o, tk)
        reference.kick_nodes(tr, act, vk, pos[3], vlast_r, tk)
        tg = oracle.tree(to, P).orderbuf
        for theta, aold in ((0.7, None), (0.0, np.full(tg.shape[0], 1e-3))):
            for batch in (0, 8):
                a = oracle.accel_grouped_tj(to, P, tg, theta, 1.7, batch, aold, vk, True)
                b = reference.accel_grouped_tj(tr, rp, tg, theta, 1.7, batch, aold, vk, True)
                for u, w in zip(a, b):
                    assert np.array_equal(u, w)
        # without velocities: the tidal tensor alone, jerk zero
        a = oracle.accel_grouped_tj(to, P, tg, 0.7, 1.0, 0, None, None, True)
        b = reference.accel_grouped_tj(tr, rp, tg, 0.7, 1.0, 0, None, None, True)
        assert np.array_equal(a[3], b[3]) and not a[4].any() and not b[4].any()
    finally:
        oracle.free(to); reference.free(tr); reference.free_particles(rp)


def _direct_accel(x, y, z, m, soft, i, px, py, pz, oracle):
    d = np.stack([x - px, y - py, z - pz], 1)
    r = np.linalg.norm(d, axis=1)
    ok = (r > 0) & (np.arange(x.shape[0]) != i)
    h = np.maximum(np.maximum(soft, soft[i]), 1e-300)
    g = np.array([oracle.spline_force_over_r(rr, hh) for rr, hh in zip(r[ok], h[ok])])
    return (d[ok] * (m[ok] * g)[:, None]).sum(0)


def test_tidal_and_jerk_are_derivatives_of_the_direct_sum(oracle):
    """theta -> 0 opens every node, so the walk is the direct sum and its tidal tensor and jerk
    must be the derivatives of the direct-sum acceleration."""
    x, y, z, m, soft = tree_cases.plummer_small(400, seed=9, eps=0.05)
    rng = np.random.default_rng(2)
    v = tuple(rng.normal(scale=0.3, size=400) for _ in range(3))
    P = HostParticles(x, y, z, m, soft)
    to = oracle.build_vel(P, v)
    try:
        tg = np.arange(0, 400, 37, dtype=np.uint32)
        _, _, _, tid, jk = oracle.accel_grouped_tj(to, P, tg, 1e-9, 1.0, 0, None, v, True)
    finally:
        oracle.free(to)
    comp = ((0, 0), (1, 1), (2, 2), (0, 1), (1, 2), (0, 2))
    for k, i in enumerate(tg):
        p0 = np.array([x[i], y[i], z[i]])
        hstep = 1e-5
        grad = np.zeros((3, 3))
        for l in range(3):
            dp = np.zeros(3)
            dp[l] = hstep
            ap = _direct_accel(x, y, z, m, soft, i, *(p0 + dp), oracle)
            am = _direct_accel(x, y, z, m, soft, i, *(p0 - dp), oracle)
            grad[:, l] = (ap - am) / (2 * hstep)
        scale = np.abs(grad).max()
        for c, (a_, b_) in enumerate(comp):
            assert tid[c, k] == pytest.approx(grad[a_, b_], abs=1e-6 * scale)
        # jerk: every particle moves ballistically; d/dt of the target's acceleration
        dt = 1e-6
        xs = [np.asarray(q) for q in (x, y, z)]
        ap = _direct_accel(*(q + dt * vq for q, vq in zip(xs, v)), m, soft, i,
                           *(p0 + dt * np.array([v[0][i], v[1][i], v[2][i]])), oracle)
        am = _direct_accel(*(q - dt * vq for q, vq in zip(xs, v)), m, soft, i,
                           *(p0 - dt * np.array([v[0][i], v[1][i], v[2][i]])), oracle)
        fd = (ap - am) / (2 * dt)
        np.testing.assert_allclose(jk[:, k], fd, atol=1e-6 * np.abs(fd).max())
