"""Tree reuse between rebuilds in the oracle, pinned three ways:

* bit for bit against the reference (tree.cc:211-319 with the kick feed of mfm.cc:3798-3811),
  both run on one thread so the node numbering and the order of every sum agree;
* the refreshed centre of mass against the centre of mass of the drifted particles: particles
  move ballistically between kicks, which makes the incremental moments exact;
* vmax against max |v| over each node's particles, over every velocity the node has seen.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle.bindings import HostParticles
from tests import tree_cases


def _velocities(n, seed, scale=1.0):
    rng = np.random.default_rng(seed)
    return tuple(rng.normal(scale=scale, size=n) for _ in range(3))


def _scenario(n=4000, seed=8):
    """Build at t = 0, kick a third of the particles at t1 and another third at t2, refresh at
    t3. Returns the particles, velocities and the kick schedule."""
    x, y, z, m, soft = tree_cases.plummer_small(n, seed=seed, eps=0.01)
    v0 = _velocities(n, seed + 1, 0.3)
    rng = np.random.default_rng(seed + 2)
    k1 = np.sort(rng.choice(n, n // 3, replace=False)).astype(np.uint32)
    k2 = np.sort(rng.choice(n, n // 3, replace=False)).astype(np.uint32)
    v1 = tuple(a.copy() for a in v0)
    for a in v1:
        a[k1] += rng.normal(scale=0.2, size=k1.shape[0])
    v2 = tuple(a.copy() for a in v1)
    for a in v2:
        a[k2] += rng.normal(scale=0.2, size=k2.shape[0])
    return (x, y, z, m, soft), v0, [(0.05, k1, v1), (0.12, k2, v2)], 0.2


def _drifted(pos, v0, kicks, t):
    """Positions at t: ballistic, with the velocity changes of the kick schedule."""
    x = [p.copy() for p in pos[:3]]
    t_prev, v = 0.0, v0
    for tk, _, vk in kicks:
        for d in range(3):
            x[d] += v[d] * (tk - t_prev)
        t_prev, v = tk, vk
    for d in range(3):
        x[d] += v[d] * (t - t_prev)
    return x


def _run(lib, handle_of, P, v0, kicks, t3, m):
    t = handle_of(v0)
    vlast = tuple(a.copy() for a in v0)
    for tk, act, vk in kicks:
        lib.kick_nodes(t, act, vk, m, vlast, tk)
    lib.refresh_walk_nodes(t, t3)
    return t, lib.reuse(t)


@pytest.fixture()
def one_thread(reference):
    prev = reference.max_threads()
    reference.set_threads(1)
    yield
    reference.set_threads(prev)


def test_reuse_matches_reference_bitwise(oracle, reference, one_thread):
    pos, v0, kicks, t3 = _scenario()
    P = HostParticles(*pos)
    rp = reference.particles(P)
    to, a = _run(oracle, lambda v: oracle.build_vel(P, v), P, v0, kicks, t3, P.m)
    tr, b = _run(reference, lambda v: reference.build_vel(rp, v), P, v0, kicks, t3, P.m)
    try:
        for f in ("vmax", "vcom", "dp", "dq", "wcom", "wmass", "wn_s", "wn_soft", "wn_len"):
            assert np.array_equal(getattr(a, f), getattr(b, f)), f
        # the walk over the refreshed nodes, on the drifted particles
        xt = _drifted(pos, v0, kicks, t3)
        Pt = HostParticles(xt[0], xt[1], xt[2], pos[3], pos[4])
        rpt = reference.particles(Pt)
        tg = oracle.tree(to, P).orderbuf
        ga = oracle.accel_grouped(to, Pt, tg, 0.6, 1.0, 0)
        gb = reference.accel_grouped(tr, rpt, tg, 0.6, 1.0, 0)
        reference.free_particles(rpt)
        for u, w in zip(ga, gb):
            assert np.array_equal(u, w)
        # raise_soft: a few particles' softening grows
        soft2 = pos[4].copy()
        grow = np.arange(0, P.n, 97, dtype=np.uint32)
        soft2[grow] *= 3.0
        oracle.raise_soft(to, grow, soft2)
        reference.raise_soft(tr, grow, soft2)
        a2, b2 = oracle.reuse(to), reference.reuse(tr)
        assert np.array_equal(a2.wn_soft, b2.wn_soft)
        ta, tb = oracle.tree(to, P), reference.tree(tr, P.n)
        assert np.array_equal(ta.soft, tb.soft)
    finally:
        oracle.free(to)
        reference.free(tr)
        reference.free_particles(rp)


def test_refreshed_com_is_the_drifted_com(oracle):
    pos, v0, kicks, t3 = _scenario(n=3000, seed=21)
    P = HostParticles(*pos)
    to, r = _run(oracle, lambda v: oracle.build_vel(P, v), P, v0, kicks, t3, P.m)
    try:
        ta = oracle.tree(to, P)
        xt = _drifted(pos, v0, kicks, t3)
        m = pos[3]
        for i in range(ta.nnodes):
            idx = ta.orderbuf[ta.plo[i]:ta.phi[i]]
            M = m[idx].sum()
            com = np.array([np.dot(m[idx], xt[d][idx]) / M for d in range(3)])
            assert r.wmass[i] == pytest.approx(M, rel=1e-13)
            assert np.linalg.norm(r.wcom[i] - com) <= 1e-12 * (1.0 + np.linalg.norm(com))
        # vmax: the largest speed each node's particles have had since the build
        speeds = [np.sqrt(v[0] ** 2 + v[1] ** 2 + v[2] ** 2) for v in
                  [v0] + [k[2] for k in kicks]]
        seen = speeds[0].copy()
        for (_, act, _), sp in zip(kicks, speeds[1:]):
            seen[act] = np.maximum(seen[act], sp[act])
        for i in range(ta.nnodes):
            idx = ta.orderbuf[ta.plo[i]:ta.phi[i]]
            assert r.vmax[i] == np.float32(seen[idx].max()), i
        # the opening radius and side grow by 2 vmax tau
        grow = 2.0 * r.vmax.astype(np.float64) * t3
        np.testing.assert_array_equal(r.wn_s, ta.size + ta.delta + grow)
        np.testing.assert_array_equal(r.wn_len, (ta.size + grow).astype(np.float32))
    finally:
        oracle.free(to)


def test_refresh_without_velocities_is_identity(oracle):
    P = HostParticles(*tree_cases.plummer_small(2000, seed=5))
    to = oracle.build(P)
    try:
        before = oracle.reuse(to)
        oracle.refresh_walk_nodes(to, 0.7)
        after = oracle.reuse(to)
        for f in ("wcom", "wmass", "wn_s", "wn_len", "wn_soft"):
            assert np.array_equal(getattr(before, f), getattr(after, f)), f
    finally:
        oracle.free(to)
