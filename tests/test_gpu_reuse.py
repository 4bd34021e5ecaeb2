"""Device tree reuse against the oracle (whose reuse functions are pinned to the reference in
tests/test_oracle_reuse.py): build with velocities, two kick feeds, a refresh, a walk over the
refreshed nodes on drifted particles, and raised softenings.

Tolerances: maxima (vmax, softening bounds) and everything computed from the build alone are bit
for bit; the momentum feed is a sum of FP64 atomic additions whose order is not fixed, so dp, dq
and the refreshed centres of mass agree to 1e-12 relative, and the walk over them to 1e-10.
"""
from __future__ import annotations

import types

import numpy as np
import pytest

from tests import tree_cases
from tests.test_oracle_reuse import _drifted, _scenario
from tests.treecmp import canonical

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def gz():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from gizmo_shared_memory_amd import tree
    return tree


def _dev(a):
    return torch.as_tensor(np.ascontiguousarray(a), device="cuda:0")


def _perm_device(d):
    return canonical(types.SimpleNamespace(**d))[0]


def _perm_oracle(oracle, to, hp):
    return canonical(oracle.tree(to, hp))[0]


def _close(a, b, rtol, scale=None):
    scale = np.maximum(np.abs(b), 1e-300) if scale is None else scale
    assert np.all(np.abs(a - b) <= rtol * scale), float(np.max(np.abs(a - b) / scale))


def test_reuse_matches_oracle(gz, oracle):
    from oracle.bindings import HostParticles
    pos, v0, kicks, t3 = _scenario(n=6000, seed=31)
    hp = HostParticles(*pos)
    P = gz.Particles(*pos)
    T = gz.build(P, vel=v0)
    to = oracle.build_vel(hp, v0)
    try:
        d = T.download()
        pd, po = _perm_device(d), _perm_oracle(oracle, to, hp)
        # build: vmax and vcom bit for bit
        r = T.download_reuse()
        o = oracle.reuse(to)
        assert np.array_equal(r["vmax"][pd], o.vmax[po])
        assert np.array_equal(r["vcom"][pd], o.vcom[po])
        # two kick feeds
        vlast_d = tuple(_dev(a.copy()) for a in v0)
        vlast_o = tuple(a.copy() for a in v0)
        for tk, act, vk in kicks:
            T.kick_nodes(act, vk, None, vlast_d, tk)
            oracle.kick_nodes(to, act, vk, pos[3], vlast_o, tk)
        for u, w in zip(vlast_d, vlast_o):
            assert np.array_equal(u.cpu().numpy(), w)
        r = T.download_reuse()
        o = oracle.reuse(to)
        assert np.array_equal(r["vmax"][pd], o.vmax[po])
        mom = np.abs(o.dp[po]).max() + 1e-300
        _close(r["dp"][pd], o.dp[po], 1e-12, mom)
        _close(r["dq"][pd], o.dq[po], 1e-12, np.abs(o.dq[po]).max() + 1e-300)
        # refresh
        T.refresh_walk_nodes(t3)
        oracle.refresh_walk_nodes(to, t3)
        d = T.download()
        o = oracle.reuse(to)
        assert np.array_equal(d["mass"][pd], o.wmass[po])
        com_d = np.stack([d["cx"], d["cy"], d["cz"]], 1)[pd]
        _close(com_d, o.wcom[po], 1e-12, 1.0 + np.abs(o.wcom[po]))
        assert np.array_equal(d["wn_s"][pd], o.wn_s[po])
        assert np.array_equal(d["wn_len"][pd], o.wn_len[po])
        # walk over the refreshed tree, particles drifted to t3
        xt = _drifted(pos, v0, kicks, t3)
        Pt = gz.Particles(xt[0], xt[1], xt[2], pos[3], pos[4])
        T.update_sources(Pt)
        hpt = HostParticles(xt[0], xt[1], xt[2], pos[3], pos[4])
        tg = d["orderbuf"]
        gx = gz.accel_grouped(T, Pt, tg, 0.6, 1.0, 0, counts=True)
        ox, oy, oz, oc = oracle.accel_grouped(to, hpt, tg, 0.6, 1.0, 0, None, want_counts=True)
        assert [gx[3]["nodes_visited"], gx[3]["multipole"], gx[3]["pairs"]] == \
            [int(v) for v in oc]
        g = np.stack([t.cpu().numpy() for t in gx[:3]], 1)
        ob = np.stack([ox, oy, oz], 1)
        err = np.linalg.norm(g - ob, axis=1) / np.linalg.norm(ob, axis=1)
        assert err.max() < 1e-10
        # raise_soft
        soft2 = pos[4].copy()
        grow = np.arange(0, hp.n, 53, dtype=np.uint32)
        soft2[grow] *= 2.5
        T.raise_soft(grow, soft2)
        oracle.raise_soft(to, grow, soft2)
        d = T.download()
        ta = oracle.tree(to, hp)
        o = oracle.reuse(to)
        assert np.array_equal(d["wn_soft"][pd], o.wn_soft[po])
        assert np.array_equal(d["soft"][pd], ta.soft[po])
    finally:
        oracle.free(to)


def test_refresh_without_velocities_is_identity(gz):
    case = tree_cases.plummer_small(5000, seed=3, eps=0.01)
    P = gz.Particles(*case)
    T = gz.build(P)
    before = T.download()
    T.refresh_walk_nodes(0.8)
    after = T.download()
    for k in ("mass", "cx", "cy", "cz", "wn_s", "wn_len", "wn_soft"):
        assert np.array_equal(before[k], after[k]), k
    r = T.download_reuse()
    assert not r["vmax"].any() and not r["vcom"].any()


def test_kick_nodes_rejects_bad_index(gz):
    case = tree_cases.plummer_small(1000, seed=4)
    P = gz.Particles(*case)
    T = gz.build(P)
    v = tuple(np.zeros(1000) for _ in range(3))
    with pytest.raises(Exception, match="outside"):
        T.kick_nodes(np.array([5, 1000], dtype=np.uint32), v)


@pytest.mark.parametrize("batch", [0, 8])
def test_tidal_and_jerk_match_oracle(gz, oracle, batch):
    """The tidal tensor and jerk of the FP64 grouped walk (tree.cc:811-871) against the oracle
    (pinned to the reference and to finite differences of the direct sum), on a tree built with
    velocities and fed one kick."""
    from oracle.bindings import HostParticles
    pos, v0, kicks, _ = _scenario(n=5000, seed=43)
    hp = HostParticles(*pos)
    P = gz.Particles(*pos)
    T = gz.build(P, vel=v0)
    to = oracle.build_vel(hp, v0)
    try:
        tk, act, vk = kicks[0]
        T.kick_nodes(act, vk, None, tuple(_dev(a.copy()) for a in v0), tk)
        oracle.kick_nodes(to, act, vk, pos[3], tuple(a.copy() for a in v0), tk)
        tg = oracle.tree(to, hp).orderbuf
        for theta, aold in ((0.7, None), (0.0, np.full(tg.shape[0], 1e-3))):
            g = gz.accel_grouped_tidal(T, P, tg, theta, 1.3, batch, aold, vel=vk, jerk=True)
            o = oracle.accel_grouped_tj(to, hp, tg, theta, 1.3, batch, aold, vk, True)
            acc_g = np.stack([_np(v) for v in g[:3]], 1)
            acc_o = np.stack(o[:3], 1)
            e = np.linalg.norm(acc_g - acc_o, axis=1) / np.linalg.norm(acc_o, axis=1)
            assert e.max() < 1e-12
            tid_g, tid_o = _np(g[3]), o[3]
            scale = np.abs(tid_o).max(0)
            assert np.all(np.abs(tid_g - tid_o) <= 1e-11 * scale), \
                np.max(np.abs(tid_g - tid_o) / scale)
            jk_g, jk_o = _np(g[4]), o[4]
            jscale = np.linalg.norm(jk_o, axis=0)
            assert np.all(np.linalg.norm(jk_g - jk_o, axis=0) <= 1e-10 * jscale)
    finally:
        oracle.free(to)


def _np(t):
    return t.detach().cpu().numpy()
