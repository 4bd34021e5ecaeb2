"""The device neighbour search against the oracle (pinned to the reference's ngb_search and to a
brute-force distance test): identical CSR lists, Morton order within each query."""
from __future__ import annotations

import numpy as np
import pytest

from tests import tree_cases
from tests.test_oracle_ngb import _queries
from tests.test_oracle_reuse import _scenario

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def gz():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from gizmo_shared_memory_amd import tree
    return tree


@pytest.mark.parametrize("name", ["plummer20k", "collocated", "tombstones", "n1", "n17"])
def test_ngb_search_matches_oracle(gz, oracle, name):
    from oracle.bindings import HostParticles
    x, y, z, m, soft = tree_cases.CASES[name]()
    print(f"[ngb] case={name}", flush=True)
    hp = HostParticles(x, y, z, m, soft)
    P = gz.Particles(x, y, z, m, soft)
    T = gz.build(P)
    nq = min(300, x.shape[0])
    c, r = _queries(x, y, z, nq, 3, 0.05)
    # Morton-ordered queries, as a hydro step would pass them, and also scattered ones
    for order in (np.argsort(np.asarray(T.download()["rank"])[:nq]), np.arange(nq)):
        cq, rq = c[order], r[order]
        s_d, f_d = gz.ngb_search(T, cq, rq)
        to = oracle.build(hp)
        try:
            s_o, f_o = oracle.ngb_search(to, hp, cq, rq)
        finally:
            oracle.free(to)
        assert np.array_equal(s_d.cpu().numpy(), s_o)
        assert np.array_equal(f_d.cpu().numpy(), f_o)


def test_ngb_search_on_an_aged_tree(gz, oracle):
    from oracle.bindings import HostParticles
    pos, v0, _, _ = _scenario(n=5000, seed=29)
    tau = 0.05
    xt = [pos[d] + v0[d] * tau for d in range(3)]
    P0 = gz.Particles(*pos)
    T = gz.build(P0, vel=v0)
    Pt = gz.Particles(xt[0], xt[1], xt[2], pos[3], pos[4])
    T.update_sources(Pt)
    c, r = _queries(xt[0], xt[1], xt[2], 400, 8, 0.1)
    s_d, f_d = gz.ngb_search(T, c, r, tau)
    hp0 = HostParticles(*pos)
    hpt = HostParticles(xt[0], xt[1], xt[2], pos[3], pos[4])
    to = oracle.build_vel(hp0, v0)
    try:
        s_o, f_o = oracle.ngb_search(to, hpt, c, r, tau)
    finally:
        oracle.free(to)
    assert np.array_equal(s_d.cpu().numpy(), s_o)
    assert np.array_equal(f_d.cpu().numpy(), f_o)


def test_ngb_search_empty_and_large(gz):
    x, y, z, m, soft = tree_cases.plummer_small(20000, seed=2)
    P = gz.Particles(x, y, z, m, soft)
    T = gz.build(P)
    s, f = gz.ngb_search(T, np.zeros((0, 3)), np.zeros(0))
    assert s.shape[0] == 1 and f.shape[0] == 0
    # one query covering everything: every live particle, in Morton order
    s, f = gz.ngb_search(T, np.zeros((1, 3)), np.array([1e6]))
    assert np.array_equal(f.cpu().numpy(), T.download()["orderbuf"])


@pytest.mark.parametrize("name", ["plummer20k", "collocated", "tombstones", "n17"])
def test_ngb_search_reverse_matches_oracle(gz, oracle, name):
    """The reverse search (hydro.cc:110-148): device lists equal to the oracle's, which are the
    reference's and the brute-force h_q <= r < h_j sets."""
    from oracle.bindings import HostParticles
    x, y, z, m, soft = tree_cases.CASES[name]()
    print(f"[ngb-rev] case={name}", flush=True)
    n = x.shape[0]
    rng = np.random.default_rng(11)
    h = 0.04 * rng.uniform(0.3, 2.0, n)
    h[rng.random(n) < 0.05] = 0.0
    n_gas = max(1, int(0.7 * n))
    hp = HostParticles(x, y, z, m, soft)
    P = gz.Particles(x, y, z, m, soft)
    T = gz.build(P)
    T.init_hmax(h, n_gas)
    nq = min(300, n)
    idx = rng.choice(n, nq, replace=False)
    c = np.stack([x[idx], y[idx], z[idx]], 1)
    hq = np.where(h[idx] > 0, h[idx], 0.02)
    to = oracle.build(hp)
    try:
        oracle.init_hmax(to, h, n_gas)
        s_o, f_o = oracle.ngb_search_reverse(to, hp, h, n_gas, c, hq)
    finally:
        oracle.free(to)
    s_d, f_d = gz.ngb_search_reverse(T, h, n_gas, c, hq)
    assert np.array_equal(s_d.cpu().numpy(), s_o)
    assert np.array_equal(f_d.cpu().numpy(), f_o)


def test_ngb_search_reverse_on_an_aged_tree(gz, oracle):
    from oracle.bindings import HostParticles
    pos, v0, _, _ = _scenario(n=5000, seed=31)
    tau = 0.05
    xt = [pos[d] + v0[d] * tau for d in range(3)]
    n = pos[0].shape[0]
    rng = np.random.default_rng(5)
    h = 0.08 * rng.uniform(0.5, 1.5, n)
    n_gas = n // 2
    P0 = gz.Particles(*pos)
    T = gz.build(P0, vel=v0)
    T.init_hmax(h, n_gas)
    Pt = gz.Particles(xt[0], xt[1], xt[2], pos[3], pos[4])
    T.update_sources(Pt)
    idx = rng.choice(n, 400, replace=False)
    c = np.stack([xt[0][idx], xt[1][idx], xt[2][idx]], 1)
    hq = h[idx]
    s_d, f_d = gz.ngb_search_reverse(T, h, n_gas, c, hq, tau)
    hp0 = HostParticles(*pos)
    hpt = HostParticles(xt[0], xt[1], xt[2], pos[3], pos[4])
    to = oracle.build_vel(hp0, v0)
    try:
        oracle.init_hmax(to, h, n_gas)
        s_o, f_o = oracle.ngb_search_reverse(to, hpt, h, n_gas, c, hq, tau)
    finally:
        oracle.free(to)
    assert s_o[-1] > 0
    assert np.array_equal(s_d.cpu().numpy(), s_o)
    assert np.array_equal(f_d.cpu().numpy(), f_o)


def test_ngb_search_reverse_needs_hmax(gz):
    x, y, z, m, soft = tree_cases.CASES["n17"]()
    P = gz.Particles(x, y, z, m, soft)
    T = gz.build(P)
    with pytest.raises(Exception, match="init_hmax"):
        gz.ngb_search_reverse(T, np.full(x.shape[0], 0.1), x.shape[0],
                              np.zeros((2, 3)), np.full(2, 0.05))


def test_ngb_search_reverse_after_raise_hmax(gz, oracle):
    """Kernel radii that grow between rebuilds (mfm.cc:297): raise_hmax on the device and in the
    oracle, then the reverse search finds the new, wider reaches."""
    from oracle.bindings import HostParticles
    x, y, z, m, soft = tree_cases.CASES["plummer20k"]()
    n = x.shape[0]
    rng = np.random.default_rng(23)
    h = 0.03 * rng.uniform(0.5, 1.5, n)
    n_gas = int(0.8 * n)
    grow = rng.choice(n_gas, 800, replace=False).astype(np.uint32)
    h2 = h.copy()
    h2[grow] *= 2.5
    P = gz.Particles(x, y, z, m, soft)
    T = gz.build(P)
    T.init_hmax(h, n_gas)
    T.raise_hmax(grow, h2)
    idx = rng.choice(n, 300, replace=False)
    c = np.stack([x[idx], y[idx], z[idx]], 1)
    hq = h2[idx]
    s_d, f_d = gz.ngb_search_reverse(T, h2, n_gas, c, hq)
    hp = HostParticles(x, y, z, m, soft)
    to = oracle.build(hp)
    try:
        oracle.init_hmax(to, h, n_gas)
        oracle.raise_hmax(to, grow, h2)
        s_o, f_o = oracle.ngb_search_reverse(to, hp, h2, n_gas, c, hq)
    finally:
        oracle.free(to)
    assert s_o[-1] > 0
    assert np.array_equal(s_d.cpu().numpy(), s_o)
    assert np.array_equal(f_d.cpu().numpy(), f_o)


def test_searches_on_a_forest_match_brute_force(gz):
    """Both searches over a forest (one tree per cluster, spheres close enough that kernels of
    one reach into the next) give exactly the brute-force sets."""
    from gizmo_shared_memory_amd import ics
    (x, y, z, m, soft), starts = ics.spheres(n_spheres=4, lo_log2=9.0, hi_log2=11.0, eps=0.01,
                                             spacing=1.5, seed=6)
    print("[ngb-forest] 4 clusters", flush=True)
    n = x.shape[0]
    P = gz.Particles(x, y, z, m, soft)
    F = gz.build(P, clusters=starts)
    rng = np.random.default_rng(8)
    h = rng.uniform(0.05, 0.4, n)
    n_gas = int(0.75 * n)
    F.init_hmax(h, n_gas)
    idx = rng.choice(n, 256, replace=False)
    c = np.stack([x[idx], y[idx], z[idx]], 1) + rng.normal(scale=0.05, size=(256, 3))
    r = rng.uniform(0.05, 0.3, 256)
    s_f, f_f = gz.ngb_search(F, c, r)
    s_r, f_r = gz.ngb_search_reverse(F, h, n_gas, c, r)
    s_f, f_f, s_r, f_r = (t.cpu().numpy() for t in (s_f, f_f, s_r, f_r))
    live = m > 0
    gas = np.arange(n) < n_gas
    for q in range(c.shape[0]):
        d2 = (x - c[q, 0]) ** 2 + (y - c[q, 1]) ** 2 + (z - c[q, 2]) ** 2
        want_f = np.nonzero(live & (d2 < r[q] ** 2))[0]
        want_r = np.nonzero(live & gas & (d2 >= r[q] ** 2) & (d2 < h ** 2))[0]
        assert np.array_equal(np.sort(f_f[s_f[q]:s_f[q + 1]]), want_f), q
        assert np.array_equal(np.sort(f_r[s_r[q]:s_r[q + 1]]), want_r), q
    assert s_f[-1] > 0 and s_r[-1] > 0
