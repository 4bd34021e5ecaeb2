"""Neighbour search (hydro.cc:37-108) in the oracle: the same lists, in the same order, as the
reference's ngb_search, and as sets the brute-force distance test."""
from __future__ import annotations

import numpy as np

from oracle.bindings import HostParticles
from tests import tree_cases
from tests.test_oracle_reuse import _scenario


def _queries(x, y, z, n_q, seed, scale):
    rng = np.random.default_rng(seed)
    idx = rng.choice(x.shape[0], n_q, replace=False)
    c = np.stack([x[idx], y[idx], z[idx]], 1) + rng.normal(scale=scale * 0.1, size=(n_q, 3))
    r = scale * rng.uniform(0.3, 1.5, n_q)
    return c, r


def test_ngb_search_matches_reference_and_brute_force(oracle, reference):
    for name in ("plummer20k", "collocated", "tombstones"):
        x, y, z, m, soft = tree_cases.CASES[name]()
        P = HostParticles(x, y, z, m, soft)
        c, r = _queries(x, y, z, 200, 3, 0.05)
        to = oracle.build(P)
        rp = reference.particles(P)
        tr = reference.build(rp)
        try:
            so, fo = oracle.ngb_search(to, P, c, r)
            sr, fr = reference.ngb_search(tr, rp, c, r)
        finally:
            oracle.free(to); reference.free(tr); reference.free_particles(rp)
        assert np.array_equal(so, sr), name
        keys, _ = oracle.keys(P)
        ties = np.unique(keys[m > 0]).shape[0] < int((m > 0).sum())
        if not ties:
            assert np.array_equal(fo, fr), name
        for q in range(c.shape[0]):
            # traversal order is Morton order (ties inside a run of equal keys are unspecified
            # in the reference's sort)
            a, b = fo[so[q]:so[q + 1]], fr[sr[q]:sr[q + 1]]
            assert np.all(np.diff(keys[a].astype(np.int64)) >= 0) and \
                np.all(np.diff(keys[b].astype(np.int64)) >= 0)
            assert np.array_equal(np.sort(a), np.sort(b))
        live = m > 0
        for q in range(c.shape[0]):
            d2 = (x - c[q, 0]) ** 2 + (y - c[q, 1]) ** 2 + (z - c[q, 2]) ** 2
            want = np.nonzero((d2 < r[q] ** 2) & live)[0]
            assert np.array_equal(np.sort(fo[so[q]:so[q + 1]]), want), (name, q)


def test_ngb_search_on_an_aged_tree(oracle, reference):
    """After a drift the search still finds every neighbour at the current positions: the node
    boxes are widened by vmax * tau (the tree built with velocities, the reference likewise)."""
    pos, v0, _, _ = _scenario(n=4000, seed=17)
    P0 = HostParticles(*pos)
    tau = 0.05
    xt = [pos[d] + v0[d] * tau for d in range(3)]
    Pt = HostParticles(xt[0], xt[1], xt[2], pos[3], pos[4])
    c, r = _queries(xt[0], xt[1], xt[2], 150, 5, 0.1)
    to = oracle.build_vel(P0, v0)
    rp0 = reference.particles(P0)
    tr = reference.build_vel(rp0, v0)
    rpt = reference.particles(Pt)
    try:
        so, fo = oracle.ngb_search(to, Pt, c, r, tau)
        sr, fr = reference.ngb_search(tr, rpt, c, r, tau)
    finally:
        oracle.free(to); reference.free(tr)
        reference.free_particles(rp0); reference.free_particles(rpt)
    assert np.array_equal(so, sr) and np.array_equal(fo, fr)
    for q in range(c.shape[0]):
        d2 = (xt[0] - c[q, 0]) ** 2 + (xt[1] - c[q, 1]) ** 2 + (xt[2] - c[q, 2]) ** 2
        assert np.array_equal(np.sort(fo[so[q]:so[q + 1]]), np.nonzero(d2 < r[q] ** 2)[0])


def _h_as_bound(h):
    """h rounded up to float32 (tree.cc:286-290)."""
    f = h.astype(np.float32)
    up = f.astype(np.float64) < h
    f[up] = np.nextafter(f[up], np.float32(1e30))
    return f


def _hydro_case(name, seed=9):
    x, y, z, m, soft = tree_cases.CASES[name]()
    rng = np.random.default_rng(seed)
    n = x.shape[0]
    h = 0.04 * rng.uniform(0.3, 2.0, n)
    h[rng.random(n) < 0.05] = 0.0  # gas with no kernel yet never raises hmax
    n_gas = int(0.7 * n)
    return HostParticles(x, y, z, m, soft), h, n_gas


def test_hmax_matches_reference_and_subtree_maximum(oracle, reference):
    for name in ("plummer20k", "collocated", "tombstones"):
        P, h, n_gas = _hydro_case(name)
        to = oracle.build(P)
        rp = reference.particles(P)
        tr = reference.build(rp)
        try:
            ho = oracle.init_hmax(to, h, n_gas)
            hr = reference.init_hmax(tr, h, n_gas)
            T = oracle.tree(to, P)
        finally:
            oracle.free(to); reference.free(tr); reference.free_particles(rp)
        assert np.array_equal(ho.view(np.uint32), hr.view(np.uint32)), name
        # closed form: the largest rounded-up h of the gas particles below each node
        hb = _h_as_bound(h)
        gas = (np.arange(P.n) < n_gas) & (h > 0)
        for i in range(T.nnodes):
            js = T.orderbuf[T.plo[i]:T.phi[i]]
            js = js[gas[js]]
            want = hb[js].max() if js.size else np.float32(0)
            assert ho[i] == want, (name, i)


def test_raise_hmax_matches_reference(oracle, reference):
    """raise_hmax (tree.cc:293-301) after init_hmax: bit for bit the reference's, and the subtree
    maximum of the raised radii (h only grows here, so the raises reach what init would)."""
    P, h, n_gas = _hydro_case("plummer20k")
    rng = np.random.default_rng(21)
    grow = rng.choice(n_gas, 500, replace=False).astype(np.uint32)
    h2 = h.copy()
    h2[grow] = h2[grow] * 1.7 + 0.01
    to = oracle.build(P)
    rp = reference.particles(P)
    tr = reference.build(rp)
    to2 = oracle.build(P)
    try:
        oracle.init_hmax(to, h, n_gas)
        reference.init_hmax(tr, h, n_gas)
        oracle.raise_hmax(to, grow, h2)
        hr = reference.raise_hmax(tr, grow, h2)
        ho = oracle.hmax(to)
        fresh = oracle.init_hmax(to2, h2, n_gas)
    finally:
        oracle.free(to); oracle.free(to2); reference.free(tr); reference.free_particles(rp)
    assert np.array_equal(ho.view(np.uint32), hr.view(np.uint32))
    assert np.array_equal(ho.view(np.uint32), fresh.view(np.uint32))


def test_ngb_search_reverse_matches_reference_and_brute_force(oracle, reference):
    for name in ("plummer20k", "collocated", "tombstones"):
        P, h, n_gas = _hydro_case(name)
        x, y, z, m = P.x, P.y, P.z, P.m
        rng = np.random.default_rng(4)
        idx = rng.choice(P.n, 200, replace=False)
        c = np.stack([x[idx], y[idx], z[idx]], 1)
        hq = np.where(h[idx] > 0, h[idx], 0.02)
        to = oracle.build(P)
        rp = reference.particles(P)
        tr = reference.build(rp)
        try:
            oracle.init_hmax(to, h, n_gas)
            reference.init_hmax(tr, h, n_gas)
            so, fo = oracle.ngb_search_reverse(to, P, h, n_gas, c, hq)
            sr, fr = reference.ngb_search_reverse(tr, rp, h, n_gas, c, hq)
        finally:
            oracle.free(to); reference.free(tr); reference.free_particles(rp)
        assert np.array_equal(so, sr) and np.array_equal(fo, fr), name
        if name == "plummer20k":  # collocated particles sit inside every kernel: no pairs
            assert so[-1] > 0, name
        live = m > 0
        gas = np.arange(P.n) < n_gas
        for q in range(c.shape[0]):
            d2 = (x - c[q, 0]) ** 2 + (y - c[q, 1]) ** 2 + (z - c[q, 2]) ** 2
            want = np.nonzero(live & gas & (d2 >= hq[q] ** 2) & (d2 < h ** 2))[0]
            assert np.array_equal(np.sort(fo[so[q]:so[q + 1]]), want), (name, q)
