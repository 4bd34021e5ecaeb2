"""Events on a reused tree that are not a drift or a kick — Tree::mass_event, shift_node,
raise_pad and raise_vmax (tree.cc:239-258, 276-285, 323-330) — in the oracle, pinned:

* bit for bit against the reference's own methods called in the same order (one thread), through
  the incremental moments, the pads, the speed bounds and the refreshed walk nodes, in an open
  and in a periodic box;
* the refreshed mass and centre of mass against those of the particle set with the event
  applied (a particle removed and put back with another mass, position and velocity — the
  accretion merge of mfm.cc:1676-1687 — or displaced without a velocity change, the Hermite
  correction of mfm.cc:2350-2356), computed directly from the particles;
* an event followed by its negation cancelling;
* a leaf's pad against the float-rounded-up running sum of its displacements, every node's pad
  against the maximum over the leaves below it, and every speed bound against the largest speed
  reported below the node.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle.bindings import HostParticles
from tests import tree_cases


@pytest.fixture()
def one_thread(reference):
    prev = reference.max_threads()
    reference.set_threads(1)
    yield
    reference.set_threads(prev)


def _setup(n=3000, seed=12):
    x, y, z, m, soft = tree_cases.plummer_small(n, seed=seed, eps=0.01)
    rng = np.random.default_rng(seed + 1)
    v = tuple(rng.normal(scale=0.3, size=n) for _ in range(3))
    return (x, y, z, m, soft), v, rng


def _ancestors(ta, leaf):
    out = []
    no = int(leaf)
    while no >= 0:
        out.append(no)
        no = int(ta.parent[no])
    return out


def _events(pos, v, rng, tau, n_ev=40, box=0.0):
    """Accretion-like pairs (remove particle j at its position at tau, put back a heavier one
    nearby with a new velocity), then Hermite-like displacements of other particles."""
    n = pos[0].shape[0]
    js = rng.choice(n, n_ev, replace=False).astype(np.uint32)
    xj = [pos[d][js] + v[d][js] * tau for d in range(3)]
    vj = [v[d][js] for d in range(3)]
    m0 = pos[3][js]
    m1 = m0 * (1.0 + rng.uniform(0.1, 2.0, n_ev))
    x1 = [a + rng.normal(scale=1e-3, size=n_ev) for a in xj]
    if box > 0:  # some events reported at a periodic image of the position
        flip = rng.random(n_ev) < 0.5
        x1[0] = np.where(flip, x1[0] + box, x1[0])
    v1 = [a + rng.normal(scale=0.05, size=n_ev) for a in vj]
    idx = np.repeat(js, 2)
    dm = np.empty(2 * n_ev)
    dm[0::2], dm[1::2] = -m0, m1
    ex = [np.empty(2 * n_ev) for _ in range(3)]
    ev = [np.empty(2 * n_ev) for _ in range(3)]
    for d in range(3):
        ex[d][0::2], ex[d][1::2] = xj[d], x1[d]
        ev[d][0::2], ev[d][1::2] = vj[d], v1[d]
    ks = np.setdiff1d(rng.choice(n, n_ev, replace=False), js).astype(np.uint32)
    dx = [rng.normal(scale=2e-3, size=ks.shape[0]) for _ in range(3)]
    return dict(idx=idx, dm=dm, x=ex, v=ev, js=js, m1=m1, x1=x1, v1=v1, ks=ks, dx=dx)


def _apply(lib, t, e, m, tau, box):
    lib.mass_events(t, e["idx"], e["dm"], e["x"], e["v"], tau, box)
    lib.shift_nodes(t, e["ks"], m[e["ks"]], e["dx"])
    d = np.sqrt(sum(a * a for a in e["dx"]))
    lib.raise_pad(t, e["ks"], d)
    js = e["js"]
    sp = (np.sqrt(sum(a * a for a in e["v1"])) * (1.0 + 1e-6)).astype(np.float32)
    lib.raise_vmax(t, js, sp)


@pytest.mark.parametrize("box", [0.0, 40.0])
def test_events_match_reference_bitwise(oracle, reference, one_thread, box):
    pos, v, rng = _setup()
    P = HostParticles(*pos)
    rp = reference.particles(P)
    tau, t3 = 0.07, 0.11
    e = _events(pos, v, rng, tau, box=box)
    to = oracle.build_vel(P, v)
    tr = reference.build_vel(rp, v)
    try:
        for lib, t in ((oracle, to), (reference, tr)):
            _apply(lib, t, e, pos[3], tau, box)
            lib.refresh_walk_nodes(t, t3)
        a, b = oracle.reuse(to), reference.reuse(tr)
        for f in ("vmax", "dp", "dq", "wcom", "wmass", "wn_s", "wn_len"):
            assert np.array_equal(getattr(a, f), getattr(b, f)), f
        pa, pb = oracle.pads(to), reference.pads(tr)
        assert np.array_equal(pa[0], pb[0]) and np.array_equal(pa[1], pb[1])
        assert np.any(pa[0] > 0) and np.any(pa[1] != 0)
    finally:
        oracle.free(to)
        reference.free(tr)
        reference.free_particles(rp)


def test_refreshed_moments_are_those_of_the_changed_set(oracle):
    pos, v, rng = _setup(n=2500, seed=33)
    P = HostParticles(*pos)
    tau, t3 = 0.05, 0.13
    e = _events(pos, v, rng, tau, n_ev=25)
    to = oracle.build_vel(P, v)
    try:
        _apply(oracle, to, e, pos[3], tau, 0.0)
        oracle.refresh_walk_nodes(to, t3)
        r = oracle.reuse(to)
        ta = oracle.tree(to, P)
        # the particle set at t3: ballistic, except the merged particles (new mass, position at
        # tau and velocity) and the displaced ones (shifted by dx)
        m = pos[3].copy()
        xt = [pos[d] + v[d] * t3 for d in range(3)]
        js, ks = e["js"], e["ks"]
        m[js] = e["m1"]
        for d in range(3):
            xt[d][js] = e["x1"][d] + e["v1"][d] * (t3 - tau)
            xt[d][ks] += e["dx"][d]
        for i in range(ta.nnodes):
            idx = ta.orderbuf[ta.plo[i]:ta.phi[i]]
            M = m[idx].sum()
            com = np.array([np.dot(m[idx], xt[d][idx]) / M for d in range(3)])
            assert r.wmass[i] == pytest.approx(M, rel=1e-13), i
            assert np.linalg.norm(r.wcom[i] - com) <= 1e-12 * (1.0 + np.linalg.norm(com)), i
    finally:
        oracle.free(to)


def test_event_and_its_negation_cancel(oracle):
    pos, v, rng = _setup(n=1500, seed=4)
    P = HostParticles(*pos)
    to = oracle.build_vel(P, v)
    try:
        n_ev = 30
        j = rng.choice(P.n, n_ev).astype(np.uint32)
        dm = rng.uniform(0.1, 1.0, n_ev) * pos[3][j]
        x = [rng.normal(size=n_ev) for _ in range(3)]
        vv = [rng.normal(size=n_ev) for _ in range(3)]
        idx = np.concatenate([j, j])
        oracle.mass_events(to, idx, np.concatenate([dm, -dm]), [np.tile(a, 2) for a in x],
                           [np.tile(a, 2) for a in vv], 0.04, 0.0)
        oracle.shift_nodes(to, j, dm, x)
        oracle.shift_nodes(to, j, dm, [-a for a in x])
        r = oracle.reuse(to)
        _, dmass = oracle.pads(to)
        scale = float(np.abs(dm).max() * 10.0)
        assert np.all(np.abs(dmass) <= 1e-15 * scale)
        assert np.all(np.abs(r.dp) <= 1e-15 * scale)
        assert np.all(np.abs(r.dq) <= 1e-14 * scale)
    finally:
        oracle.free(to)


def _float_up(d):
    f = np.float32(d)
    return f if float(f) >= d else np.nextafter(f, np.float32(np.inf))


def test_pads_and_speed_bounds(oracle):
    pos, v, rng = _setup(n=2000, seed=9)
    P = HostParticles(*pos)
    to = oracle.build_vel(P, v)
    try:
        ta = oracle.tree(to, P)
        vmax0 = oracle.reuse(to).vmax.copy()
        # repeated displacements of a few particles, some sharing a leaf
        ks = rng.choice(P.n, 12, replace=False).astype(np.uint32)
        seq = np.concatenate([ks, ks[:5], ks[:2]])
        d = rng.uniform(1e-4, 1e-2, seq.shape[0])
        d[3] = 0.0  # a zero displacement changes nothing
        oracle.raise_pad(to, seq, d)
        speeds = rng.uniform(0.0, 3.0, seq.shape[0]).astype(np.float32)
        oracle.raise_vmax(to, seq, speeds)
        xpad, _ = oracle.pads(to)
        vmax = oracle.reuse(to).vmax
        leaf_pad = {}
        for k, dk in zip(seq, d):
            leaf = int(ta.leaf_of[k])
            if dk > 0:
                leaf_pad[leaf] = _float_up(float(leaf_pad.get(leaf, np.float32(0))) + dk)
        want = np.zeros(ta.nnodes, np.float32)
        want_v = vmax0.copy()
        for leaf, p in leaf_pad.items():
            for no in _ancestors(ta, leaf):
                want[no] = max(want[no], p)
        for k, s in zip(seq, speeds):
            for no in _ancestors(ta, ta.leaf_of[k]):
                want_v[no] = max(want_v[no], s)
        assert np.array_equal(xpad, want)
        assert np.array_equal(vmax, want_v)
        # the refresh widens the opening radius and side by 2 (vmax tau + xpad)
        tau = 0.02
        oracle.refresh_walk_nodes(to, tau)
        r = oracle.reuse(to)
        grow = 2.0 * (vmax.astype(np.float64) * tau + xpad.astype(np.float64))
        np.testing.assert_array_equal(r.wn_s, ta.size + ta.delta + grow)
        np.testing.assert_array_equal(r.wn_len, (ta.size + grow).astype(np.float32))
    finally:
        oracle.free(to)
