"""Device mass events, shifts, pads and speed bounds (Tree::mass_event, shift_node, raise_pad,
raise_vmax; tree.cc:239-258, 276-285, 323-330) against the oracle, whose versions are pinned to
the reference's in tests/test_oracle_events.py.

The device applies the events in the caller's order with the oracle's operations, so on a tree
whose incremental moments hold no atomic sums (no kick feed yet) every moment, pad and speed bound
— and the walk nodes refreshed from them — agree bit for bit. After a kick feed (FP64 atomics)
the moments agree to 1e-12 relative. A pad widens the neighbour search's node test exactly as the
oracle's does, so a particle displaced out of its leaf's box is still found.
"""
from __future__ import annotations

import types

import numpy as np
import pytest

from tests.test_oracle_events import _events, _setup
from tests.treecmp import canonical

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def gz():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from gizmo_shared_memory_amd import tree
    return tree


def _perms(T, oracle, to, hp):
    pd = canonical(types.SimpleNamespace(**T.download()))[0]
    po = canonical(oracle.tree(to, hp))[0]
    return pd, po


def _apply_device(T, e, m, tau, box):
    T.mass_events(e["idx"], e["dm"], e["x"], e["v"], tau, box)
    T.shift_nodes(e["ks"], m[e["ks"]], e["dx"])
    T.raise_pad(e["ks"], np.sqrt(sum(a * a for a in e["dx"])))
    sp = (np.sqrt(sum(a * a for a in e["v1"])) * (1.0 + 1e-6)).astype(np.float32)
    T.raise_vmax(e["js"], sp)


def _apply_oracle(oracle, to, e, m, tau, box):
    oracle.mass_events(to, e["idx"], e["dm"], e["x"], e["v"], tau, box)
    oracle.shift_nodes(to, e["ks"], m[e["ks"]], e["dx"])
    oracle.raise_pad(to, e["ks"], np.sqrt(sum(a * a for a in e["dx"])))
    sp = (np.sqrt(sum(a * a for a in e["v1"])) * (1.0 + 1e-6)).astype(np.float32)
    oracle.raise_vmax(to, e["js"], sp)


@pytest.mark.parametrize("box", [0.0, 40.0])
def test_events_match_oracle_bitwise(gz, oracle, box):
    from oracle.bindings import HostParticles
    pos, v, rng = _setup(n=6000, seed=41)
    print(f"[events] n=6000 box={box}", flush=True)
    hp = HostParticles(*pos)
    tau, t3 = 0.06, 0.1
    e = _events(pos, v, rng, tau, n_ev=150, box=box)  # 300 mass events: more than one chunk of 64
    T = gz.build(gz.Particles(*pos), vel=v)
    to = oracle.build_vel(hp, v)
    try:
        pd, po = _perms(T, oracle, to, hp)
        _apply_device(T, e, pos[3], tau, box)
        _apply_oracle(oracle, to, e, pos[3], tau, box)
        r, o = T.download_reuse(), oracle.reuse(to)
        for f in ("vmax", "dp", "dq"):
            assert np.array_equal(r[f][pd], getattr(o, f)[po]), f
        xd, md = T.download_pads()
        xo, mo = oracle.pads(to)
        assert np.array_equal(xd[pd], xo[po])
        assert np.array_equal(md[pd], mo[po])
        assert np.count_nonzero(xo) > 0 and np.count_nonzero(mo) > 0
        T.refresh_walk_nodes(t3)
        oracle.refresh_walk_nodes(to, t3)
        d, o = T.download(), oracle.reuse(to)
        assert np.array_equal(d["mass"][pd], o.wmass[po])
        com = np.stack([d["cx"], d["cy"], d["cz"]], 1)[pd]
        assert np.array_equal(com, o.wcom[po])
        assert np.array_equal(d["wn_s"][pd], o.wn_s[po])
        assert np.array_equal(d["wn_len"][pd], o.wn_len[po])
    finally:
        oracle.free(to)


def test_events_after_a_kick_feed_and_at_the_device_age(gz, oracle):
    """Kicks first (atomic sums), then events reported with tau < 0: the tree's age on the
    device, advanced by gz_tree_drift."""
    from oracle.bindings import HostParticles
    pos, v, rng = _setup(n=5000, seed=43)
    hp = HostParticles(*pos)
    n = hp.n
    dt = 0.04
    T = gz.build(gz.Particles(*pos), vel=v)
    to = oracle.build_vel(hp, v)
    try:
        pd, po = _perms(T, oracle, to, hp)
        act = np.sort(rng.choice(n, n // 2, replace=False)).astype(np.uint32)
        v1 = tuple(a.copy() for a in v)
        for a in v1:
            a[act] += rng.normal(scale=0.1, size=act.shape[0])
        vl_d = tuple(torch.as_tensor(a.copy(), device="cuda:0") for a in v)
        vl_o = tuple(a.copy() for a in v)
        T.kick_nodes(act, v1, None, vl_d, 0.0)
        oracle.kick_nodes(to, act, v1, pos[3], vl_o, 0.0)
        T.drift(dt)
        e = _events(pos, v1, rng, dt, n_ev=40)
        _apply_device(T, e, pos[3], -1.0, 0.0)
        _apply_oracle(oracle, to, e, pos[3], dt, 0.0)
        r, o = T.download_reuse(), oracle.reuse(to)
        assert np.array_equal(r["vmax"][pd], o.vmax[po])
        for f in ("dp", "dq"):
            ref = getattr(o, f)[po]
            scale = np.abs(ref).max() + 1e-300
            assert np.all(np.abs(r[f][pd] - ref) <= 1e-12 * scale), f
        xd, md = T.download_pads()
        xo, mo = oracle.pads(to)
        assert np.array_equal(xd[pd], xo[po])
        assert np.array_equal(md[pd], mo[po])
    finally:
        oracle.free(to)


def test_pad_keeps_a_displaced_particle_in_the_neighbour_search(gz, oracle):
    from oracle.bindings import HostParticles
    pos, v, rng = _setup(n=4000, seed=47)
    x, y, z, m, soft = (a.copy() for a in pos)
    T = gz.build(gz.Particles(x, y, z, m, soft))
    to = oracle.build(HostParticles(x, y, z, m, soft))
    try:
        # move a few particles well outside their leaves (no velocities: only the pad covers it)
        d = T.download()
        ks = rng.choice(x.shape[0], 8, replace=False).astype(np.uint32)
        side = d["size"][d["leaf_of"][ks]]
        dx = [np.full(8, 0.0) for _ in range(3)]
        dx[0][:] = 3.0 * side + 1e-3
        x2 = x.copy()
        x2[ks] += dx[0]
        P2 = gz.Particles(x2, y, z, m, soft)
        T.update_sources(P2)
        T.raise_pad(ks, dx[0])
        oracle.raise_pad(to, ks, dx[0])
        centres = np.stack([x2[ks], y[ks], z[ks]], 1)
        radius = np.full(8, 1e-4)
        s_d, f_d = gz.ngb_search(T, centres, radius)
        s_o, f_o = oracle.ngb_search(to, HostParticles(x2, y, z, m, soft), centres, radius)
        assert np.array_equal(s_d.cpu().numpy(), s_o)
        assert np.array_equal(f_d.cpu().numpy(), f_o)
        for q in range(8):
            assert ks[q] in f_o[s_o[q]:s_o[q + 1]]
    finally:
        oracle.free(to)


def test_event_errors_and_trees_without_velocities(gz):
    pos, v, rng = _setup(n=2000, seed=49)
    T = gz.build(gz.Particles(*pos))
    n = pos[0].shape[0]
    one = np.ones(1)
    # mass events and shifts need the incremental moments: a no-op without them, as in the
    # reference
    T.mass_events([3], one, [one, one, one], [one, one, one], 0.1)
    T.shift_nodes([3], one, [one, one, one])
    assert not np.any(T.download_pads()[1])
    with pytest.raises(RuntimeError):
        T.raise_pad([n], one)
    with pytest.raises(RuntimeError):
        T.raise_vmax([1], np.array([-1.0], np.float32))
    with pytest.raises(RuntimeError):
        T.raise_vmax([n + 5], np.array([1.0], np.float32))
    Tv = gz.build(gz.Particles(*pos), vel=v)
    with pytest.raises(RuntimeError):
        Tv.mass_events([n], one, [one, one, one], [one, one, one], 0.1)
