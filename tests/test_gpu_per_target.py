"""The per-target walks (gz_tree_accel = shmem::accel, tree.cc:528-582; gz_tree_potential =
shmem::potential, tree.cc:899-944) on the device: 64 Morton-neighbouring targets per wavefront,
one lane each, walking the union of their traversals.

Each lane adds its own target's terms in its own depth-first order, so a target's result does
not depend on which targets share its wave. That gives exact checks beside the oracle's:

* against the oracle: targets in any order, with duplicates, with targets outside the tree
  (massless particles), both opening regimes: within 1e-12 relative (the monopoles take
  1/sqrt(r^2) from a refined hardware estimate instead of a square root and a division);
* a shuffled target list (walked in slot order after a device sort) gives each target exactly
  the value it gets in the sorted list;
* a forest gives each cluster's targets exactly what a tree built over that cluster alone gives.
"""
from __future__ import annotations

import numpy as np
import pytest

from gizmo_shared_memory_amd import ics
from tests import tree_cases

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def gz():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from gizmo_shared_memory_amd import tree
    return tree


def _np(t):
    return t.detach().cpu().numpy()


def _rel_vec(g, o):
    g = np.stack(g, 1)
    o = np.stack(o, 1)
    return np.linalg.norm(g - o, axis=1) / np.maximum(np.linalg.norm(o, axis=1), 1e-300)


@pytest.mark.parametrize("name", ["plummer20k", "tombstones", "collocated", "close_pair", "n17"])
def test_shuffled_targets_match_oracle(gz, oracle, name):
    from oracle.bindings import HostParticles
    x, y, z, m, soft = tree_cases.CASES[name]()
    n = x.shape[0]
    print(f"[per-target] case={name} n={n}", flush=True)
    rng = np.random.default_rng(17)
    # every particle in a random order (massless ones are outside the tree), some twice
    tg = np.concatenate([rng.permutation(n), rng.integers(0, n, n // 10)]).astype(np.uint32)
    P = gz.Particles(x, y, z, m, soft)
    T = gz.build(P)
    hp = HostParticles(x, y, z, m, soft)
    to = oracle.build(hp)
    try:
        a0 = oracle.accel(to, hp, tg, 0.6, 1.0)
        for theta, G in ((0.6, 1.0), (0.0, 2.0)):
            # the geometric criterion, then the relative one with aold from a first walk
            aold = None if theta > 0 else 0.0025 * np.sqrt(a0[0] ** 2 + a0[1] ** 2 + a0[2] ** 2)
            ref = oracle.accel(to, hp, tg, theta, G, aold)
            got = gz.accel(T, P, tg, theta, G, aold)
            e = _rel_vec(tuple(_np(v) for v in got), ref)
            assert e.max() < 1e-12, (theta, e.max())
            phi = _np(gz.potential(T, P, tg, theta, G, aold))
            po = oracle.potential(to, hp, tg, theta, G, aold)
            assert np.all(np.abs(phi - po) <= 1e-12 * np.abs(po)), theta
    finally:
        oracle.free(to)


def test_result_does_not_depend_on_the_target_order(gz):
    x, y, z, m, soft = ics.hernquist(n=1 << 17, eps=0.005, seed=21)
    P = gz.Particles(x, y, z, m, soft)
    T = gz.build(P)
    tg = T.device_orderbuf().view(torch.int32).cpu().numpy().astype(np.uint32)
    perm = np.random.default_rng(3).permutation(tg.shape[0])
    for fn in ("accel", "potential"):
        print(f"[per-target] order independence: {fn}", flush=True)
        a = getattr(gz, fn)(T, P, tg, 0.7, 1.0)
        b = getattr(gz, fn)(T, P, tg[perm], 0.7, 1.0)
        a = a if isinstance(a, tuple) else (a,)
        b = b if isinstance(b, tuple) else (b,)
        for u, v in zip(a, b):
            assert np.array_equal(_np(u)[perm], _np(v))


def test_forest_equals_separate_trees(gz):
    (x, y, z, m, soft), starts = ics.spheres(n_spheres=5, lo_log2=9.0, hi_log2=12.0, eps=0.01,
                                             spacing=100.0, seed=4)
    P = gz.Particles(x, y, z, m, soft)
    F = gz.build(P, clusters=starts)
    n = x.shape[0]
    tg = np.random.default_rng(8).permutation(n).astype(np.uint32)  # clusters mixed in waves
    print(f"[per-target] forest of 5, n={n}", flush=True)
    fa = gz.accel(F, P, tg, 0.6, 1.0)
    fp = gz.potential(F, P, tg, 0.6, 1.0)
    pos = np.empty(n, dtype=np.int64)
    pos[tg] = np.arange(n)
    for c in range(5):
        lo, hi = int(starts[c]), int(starts[c + 1])
        Pc = gz.Particles(x[lo:hi], y[lo:hi], z[lo:hi], m[lo:hi], soft[lo:hi])
        Tc = gz.build(Pc)
        loc = np.arange(hi - lo, dtype=np.uint32)
        ca = gz.accel(Tc, Pc, loc, 0.6, 1.0)
        cp = gz.potential(Tc, Pc, loc, 0.6, 1.0)
        for u, v in zip(fa + (fp,), ca + (cp,)):
            assert np.array_equal(_np(u)[pos[lo:hi]], _np(v)), c


def test_per_target_errors(gz):
    x, y, z, m, soft = tree_cases.tiny(300)
    P = gz.Particles(x, y, z, m, soft)
    T = gz.build(P)
    with pytest.raises(RuntimeError):
        gz.accel(T, P, [0, 300], 0.7)
    with pytest.raises(RuntimeError):
        gz.potential(T, P, [5, 1 << 20], 0.7)
    assert gz.potential(T, P, [], 0.7).numel() == 0
    assert gz.accel(T, P, [], 0.7)[0].numel() == 0


def test_hernquist_1m_sampled(gz, oracle):
    """Both per-target walks on a 1M Hernquist sphere, every particle a target (65,536 waves'
    worth of unions), checked on a random sample against the oracle in both regimes."""
    from oracle.bindings import HostParticles
    x, y, z, m, soft = ics.hernquist(n=1 << 20, eps=0.005, seed=31)
    n = x.shape[0]
    print(f"[per-target] hernquist n={n}", flush=True)
    P = gz.Particles(x, y, z, m, soft)
    T = gz.build(P)
    tg = T.device_orderbuf()
    pick = np.sort(np.random.default_rng(6).choice(n, 2000, replace=False))
    tg_np = tg.view(torch.int32).cpu().numpy().astype(np.uint32)
    hp = HostParticles(x, y, z, m, soft)
    to = oracle.build(hp)
    try:
        # aold for every target (by list position) from the grouped walk at theta = 0.7
        boot = gz.accel_grouped(T, P, tg, 0.7, 1.0, 0)
        aold = 0.0025 * np.linalg.norm(np.stack([_np(v) for v in boot], 1), axis=1)
        assert np.all(aold > 0)
        aold_s = aold[pick]
        for theta, ao, ao_s in ((0.7, None, None), (0.0, aold, aold_s)):
            got = gz.accel(T, P, tg, theta, 1.0, ao)
            g = np.stack([_np(v)[pick] for v in got], 1)
            o = np.stack(oracle.accel(to, hp, tg_np[pick], theta, 1.0, ao_s), 1)
            e = np.linalg.norm(g - o, axis=1) / np.linalg.norm(o, axis=1)
            assert e.max() < 1e-12, (theta, e.max())
            phi = _np(gz.potential(T, P, tg, theta, 1.0, ao))[pick]
            po = oracle.potential(to, hp, tg_np[pick], theta, 1.0, ao_s)
            assert np.all(np.abs(phi - po) <= 1e-12 * np.abs(po)), theta
    finally:
        oracle.free(to)
