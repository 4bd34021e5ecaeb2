"""The caller side of a force evaluation and the root-box options, on the device.

* root box with RANDOMIZE_GRAVTREE (a seed) and SHMEM_TREE_SHIFT (three fractions),
  tree.cc:381-407: the device tree equals the oracle's bit for bit (keys, order, topology, every
  moment), and the FP64 grouped walk on it equals the oracle's walk (counts, 1e-12);
* the write-back a_grav[targets[k]] = a[k] (mfm.cc:828-831): bit for bit;
* the KDK kick (mfm.cc:3749-3767): two kicks in a row, the second closing the half kick the
  first opened, against the same expressions evaluated in numpy, bit for bit (the library is
  built without contraction, so the operations round identically).
"""
from __future__ import annotations

import types

import numpy as np
import pytest

from tests import tree_cases
from tests.treecmp import BIT_FIELDS, assert_same_tree

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


def _relerr(g, o):
    g = np.stack(g, 1)
    o = np.stack(o, 1)
    return np.linalg.norm(g - o, axis=1) / np.maximum(np.linalg.norm(o, axis=1), 1e-300)


ROOTS = [(17, (0.0, 0.0, 0.0)), (-1, (0.13, -0.07, 0.21)), (123456789, (-0.3, 0.0, 0.05))]


@pytest.mark.parametrize("seed,shift", ROOTS)
def test_seeded_and_shifted_root_matches_oracle(gz, oracle, seed, shift):
    from oracle.bindings import HostParticles
    case = tree_cases.plummer_small(12000, seed=44, eps=0.01)
    print(f"[root] seed={seed} shift={shift}", flush=True)
    x, y, z, m, soft = case
    P = gz.Particles(x, y, z, m, soft)
    T = gz.build(P, randomize_seed=seed, shift=shift)
    d = T.download()
    hp = HostParticles(x, y, z, m, soft)
    to = oracle.build(hp, seed, shift=shift)
    try:
        o = oracle.tree(to, hp)
        assert np.array_equal(d["skey"], oracle.skeys(to, o.n_live)), "sorted keys differ"
        assert np.array_equal(d["orderbuf"], o.orderbuf)
        assert np.array_equal(d["rank"], o.rank)
        assert_same_tree(types.SimpleNamespace(**d), o, fields=BIT_FIELDS,
                         what=("device", "oracle"))
        tg = o.orderbuf
        gx, gy, gz_, c = gz.accel_grouped(T, P, tg, 0.7, 1.0, 0, counts=True)
        ox, oy, oz, oc = oracle.accel_grouped(to, hp, tg, 0.7, 1.0, 0, None, want_counts=True)
        assert [c["nodes_visited"], c["multipole"], c["pairs"]] == [int(v) for v in oc]
        assert _relerr((_np(gx), _np(gy), _np(gz_)), (ox, oy, oz)).max() < 1e-12
    finally:
        oracle.free(to)


def test_scatter_accel(gz):
    rng = np.random.default_rng(3)
    n = 50000
    targets = rng.permutation(n)[:31337].astype(np.uint32)
    a = [rng.standard_normal(targets.shape[0]) for _ in range(3)]
    init = rng.standard_normal((n, 3))
    ag = torch.as_tensor(init.copy(), device="cuda:0")
    gz.scatter_accel(targets, *a, ag)
    want = init.copy()
    want[targets.astype(np.int64)] = np.stack(a, 1)
    assert np.array_equal(_np(ag), want)


def test_kick_two_steps_matches_kdk(gz):
    """v += a (close_dt + open_dt) in the reference's two half kicks, pending_half_kick handed
    from one kick of a particle to its next, dt by particle (mfm.cc:3709-3766)."""
    rng = np.random.default_rng(9)
    n = 40000
    v0 = rng.standard_normal((3, n))
    ag = rng.standard_normal((n, 3))
    ah = rng.standard_normal((n, 3)) * 0.1
    dt = 2.0 ** -rng.integers(4, 12, n).astype(np.float64)
    pend = 0.5 * dt * (rng.uniform(size=n) < 0.7)
    act1 = np.sort(rng.choice(n, 12345, replace=False)).astype(np.uint32)
    act2 = rng.permutation(act1)[:5000].astype(np.uint32)  # a subset, in another order

    dev = lambda a: torch.as_tensor(np.ascontiguousarray(a), device="cuda:0")  # noqa: E731
    v = [dev(v0[k].copy()) for k in range(3)]
    p = dev(pend.copy())
    gz.kick(act1, dev(ag), dev(dt), p, v, a_extra=dev(ah))
    # the second kick uses new accelerations and no extra term
    ag2 = rng.standard_normal((n, 3))
    gz.kick(act2, dev(ag2), dev(dt), p, v)

    hv = v0.copy()
    hp = pend.copy()
    for act, acc in ((act1, ag + ah), (act2, ag2)):
        i = act.astype(np.int64)
        close = hp[i]
        open_ = 0.5 * dt[i]
        for k in range(3):
            hv[k, i] = hv[k, i] + acc[i, k] * close
            hv[k, i] = hv[k, i] + acc[i, k] * open_
        hp[i] = open_
    for k in range(3):
        assert np.array_equal(_np(v[k]), hv[k])
    assert np.array_equal(_np(p), hp)
    # particles never active keep their state
    idle = np.setdiff1d(np.arange(n), act1)
    assert np.array_equal(_np(v[0])[idle], v0[0][idle])
