"""Device parity: the HIP build, walks and direct sum against the C oracle on the same inputs.

* keys, sorted order, topology and every node moment: bit for bit (both sides sort stably by
  (key, index) and compute moments with contraction off);
* grouped walk in FP64: the same interaction set (node visits, multipole and pair counts equal)
  and accelerations within 1e-12 relative (only the summation order differs);
* direct sum: same order and expressions as the oracle, within 1e-14 relative.
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


def _device_tree(gz, case):
    x, y, z, m, soft = case
    P = gz.Particles(x, y, z, m, soft, device="cuda:0")
    T = gz.build(P)
    return P, T


def _ns(d):
    return types.SimpleNamespace(**d)


def _relerr(g, o):
    g = np.stack(g, 1)
    o = np.stack(o, 1)
    den = np.maximum(np.linalg.norm(o, axis=1), 1e-300)
    return np.linalg.norm(g - o, axis=1) / den


def _np(t):
    return t.detach().cpu().numpy()


@pytest.mark.parametrize("name", sorted(tree_cases.CASES))
def test_build_matches_oracle_bitwise(gz, oracle, name):
    case = tree_cases.CASES[name]()
    print(f"[build] case={name} n={case[0].shape[0]}", flush=True)
    P, T = _device_tree(gz, case)
    d = T.download()
    hp = _host(case)
    to = oracle.build(hp)
    try:
        o = oracle.tree(to, hp)
        skeys = oracle.skeys(to, o.n_live)
        assert d["n_live"] == o.n_live
        assert np.array_equal(d["skey"], skeys), "sorted keys differ"
        assert np.array_equal(d["orderbuf"], o.orderbuf), "orderbuf differs"
        assert np.array_equal(d["rank"], o.rank), "rank differs"
        assert_same_tree(_ns(d), o, fields=BIT_FIELDS, what=("device", "oracle"))
        # children are contiguous and in octant order in the device numbering
        first, nch = d["first"], d["nchild"]
        for i in np.nonzero(first >= 0)[0][:2000]:
            kids = np.arange(first[i], first[i] + nch[i])
            assert np.all(d["parent"][kids] == i)
            assert np.all(np.diff(d["plo"][kids]) > 0)
    finally:
        oracle.free(to)


def _host(case):
    from oracle.bindings import HostParticles
    x, y, z, m, soft = case
    return HostParticles(x, y, z, m, soft)


WALK_CASES = ["plummer20k", "collocated", "tombstones", "close_pair", "n1", "n2", "n17", "n300"]


@pytest.mark.parametrize("name", WALK_CASES)
@pytest.mark.parametrize("batch", [0, 8])
def test_accel_grouped_fp64_matches_oracle(gz, oracle, name, batch):
    case = tree_cases.CASES[name]()
    print(f"[walk] case={name} batch={batch}", flush=True)
    P, T = _device_tree(gz, case)
    hp = _host(case)
    to = oracle.build(hp)
    try:
        tg = oracle.tree(to, hp).orderbuf
        # bootstrap: geometric theta = 0.7, no aold
        gx, gy, gz_, c = gz.accel_grouped(T, P, tg, 0.7, 1.0, batch, counts=True)
        ox, oy, oz, oc = oracle.accel_grouped(to, hp, tg, 0.7, 1.0, batch, None, want_counts=True)
        assert [c["nodes_visited"], c["multipole"], c["pairs"]] == [int(v) for v in oc]
        if tg.size:
            assert _relerr((_np(gx), _np(gy), _np(gz_)), (ox, oy, oz)).max() < 1e-12
        # relative criterion, aold from the bootstrap (ErrTolForceAcc = 0.0025)
        aold = 0.0025 * np.sqrt(ox * ox + oy * oy + oz * oz)
        gx, gy, gz_, c = gz.accel_grouped(T, P, tg, 0.0, 1.0, batch, aold=aold, counts=True)
        ox, oy, oz, oc = oracle.accel_grouped(to, hp, tg, 0.0, 1.0, batch, aold,
                                              want_counts=True)
        assert [c["nodes_visited"], c["multipole"], c["pairs"]] == [int(v) for v in oc]
        if tg.size:
            assert _relerr((_np(gx), _np(gy), _np(gz_)), (ox, oy, oz)).max() < 1e-12
    finally:
        oracle.free(to)


def test_accel_grouped_random_targets(gz, oracle):
    case = tree_cases.plummer_small(8000, seed=21)
    P, T = _device_tree(gz, case)
    hp = _host(case)
    to = oracle.build(hp)
    try:
        rng = np.random.default_rng(5)
        tg = rng.choice(hp.n, 777, replace=False).astype(np.uint32)
        gx, gy, gz_, c = gz.accel_grouped(T, P, tg, 0.5, 2.0, 0, counts=True)
        ox, oy, oz, oc = oracle.accel_grouped(to, hp, tg, 0.5, 2.0, 0, None, want_counts=True)
        assert [c["nodes_visited"], c["multipole"], c["pairs"]] == [int(v) for v in oc]
        assert _relerr((_np(gx), _np(gy), _np(gz_)), (ox, oy, oz)).max() < 1e-12
    finally:
        oracle.free(to)


def test_accel_per_target_matches_oracle(gz, oracle):
    case = tree_cases.plummer_small(6000, seed=9, eps=0.05)
    P, T = _device_tree(gz, case)
    hp = _host(case)
    to = oracle.build(hp)
    try:
        tg = np.arange(0, hp.n, 7, dtype=np.uint32)
        gx, gy, gz_ = gz.accel(T, P, tg, 0.5, 1.0)
        ox, oy, oz = oracle.accel(to, hp, tg, 0.5, 1.0)
        assert _relerr((_np(gx), _np(gy), _np(gz_)), (ox, oy, oz)).max() < 1e-12
    finally:
        oracle.free(to)


@pytest.mark.parametrize("name", ["plummer20k", "collocated", "tombstones", "n2"])
def test_direct_sum_matches_oracle(gz, oracle, name):
    case = tree_cases.CASES[name]()
    hp = _host(case)
    P = gz.Particles(*case, device="cuda:0")
    tg = np.arange(0, hp.n, max(1, hp.n // 1500), dtype=np.uint32)
    gx, gy, gz_ = gz.accel_brute(P, tg, 1.0)
    ox, oy, oz = oracle.accel_brute(hp, tg, 1.0)
    g = np.stack([_np(gx), _np(gy), _np(gz_)], 1)
    o = np.stack([ox, oy, oz], 1)
    exact = np.all(g.view(np.uint64) == o.view(np.uint64))
    rel = _relerr((g[:, 0], g[:, 1], g[:, 2]), (ox, oy, oz))
    assert exact or rel.max() < 1e-14, rel.max()


def test_fp32_walk_same_interactions_and_close(gz, oracle):
    case = tree_cases.plummer_small(20000, seed=3, eps=0.01)
    P, T = _device_tree(gz, case)
    tg = T.device_orderbuf()
    d64 = gz.accel_grouped(T, P, tg, 0.7, 1.0, 0, counts=True)
    d32 = gz.accel_grouped(T, P, tg, 0.7, 1.0, 0, precision="fp32", counts=True)
    for k in ("nodes_visited", "multipole", "pairs"):
        assert d64[3][k] == d32[3][k]
    rel = _relerr(tuple(_np(v) for v in d32[:3]), tuple(_np(v) for v in d64[:3]))
    assert np.median(rel) < 1e-6 and rel.max() < 1e-4, (np.median(rel), rel.max())


@pytest.mark.parametrize("group_size", [0, 32, 64])
def test_fp32_walk_matches_fp64_with_mixed_softening(gz, group_size):
    """Softenings spread over two decades, relative criterion: the FP32 walk moves a leaf
    particle to its monopole list only when the particle is outside every target's kernel, so it
    makes exactly the FP64 walk's interactions and stays within FP32 rounding of its forces."""
    rng = np.random.default_rng(12)
    x, y, z, m, _ = tree_cases.plummer_small(20000, seed=12, eps=0.01)
    soft = 10.0 ** rng.uniform(-3.0, -1.0, x.shape[0])
    P, T = _device_tree(gz, (x, y, z, m, soft))
    tg = T.device_orderbuf()
    boot = gz.accel_grouped(T, P, tg, 0.7, 1.0, 0)
    aold = 0.0025 * torch.linalg.norm(torch.stack(boot, 1), dim=1)
    d64 = gz.accel_grouped(T, P, tg, 0.0, 1.0, 0, aold=aold, group_size=group_size, counts=True)
    d32 = gz.accel_grouped(T, P, tg, 0.0, 1.0, 0, aold=aold, group_size=group_size,
                           precision="fp32", counts=True)
    for k in ("nodes_visited", "multipole", "pairs"):
        assert d64[3][k] == d32[3][k], k
    rel = _relerr(tuple(_np(v) for v in d32[:3]), tuple(_np(v) for v in d64[:3]))
    assert np.median(rel) < 1e-6 and rel.max() < 1e-4, (np.median(rel), rel.max())


def test_fp32_walk_precision_far_from_origin(gz):
    """Group-relative FP32 keeps its accuracy when the system sits far from the origin."""
    x, y, z, m, soft = tree_cases.plummer_small(8000, seed=4, eps=0.01)
    off = 1.0e4
    P0 = gz.Particles(x, y, z, m, soft)
    P1 = gz.Particles(x + off, y - off, z + 0.5 * off, m, soft)
    T0, T1 = gz.build(P0), gz.build(P1)
    tg = T0.device_orderbuf()
    a0 = gz.accel_grouped(T0, P0, tg, 0.7, 1.0, 0, precision="fp64")
    a1 = gz.accel_grouped(T1, P1, T1.device_orderbuf(), 0.7, 1.0, 0, precision="fp32")
    # map the second run's target order onto the first's
    o0 = _np(tg.view(torch.int32))
    o1 = _np(T1.device_orderbuf().view(torch.int32))
    pos = np.empty(P0.n, dtype=np.int64)
    pos[o1] = np.arange(o1.shape[0])
    g1 = [(_np(v))[pos[o0]] for v in a1]
    rel = _relerr(tuple(g1), tuple(_np(v) for v in a0))
    assert np.percentile(rel, 99) < 1e-2


def test_persistent_matches_one_group_per_workgroup(gz):
    case = tree_cases.plummer_small(20000, seed=8, eps=0.01)
    P, T = _device_tree(gz, case)
    tg = T.device_orderbuf()
    a = gz.accel_grouped(T, P, tg, 0.7, 1.0, 0, persistent=False)
    b = gz.accel_grouped(T, P, tg, 0.7, 1.0, 0, persistent=True)
    for u, v in zip(a, b):
        assert torch.equal(u, v)


@pytest.mark.parametrize("gs", [16, 32, 64, 128])
def test_group_size_walk_is_accurate(gz, gs):
    """Larger groups only open more (the box test is conservative): force error vs direct sum
    stays at or below the leaf-grouped walk's."""
    case = tree_cases.plummer_small(20000, seed=12, eps=0.01)
    P, T = _device_tree(gz, case)
    tg = T.device_orderbuf()
    probe = tg[::50]
    ex = gz.accel_brute(P, probe)
    base = gz.accel_grouped(T, P, tg, 0.7, 1.0, 0, counts=True)
    big = gz.accel_grouped(T, P, tg, 0.7, 1.0, 0, group_size=gs, counts=True)
    e_base = _relerr(tuple(_np(v)[::50] for v in base[:3]), tuple(_np(v) for v in ex))
    e_big = _relerr(tuple(_np(v)[::50] for v in big[:3]), tuple(_np(v) for v in ex))
    assert big[3]["groups"] <= base[3]["groups"]
    assert np.percentile(e_big, 99) <= np.percentile(e_base, 99) * 1.05 + 1e-12


def test_empty_and_errors(gz):
    x, y, z, m, soft = tree_cases.tiny(0)
    P = gz.Particles(x, y, z, m, soft)
    T = gz.build(P)
    assert T.nnodes == 1 and T.n_live == 0
    ax, ay, az = gz.accel_grouped(T, P, np.zeros(0, dtype=np.uint32), 0.7)
    assert ax.numel() == 0
    from gizmo_shared_memory_amd._lib import GzError
    case = tree_cases.tiny(100)
    P2 = gz.Particles(*case)
    with pytest.raises(GzError):
        gz.accel_grouped(T, P2, np.arange(10, dtype=np.uint32), 0.7)  # tree of another set
    T2 = gz.build(P2)
    with pytest.raises(GzError):
        gz.accel_grouped(T2, P2, np.array([3, 100], dtype=np.uint32), 0.7)  # index >= n
    with pytest.raises(GzError):
        gz.accel_grouped(T2, P2, np.arange(10, dtype=np.uint32), 0.7, precision="fp64",
                         group_size=-1)


def test_arena_growth_gives_same_tree(gz, oracle):
    case = tree_cases.dense_cloud_in_box()
    x, y, z, m, soft = case
    P = gz.Particles(*case)
    T = gz.build(P)
    hp = _host(case)
    to = oracle.build(hp)
    try:
        assert_same_tree(_ns(T.download()), oracle.tree(to, hp), what=("device", "oracle"))
    finally:
        oracle.free(to)
    from gizmo_shared_memory_amd._lib import GzError
    with pytest.raises(GzError):
        gz.build(P, max_nodes=100)


def test_order_targets_and_aold(gz, oracle):
    case = tree_cases.with_tombstones()
    P, T = _device_tree(gz, case)
    d = T.download()
    rng = np.random.default_rng(2)
    active = rng.choice(P.n, 500, replace=False).astype(np.uint32)
    tg = _np(gz.order_targets(T, active).view(torch.int32)).astype(np.uint32)
    live = active[case[3][active] > 0]
    expect = live[np.argsort(d["rank"][live])]
    assert np.array_equal(tg, expect)
    a_prev = torch.as_tensor(rng.normal(size=(P.n, 3)), device="cuda:0")
    ao = _np(gz.aold_from_accel(T, a_prev, tg, 0.0025, 2.0))
    ap = _np(a_prev)[tg]
    np.testing.assert_allclose(ao, 0.0025 * np.sqrt((ap * ap).sum(1)) / 2.0, rtol=1e-15)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_stack_spill_rewalk_is_identical(gz, prec):
    """Groups whose node stack outgrows the on-chip stack are walked again with a global-memory
    stack; forcing a tiny on-chip stack must not change a single bit of the result."""
    case = tree_cases.plummer_small(20000, seed=31, eps=0.01)
    P, T = _device_tree(gz, case)
    tg = T.device_orderbuf()
    a = gz.accel_grouped(T, P, tg, 0.5, 1.0, 0, precision=prec, counts=True)
    b = gz.accel_grouped(T, P, tg, 0.5, 1.0, 0, precision=prec, counts=True, stack_limit=16)
    assert a[3]["retries"] == 0 and b[3]["retries"] > 0
    assert a[3]["max_stack"] <= 1024
    for k in ("nodes_visited", "multipole", "pairs"):
        assert a[3][k] == b[3][k]
    for u, v in zip(a[:3], b[:3]):
        assert torch.equal(u, v)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_stack_spill_tier_is_identical(gz, prec):
    """A persistent walk moves the bottom of an overfull on-chip stack to global memory and back;
    the logical stack is unchanged, so the result is bit for bit the unlimited one, with no
    group walked twice. The per-group launch has no spill tier and walks such groups again."""
    case = tree_cases.plummer_small(20000, seed=31, eps=0.01)
    P, T = _device_tree(gz, case)
    tg = T.device_orderbuf()
    a = gz.accel_grouped(T, P, tg, 0.2, 1.0, 0, precision=prec, counts=True)
    b = gz.accel_grouped(T, P, tg, 0.2, 1.0, 0, precision=prec, counts=True, stack_limit=512)
    c = gz.accel_grouped(T, P, tg, 0.2, 1.0, 0, precision=prec, counts=True, stack_limit=512,
                         persistent=False)
    assert a[3]["max_stack"] > 512  # the limited walks had to spill
    assert b[3]["retries"] == 0 and c[3]["retries"] > 0
    for r in (b, c):
        for k in ("nodes_visited", "multipole", "pairs", "max_stack"):
            assert a[3][k] == r[3][k], k
        for u, v in zip(a[:3], r[:3]):
            assert torch.equal(u, v)


def test_packed_groups_are_compact_and_cover_targets(gz):
    case = tree_cases.plummer_small(20000, seed=13, eps=0.01)
    P, T = _device_tree(gz, case)
    tg = T.device_orderbuf()
    plain = gz.accel_grouped(T, P, tg, 0.7, 1.0, 0, group_size=64, counts=True)
    assert plain[3]["groups"] < T.info.nleaves
    ex = gz.accel_brute(P, tg[::97])
    e = _relerr(tuple(_np(v)[::97] for v in plain[:3]), tuple(_np(v) for v in ex))
    assert np.percentile(e, 99) < 2e-2


@pytest.mark.parametrize("name", ["plummer20k", "tombstones", "close_pair", "n17"])
def test_potential_matches_oracle(gz, oracle, name):
    """shmem::potential (tree.cc:899-944) in FP64, per target, both opening regimes."""
    case = tree_cases.CASES[name]()
    print(f"[potential] case={name}", flush=True)
    P, T = _device_tree(gz, case)
    hp = _host(case)
    to = oracle.build(hp)
    try:
        tg = oracle.tree(to, hp).orderbuf
        for theta, aold in ((0.7, None), (0.0, np.full(tg.shape[0], 1e-3))):
            g = _np(gz.potential(T, P, tg, theta, 2.0, aold))
            o = oracle.potential(to, hp, tg, theta, 2.0, aold)
            if tg.size:
                assert np.all(np.abs(g - o) <= 1e-12 * np.abs(o)), np.max(np.abs(g - o) / np.abs(o))
    finally:
        oracle.free(to)
