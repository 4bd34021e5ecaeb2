"""The named workloads at their full sizes.

* Plummer, N = 65,536 (the reference's own test_tree case): the whole device path against the
  oracle on every particle, both opening regimes, and the tree itself bit for bit;
* Hernquist, N = 4,194,304: sorted keys, orderbuf, rank, the canonical topology and every node
  moment of the device tree against the oracle's tree bit for bit; the FP64 grouped walk against
  the oracle on the full leaves of 4,096 probe particles (a leaf's targets form one group in the
  all-active walk, so walking just those leaves reproduces the all-active groups), within 1e-6
  relative; and the walk's error against the FP64 direct sum on 65,536 probes;
* clustered, N = 16,777,216: the FP32 group-relative walk against the FP64 walk, both judged by
  the FP64 direct sum on 65,536 probes.
"""
from __future__ import annotations

import numpy as np
import pytest

import types

from gizmo_shared_memory_amd import ics
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


def _host(x, y, z, m, soft):
    from oracle.bindings import HostParticles
    return HostParticles(x, y, z, m, soft)


def _assert_tree_equal(d, o, oracle_skeys):
    """Device tree (downloaded) = oracle tree: keys, order, rank, topology, every moment."""
    assert d["n_live"] == o.n_live and d["nnodes"] == o.nnodes
    assert np.array_equal(d["skey"], oracle_skeys), "sorted keys differ"
    assert np.array_equal(d["orderbuf"], o.orderbuf), "orderbuf differs"
    assert np.array_equal(d["rank"], o.rank), "rank differs"
    assert_same_tree(types.SimpleNamespace(**d), o, fields=BIT_FIELDS, what=("device", "oracle"))


def test_plummer_64k_all_active(gz, oracle):
    x, y, z, m, soft = ics.plummer()
    print(f"[config] plummer n={x.shape[0]}", flush=True)
    P = gz.Particles(x, y, z, m, soft)
    T = gz.build(P)
    hp = _host(x, y, z, m, soft)
    to = oracle.build(hp)
    try:
        o = oracle.tree(to, hp)
        d = T.download()
        _assert_tree_equal(d, o, oracle.skeys(to, o.n_live))
        tg = o.orderbuf
        gx, gy, gz_, c = gz.accel_grouped(T, P, tg, 0.7, 1.0, 0, counts=True)
        ox, oy, oz, oc = oracle.accel_grouped(to, hp, tg, 0.7, 1.0, 0, None, want_counts=True)
        assert [c["nodes_visited"], c["multipole"], c["pairs"]] == [int(v) for v in oc]
        assert _relerr((_np(gx), _np(gy), _np(gz_)), (ox, oy, oz)).max() < 1e-12
        aold = 0.0025 * np.sqrt(ox * ox + oy * oy + oz * oz)
        gx, gy, gz_, c = gz.accel_grouped(T, P, tg, 0.0, 1.0, 0, aold=aold, counts=True)
        ox, oy, oz, oc = oracle.accel_grouped(to, hp, tg, 0.0, 1.0, 0, aold, want_counts=True)
        assert [c["nodes_visited"], c["multipole"], c["pairs"]] == [int(v) for v in oc]
        assert _relerr((_np(gx), _np(gy), _np(gz_)), (ox, oy, oz)).max() < 1e-12
        # shmem::accel, per target, on a sample
        sub = tg[::16].copy()
        gx, gy, gz_ = gz.accel(T, P, sub, 0.7, 1.0)
        ox, oy, oz = oracle.accel(to, hp, sub, 0.7, 1.0)
        assert _relerr((_np(gx), _np(gy), _np(gz_)), (ox, oy, oz)).max() < 1e-12
    finally:
        oracle.free(to)


def test_hernquist_4m_probe_leaves(gz, oracle):
    x, y, z, m, soft = ics.hernquist()
    n = x.shape[0]
    print(f"[config] hernquist n={n}", flush=True)
    P = gz.Particles(x, y, z, m, soft)
    T = gz.build(P)
    tg = T.device_orderbuf()
    boot = gz.accel_grouped(T, P, tg, 0.7, 1.0, 0)
    aold = 0.0025 * torch.linalg.norm(torch.stack(boot, 1), dim=1)  # per Morton slot
    rel = gz.accel_grouped(T, P, tg, 0.0, 1.0, 0, aold=aold)
    d = T.download()
    order, leaf_of = d["orderbuf"], d["leaf_of"]
    leaf_seq = leaf_of[order]  # leaf of each Morton slot; a leaf is a run of slots
    probes = ics.probes(order.shape[0], 4096)
    keep = np.isin(leaf_seq, np.unique(leaf_seq[probes]))
    targets = order[keep]
    print(f"[config] {probes.shape[0]} probes -> {targets.shape[0]} targets in their leaves",
          flush=True)
    hp = _host(x, y, z, m, soft)
    to = oracle.build(hp)
    try:
        o = oracle.tree(to, hp)
        _assert_tree_equal(d, o, oracle.skeys(to, o.n_live))
        print(f"[config] tree bit-identical: {d['nnodes']} nodes, {o.n_live} live", flush=True)
        del o
        ao = _np(aold)[keep]
        ob = oracle.accel_grouped(to, hp, targets, 0.7, 1.0, 0, None)
        orl = oracle.accel_grouped(to, hp, targets, 0.0, 1.0, 0, ao)
    finally:
        oracle.free(to)
    eb = _relerr(tuple(_np(v)[keep] for v in boot), ob)
    er = _relerr(tuple(_np(v)[keep] for v in rel), orl)
    print(f"[config] max rel diff: bootstrap {eb.max():.3g}, relative {er.max():.3g}", flush=True)
    assert eb.max() < 1e-6 and er.max() < 1e-6
    # the walk's own error: 65,536 random probes against the FP64 direct sum
    pr = ics.probes(order.shape[0], 65536)
    pt = tg.view(torch.int32)[torch.as_tensor(pr.astype(np.int64), device="cuda:0")]
    ex = tuple(_np(v) for v in gz.accel_brute(P, pt.view(torch.uint32), 1.0))
    e = _relerr(tuple(_np(v)[pr] for v in rel), ex)
    print(f"[config] relative-criterion walk vs direct sum: median {np.median(e):.3g} "
          f"p99 {np.percentile(e, 99):.3g} max {e.max():.3g}", flush=True)
    assert np.percentile(e, 99) < 1e-2


def test_clustered_16m_fp32_vs_fp64(gz):
    x, y, z, m, soft = ics.clustered()
    n = x.shape[0]
    print(f"[config] clustered n={n}", flush=True)
    P = gz.Particles(x, y, z, m, soft)
    T = gz.build(P)
    tg = T.device_orderbuf()
    boot = gz.accel_grouped(T, P, tg, 0.7, 1.0, 0)
    aold = 0.0025 * torch.linalg.norm(torch.stack(boot, 1), dim=1)
    del boot
    w64 = gz.accel_grouped(T, P, tg, 0.0, 1.0, 0, aold=aold, precision="fp64", counts=True)
    w32 = gz.accel_grouped(T, P, tg, 0.0, 1.0, 0, aold=aold, precision="fp32", counts=True)
    # the same leaf groups in both precisions: the same interaction set
    for k in ("nodes_visited", "multipole", "pairs"):
        assert w32[3][k] == w64[3][k], k
    probes = ics.probes(n, 65536)
    pt = tg.view(torch.int32)[torch.as_tensor(probes.astype(np.int64), device="cuda:0")]
    ex = tuple(_np(v) for v in gz.accel_brute(P, pt.view(torch.uint32), 1.0))
    e64 = _relerr(tuple(_np(v)[probes] for v in w64[:3]), ex)
    e32 = _relerr(tuple(_np(v)[probes] for v in w32[:3]), ex)
    d = _relerr(tuple(_np(v)[probes] for v in w32[:3]), tuple(_np(v)[probes] for v in w64[:3]))
    for name, e in (("fp64", e64), ("fp32", e32), ("fp32-vs-fp64", d)):
        print(f"[config] {name}: median {np.median(e):.3g} p99 {np.percentile(e, 99):.3g} "
              f"max {e.max():.3g}", flush=True)
    # FP32 evaluation adds rounding far below the tree's own error (tolerances of this check)
    assert np.percentile(d, 99) < 1e-5 and d.max() < 1e-4
    assert np.percentile(e32, 99) <= np.percentile(e64, 99) * 1.01 + 1e-6
