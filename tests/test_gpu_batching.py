"""Walk plans: segmented target lists walked in one launch, and graph-captured replays."""
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


@pytest.fixture(scope="module")
def setup(gz):
    x, y, z, m, soft = tree_cases.plummer_small(30000, seed=23, eps=0.01)
    P = gz.Particles(x, y, z, m, soft)
    T = gz.build(P)
    level = ics.timestep_bins(x, y, z, n_levels=5, seed=3)
    tg = T.device_orderbuf()
    boot = gz.accel_grouped(T, P, tg, 0.7, 1.0, 0)
    a_prev = torch.stack(boot, 1)  # Morton slot order
    # per-particle |a| for aold of any target list
    amag = torch.empty(P.n, dtype=torch.float64, device="cuda:0")
    amag[tg.view(torch.int32).long()] = torch.linalg.norm(a_prev, dim=1)
    return P, T, level, amag


def _substep_targets(gz, T, level, s):
    act = ics.active_for_substep(level, s, n_levels=5)
    return gz.order_targets(T, act)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_segmented_plan_equals_separate_walks(gz, setup, prec):
    P, T, level, amag = setup
    lists = [_substep_targets(gz, T, level, s) for s in range(16)]
    seg = np.concatenate([[0], np.cumsum([int(v.shape[0]) for v in lists])])
    allt = torch.cat([v.view(torch.int32) for v in lists]).view(torch.uint32)
    aold = 0.0025 * amag[allt.view(torch.int32).long()]
    plan = gz.WalkPlan(T, P, allt, segments=seg, precision=prec, group_size=32)
    plan.launch(0.0, 1.0, aold)
    st = plan.status()
    assert st["retries"] == 0 or st["retries"] > 0  # reported
    got = [v.clone() for v in (plan.ax, plan.ay, plan.az)]
    for s in range(16):
        lo, hi = int(seg[s]), int(seg[s + 1])
        if hi == lo:
            continue
        ref = gz.accel_grouped(T, P, lists[s], 0.0, 1.0, 0, aold=aold[lo:hi], precision=prec,
                               group_size=32)
        for u, v in zip(got, ref):
            assert torch.equal(u[lo:hi], v), s


def test_graph_replays_match_eager(gz, setup):
    P, T, level, amag = setup
    plans, aolds = [], []
    for s in (0, 1, 2, 4):
        tl = _substep_targets(gz, T, level, s)
        plans.append(gz.WalkPlan(T, P, tl, precision="fp32", group_size=32))
        aolds.append(0.0025 * amag[tl.view(torch.int32).long()])
    eager = []
    for p, a in zip(plans, aolds):
        p.launch(0.0, 1.0, a)
        T.ctx.synchronize()
        eager.append([v.clone() for v in (p.ax, p.ay, p.az)])
    g = gz.Graph(T.ctx)
    g.begin()
    for p, a in zip(plans, aolds):
        p.launch(0.0, 1.0, a)
    g.end()
    for rep in range(3):
        for p in plans:  # clobber the outputs between replays
            for v in (p.ax, p.ay, p.az):
                v.fill_(float("nan"))
        torch.cuda.synchronize()
        g.launch()
        T.ctx.synchronize()
        for p, e in zip(plans, eager):
            for u, v in zip((p.ax, p.ay, p.az), e):
                assert torch.equal(u, v), rep
        # eager work in between replays
        _ = gz.accel_grouped(T, P, T.device_orderbuf()[:1000], 0.7, 1.0, 0)


def test_plan_with_learned_model_matches_walk(gz, setup):
    from gizmo_shared_memory_amd import learned
    P, T, level, amag = setup
    M = learned.load_model()
    tl = _substep_targets(gz, T, level, 0)
    aold = 0.0025 * amag[tl.view(torch.int32).long()]
    plan = gz.WalkPlan(T, P, tl, precision="fp32", group_size=32, model=M)
    plan.launch(0.0, 1.0, aold)
    st = plan.status()
    ref = gz.accel_grouped(T, P, tl, 0.0, 1.0, 0, aold=aold, precision="fp32", group_size=32,
                           model=M, counts=True)
    assert st["fallback_groups"] == ref[3]["fallback_groups"]
    for u, v in zip((plan.ax, plan.ay, plan.az), ref[:3]):
        assert torch.equal(u, v)


@pytest.mark.parametrize("prec,gs", [("fp64", 0), ("fp32", 32)])
def test_forest_equals_separate_trees(gz, prec, gs):
    """A forest built over several independent clusters walks each cluster exactly as a tree
    built over that cluster alone."""
    (x, y, z, m, soft), starts = ics.spheres(n_spheres=5, lo_log2=9.0, hi_log2=12.0, eps=0.01,
                                             spacing=100.0, seed=4)
    P = gz.Particles(x, y, z, m, soft)
    F = gz.build(P, clusters=starts)
    assert F.info.nnodes >= 5
    tg = F.device_orderbuf()
    fa = gz.accel_grouped(F, P, tg, 0.6, 1.0, 0, precision=prec, group_size=gs, counts=True)
    order = tg.view(torch.int32).cpu().numpy()
    pos = np.empty(P.n, dtype=np.int64)
    pos[order] = np.arange(order.shape[0])
    tot = 0
    for c in range(5):
        lo, hi = int(starts[c]), int(starts[c + 1])
        Pc = gz.Particles(x[lo:hi], y[lo:hi], z[lo:hi], m[lo:hi], soft[lo:hi])
        Tc = gz.build(Pc)
        tc = Tc.device_orderbuf()
        ca = gz.accel_grouped(Tc, Pc, tc, 0.6, 1.0, 0, precision=prec, group_size=gs,
                              counts=True)
        oc = tc.view(torch.int32).cpu().numpy() + lo
        for u, v in zip(fa[:3], ca[:3]):
            assert np.array_equal(u.cpu().numpy()[pos[oc]], v.cpu().numpy()), c
        tot += ca[3]["multipole"] + ca[3]["pairs"]
    assert fa[3]["multipole"] + fa[3]["pairs"] == tot
