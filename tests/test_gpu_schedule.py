"""The pieces of an integrator cycle enqueued without the host, and the walk autotuner.

* gz_tree_drift on a tree built with velocities, after kicks reported through gz_tree_kick_nodes,
  leaves the sorted copy exactly as gz_drift + gz_tree_update_sources do: the walks on the two
  trees are bit for bit the same (the kick feed here raises speed bounds only: the momentum feed's
  FP64 atomic sums round in the order the atomics land, which would make the two trees' refreshed
  centres of mass differ in their last bits); the tree's device age equals the summed drifts, and a
  refresh at tau < 0 (the device age) equals one at that tau on the same tree;
* a cycle of 16 substeps on a reused tree (drift, tree drift, refresh, aold, walk plan, write-back,
  kick, kick feed) captured into one hipGraph: a replay from a saved state gives what the same
  cycle enqueued eagerly gives from that state (to the rounding of the FP64 atomic sums in the
  kick feed);
* gz_walk_autotune: the choice is the fastest candidate it timed, a second call for the same
  workload comes from the context's cache, and the chosen launch parameters give the same forces
  as any other (the parameters change the schedule, not the interactions).
"""
from __future__ import annotations

import numpy as np
import pytest

from gizmo_shared_memory_amd import ics

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def gz():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from gizmo_shared_memory_amd import tree
    return tree


def _pair(gz, n, seed):
    x, y, z, m, soft = ics.hernquist(n=n, eps=0.005, seed=seed)
    rng = np.random.default_rng(seed)
    v = [rng.normal(0.0, 0.3, n) for _ in range(3)]
    return (x, y, z, m, soft), v


def test_tree_drift_equals_drift_and_regather(gz):
    n = 1 << 18
    (x, y, z, m, soft), v = _pair(gz, n, 5)
    rng = np.random.default_rng(1)
    trees = []
    for _ in range(2):
        P = gz.Particles(x, y, z, m, soft)
        vd = [torch.as_tensor(a.copy(), device="cuda:0") for a in v]
        T = gz.build(P, vel=vd)
        trees.append((P, T, vd, [a.clone() for a in vd]))
    tau, dt = 0.0, 2e-3
    for step in range(3):
        act = np.sort(rng.choice(n, n // 5, replace=False)).astype(np.uint32)
        dv = [rng.normal(0.0, 0.05, act.shape[0]) for _ in range(3)]
        tau += dt
        for k, (P, T, vd, vl) in enumerate(trees):
            gz.drift(P, vd, dt)
            if k == 0:
                T.update_sources()
            else:
                T.drift(dt)
            ai = torch.as_tensor(act.astype(np.int64), device="cuda:0")
            for c, d in zip(vd, dv):
                c[ai] += torch.as_tensor(d, device="cuda:0")
            T.kick_nodes(act, vd, tau=tau)
        (P0, T0, _, _), (P1, T1, _, _) = trees
        T0.refresh_walk_nodes(tau)
        T1.refresh_walk_nodes(tau)
        d0, d1 = T0.download(), T1.download()
        for f in ("mass", "cx", "cy", "cz", "wn_s", "wn_len"):
            assert np.array_equal(d0[f], d1[f]), f"step {step}: refreshed {f} differs"
        T1.refresh_walk_nodes(-1.0)  # the device age, advanced by the tree drifts
        d2 = T1.download()
        for f in ("mass", "cx", "cy", "cz", "wn_s", "wn_len"):
            assert np.array_equal(d1[f], d2[f]), f"step {step}: the device age differs from tau"
        tg = T0.device_orderbuf()
        a0 = gz.accel_grouped(T0, P0, tg, 0.7, 1.0, 0, precision="fp32", group_size=32)
        a1 = gz.accel_grouped(T1, P1, tg, 0.7, 1.0, 0, precision="fp32", group_size=32)
        for u, w in zip(a0, a1):
            assert torch.equal(u, w), f"step {step}: walks differ"
        print(f"[schedule] step {step}: tree drift = drift + regather", flush=True)


def test_captured_cycle_replays_the_eager_cycle(gz):
    from gizmo_shared_memory_amd import batching
    S = batching.Schedule(n=1 << 19, n_levels=5, precision="fp32", tune=False)
    print(f"[schedule] active per substep {S.counts}", flush=True)
    S.enqueue_cycle()  # eager: loads every kernel and reserves the scratch the capture reuses
    S.ctx.synchronize()
    saved = S.state()
    # capture one cycle (capturing runs nothing), then replay it from the saved state
    g = gz.Graph(S.ctx)
    g.begin()
    S.enqueue_cycle()
    g.end()
    g.launch()
    S.ctx.synchronize()
    via_graph = S.state()
    # restore, and run the same cycle eagerly: the tree's age and node moments have moved on by
    # one cycle in both, so rebuild the comparison from the saved particle state
    S2 = batching.Schedule(n=1 << 19, n_levels=5, precision="fp32", tune=False)
    S2.enqueue_cycle()
    S2.ctx.synchronize()
    S2.enqueue_cycle()
    S2.ctx.synchronize()
    eager = S2.state()
    for k, (u, w) in enumerate(zip(via_graph, eager)):
        scale = float(torch.max(torch.abs(w))) or 1.0
        diff = float(torch.max(torch.abs(u - w))) / scale
        print(f"[schedule] state {k}: max |graph - eager| / max = {diff:.2e}", flush=True)
        assert diff < 1e-9
    assert not all(torch.equal(a, b) for a, b in zip(saved, via_graph))


def test_autotune_picks_the_fastest_and_caches(gz):
    x, y, z, m, soft = ics.hernquist(n=1 << 20, eps=0.005, seed=9)
    P = gz.Particles(x, y, z, m, soft)
    T = gz.build(P)
    tg = T.device_orderbuf()
    boot = gz.accel_grouped(T, P, tg, 0.7, 1.0, 0)
    aold = 0.0025 * torch.linalg.norm(torch.stack(boot, 1), dim=1)
    choice, rep = gz.autotune(T, P, tg, 0.0, 1.0, aold, precision="fp32", retune=True)
    print(f"[autotune] choice {choice}; timed {rep['points']}", flush=True)
    assert not rep["from_cache"] and len(rep["points"]) >= 5
    fastest = min(rep["points"], key=lambda q: q["ms"])
    assert choice == {k: fastest[k] for k in ("group_size", "stack_limit", "waves_per_cu")}
    again, rep2 = gz.autotune(T, P, tg, 0.0, 1.0, aold, precision="fp32")
    assert rep2["from_cache"] and again == choice and not rep2["points"]
    # launch parameters change the schedule, not the forces
    ref = gz.accel_grouped(T, P, tg, 0.0, 1.0, 0, aold, precision="fp32",
                           group_size=choice["group_size"])
    got = gz.accel_grouped(T, P, tg, 0.0, 1.0, 0, aold, precision="fp32", **choice)
    for u, w in zip(ref, got):
        assert torch.equal(u, w)
