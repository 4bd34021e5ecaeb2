"""audit_tree's invariants (shmem/tree.cc:973-1119) on device trees: fresh builds at the workload
sizes (Hernquist 4,194,304; halo+disk 8,388,608; clustered 16,777,216), and a tree reused across
steps (built with velocities, particles drifted, velocities kicked, the kick fed into the nodes,
walk nodes refreshed), audited with its speed bounds and age. The audit is the oracle's
restatement, pinned against the reference's audit in tests/test_oracle_audit.py."""
from __future__ import annotations

import numpy as np
import pytest

from gizmo_shared_memory_amd import ics
from oracle.bindings import HostParticles

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def gz():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from gizmo_shared_memory_amd import tree
    return tree


WORKLOADS = {
    "hernquist_4m": lambda: ics.hernquist(),
    "halo_disk_8m": lambda: ics.halo_disk(),
    "clustered_16m": lambda: ics.clustered(),
}


@pytest.mark.parametrize("name", sorted(WORKLOADS))
def test_fresh_device_tree_passes_audit(gz, oracle, name):
    x, y, z, m, soft = WORKLOADS[name]()
    print(f"[audit] {name} n={x.shape[0]}", flush=True)
    T = gz.build(gz.Particles(x, y, z, m, soft))
    d = T.download()
    del T
    torch.cuda.empty_cache()
    bad, msg = oracle.audit_tree(d, HostParticles(x, y, z, m, soft))
    print(f"[audit] {name}: {d['nnodes']} nodes, {bad} violations {msg}", flush=True)
    assert bad == 0, msg


def test_reused_device_tree_passes_audit(gz, oracle):
    rng = np.random.default_rng(31)
    x, y, z, m, soft = ics.hernquist(n=1 << 20, eps=0.005, seed=8)
    n = x.shape[0]
    v = [rng.normal(0.0, 0.3, n) for _ in range(3)]
    P = gz.Particles(x, y, z, m, soft)
    vd = [torch.as_tensor(a.copy(), device="cuda:0") for a in v]
    vlast = [a.clone() for a in vd]
    T = gz.build(P, vel=vd)
    tau, dt = 0.0, 1e-3
    for step in range(4):
        # drift everyone, kick a random quarter, feed the kick to the nodes, refresh
        for c, vc in zip((P.x, P.y, P.z), vd):
            c.add_(vc * dt)
        tau += dt
        T.update_sources()
        act = np.sort(rng.choice(n, n // 4, replace=False)).astype(np.uint32)
        ai = torch.as_tensor(act.astype(np.int64), device="cuda:0")
        for vc in vd:
            vc[ai] += torch.as_tensor(rng.normal(0.0, 0.05, act.shape[0]), device="cuda:0")
        T.kick_nodes(act, vd, vlast=vlast, tau=tau)
        T.refresh_walk_nodes(tau)
        d = T.download()
        r = T.download_reuse()
        hp = HostParticles(P.x.cpu().numpy(), P.y.cpu().numpy(), P.z.cpu().numpy(), m, soft)
        bad, msg = oracle.audit_tree(d, hp, vel=[a.cpu().numpy() for a in vd], vmax=r["vmax"],
                                     tau=tau)
        print(f"[audit] reused step {step} tau={tau:g}: {bad} violations {msg}", flush=True)
        assert bad == 0, msg
