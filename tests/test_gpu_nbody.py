"""A gravity-only N-body run on the device, end to end (the reference engine's KDK step with
hierarchical timesteps and a reused tree, gizmo_shared_memory_amd/nbody.py), judged by the
reference's own Plummer test (test/plummer/test_plummer.py:323-349, plummer.params): N = 2^15,
G = M = a = 1, softening 0.1 (Plummer-equivalent, kernel extent 0.28), ErrTolIntAccuracy 0.005,
ErrTolForceAcc 0.0025, MaxSizeTimestep 0.1, TreeDomainUpdateFrequency 0.1, run to t = 48:
|dE| / KE_0 < 0.04, the 10/50/90% Lagrange radii within 0.20/0.15/0.30 of their start, the
centre-of-mass drift under the sanity ceiling, and audit_tree's invariants on the reused tree."""
from __future__ import annotations

import math

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


def test_drift_kernel(gz):
    rng = np.random.default_rng(4)
    n = 10000
    x, y, z = (rng.normal(size=n) for _ in range(3))
    m = np.full(n, 1.0 / n)
    m[::13] = 0.0  # tombstones stay put
    v = [rng.normal(size=n) for _ in range(3)]
    P = gz.Particles(x, y, z, m, None)
    gz.drift(P, [torch.as_tensor(a, device="cuda:0") for a in v], 0.037)
    live = m > 0
    for got, c, vc in zip((P.x, P.y, P.z), (x, y, z), v):
        want = np.where(live, c + vc * 0.037, c)
        assert np.array_equal(got.cpu().numpy(), want)


# "reference": plummer.params as it is, where more than TreeDomainUpdateFrequency of the particles
# are active at every sync, so the tree is rebuilt before every walk (mfm.cc:3333-3336);
# "reused": the same run with a rebuild only after four particle counts of forces or a drift pad of
# two softenings, so most walks run on a reused tree (kick feed, drifted sources, refreshed nodes)
RUNS = {"reference": dict(tree_update_freq=0.1, tree_pad_frac=0.25),
        "reused": dict(tree_update_freq=4.0, tree_pad_frac=2.0)}


@pytest.mark.parametrize("run", sorted(RUNS))
def test_plummer_equilibrium_run(gz, oracle, run):
    from gizmo_shared_memory_amd import nbody
    from oracle.bindings import HostParticles
    x, y, z, vx, vy, vz, m, soft = ics.plummer_equilibrium(n=1 << 15, eps=0.28, seed=42)
    print(f"[nbody] run {run}", flush=True)
    sim = nbody.GravitySim(x, y, z, vx, vy, vz, m, soft, nbody.Params(
        dt_max=0.1, eta=0.005, err_tol_force_acc=0.0025, **RUNS[run]))
    e0, ke0, pe0 = sim.energy()
    vcom0 = sim.com_velocity()
    vgrav = math.sqrt(abs(pe0) / 1.0)  # sqrt(|W| / M)
    r0 = nbody.lagrange_radii(sim.positions(), m)
    print(f"[nbody] t=0 E={e0:.6f} KE={ke0:.6f} PE={pe0:.6f} r_lag={r0}", flush=True)
    def audit():
        """audit_tree's invariants on the tree the last walk used, at its age."""
        T = sim.T
        d = T.download()
        r = T.download_reuse()
        hp = HostParticles(*(a.cpu().numpy() for a in (sim.P.x, sim.P.y, sim.P.z, sim.P.m)),
                           soft)
        bad, msg = oracle.audit_tree(d, hp, vel=[a.cpu().numpy() for a in sim.v],
                                     vmax=r["vmax"], tau=sim.t_since_build)
        print(f"[nbody] audit of the tree (age {sim.t_since_build:.3f}): {bad} {msg}",
              flush=True)
        assert bad == 0, msg

    aged = 0
    for k in range(1, 161):
        sim.run_to(0.3 * k)  # whole base steps, so the clock stays aligned with every bin
        if sim.t_since_build > 0 and aged < 3:
            audit()  # a reused tree: drifted sources, kick feed, refreshed walk nodes
            aged += 1
        if k % 40 != 0:
            continue
        e, ke, pe = sim.energy()
        drift = float(np.linalg.norm(sim.com_velocity() - vcom0)) / vgrav
        st = sim.stats
        print(f"[nbody] t={sim.time:.2f} |dE|/KE0={abs(e - e0) / ke0:.4f} drift={drift:.2e} "
              f"syncs={st.syncs} rebuilds={st.rebuilds} forces={st.forces} "
              f"walk_s={st.walk_s:.2f}", flush=True)
    audit()
    if run == "reused":
        assert aged > 0
    rf = nbody.lagrange_radii(sim.positions(), m)
    rel = np.abs(rf - r0) / r0
    print(f"[nbody] Lagrange radii {r0} -> {rf}: |dr/r| {rel}", flush=True)
    assert abs(e - e0) / ke0 < 0.04
    assert np.all(rel < np.array([0.20, 0.15, 0.30]))
    assert drift < 1.0
    if run == "reused":
        assert st.rebuilds < st.syncs / 2
