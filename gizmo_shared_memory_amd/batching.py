"""The batching workloads: uneven active sets of a hierarchical timestep, and many independent
small clusters, walked through one work queue and replayed from a captured graph.

(a) Substeps. The 8,388,608-particle halo+disk with power-of-two timestep bins (5 levels, 16
    substeps; level b holds 4^-b of the particles, placed where the dynamical time is short). At
    substep s the active set is every level whose step ends at s (the reference's hierarchical
    timebins, shmem/mfm.h:276-306); its targets are put in Morton order (gz_order_targets, the
    reference's mfm.cc:700-718) and walked with the relative criterion (aold from the bootstrap
    walk). Measured three ways:
      * eager: per substep, form the groups (a plan) and walk them, timed alone (latency);
      * graph: the 16 substeps' walks captured once into a hipGraph and replayed;
      * batched: all 16 active sets in one segmented plan, walked by one launch.
(b) Spheres. 512 independent Plummer spheres (N = floor(2^U(10, 16)), a = 1, h = 0.01), built as
    one forest (one tree per sphere) and walked by one graph-captured launch; force errors are
    judged against each sphere's own FP64 direct sum on probe particles.
"""
from __future__ import annotations

import json
import time

import numpy as np
import torch

from . import ics
from . import tree as gz

ALPHA_REF = 0.0025
THETA_BOOT = 0.7


def _sync(ctx):
    ctx.synchronize()
    torch.cuda.synchronize()


def substeps(n=1 << 23, n_levels=5, precision="fp32", group_size=32, model=None, reps=3):
    x, y, z, m, soft = ics.halo_disk(n=n)
    P = gz.Particles(x, y, z, m, soft)
    T = gz.build(P)
    ctx = T.ctx
    tg = T.device_orderbuf()
    boot = gz.accel_grouped(T, P, tg, THETA_BOOT, 1.0, 0)
    amag = torch.empty(P.n, dtype=torch.float64, device=P.device)
    amag[tg.view(torch.int32).long()] = torch.linalg.norm(torch.stack(boot, 1), dim=1)
    level = ics.timestep_bins(x, y, z, n_levels=n_levels)
    nsub = 2 ** (n_levels - 1)
    lists, aolds = [], []
    for s in range(nsub):
        t = gz.order_targets(T, ics.active_for_substep(level, s, n_levels))
        lists.append(t)
        aolds.append(ALPHA_REF * amag[t.view(torch.int32).long()])
    counts = [int(t.shape[0]) for t in lists]
    kw = dict(precision=precision, group_size=group_size, model=model)

    # eager: groups + walk per substep, each timed alone
    eager = np.zeros((reps, nsub))
    plans = []
    for r in range(reps):
        plans = []
        for s in range(nsub):
            _sync(ctx)
            t0 = time.perf_counter()
            p = gz.WalkPlan(T, P, lists[s], **kw)
            p.launch(0.0, 1.0, aolds[s])
            _sync(ctx)
            eager[r, s] = time.perf_counter() - t0
            plans.append(p)
    walk_only = np.zeros((reps, nsub))
    for r in range(reps):
        for s in range(nsub):
            _sync(ctx)
            t0 = time.perf_counter()
            plans[s].launch(0.0, 1.0, aolds[s])
            _sync(ctx)
            walk_only[r, s] = time.perf_counter() - t0

    # graph: the 16 walks captured once, replayed
    g = gz.Graph(ctx)
    g.begin()
    for s in range(nsub):
        plans[s].launch(0.0, 1.0, aolds[s])
    g.end()
    g.launch()
    _sync(ctx)
    t_graph = []
    for _ in range(reps):
        _sync(ctx)
        t0 = time.perf_counter()
        g.launch()
        _sync(ctx)
        t_graph.append(time.perf_counter() - t0)

    # batched: all active sets in one segmented plan and one launch
    seg = np.concatenate([[0], np.cumsum(counts)])
    allt = torch.cat([t.view(torch.int32) for t in lists]).view(torch.uint32)
    alla = torch.cat(aolds)
    _sync(ctx)
    t0 = time.perf_counter()
    bp = gz.WalkPlan(T, P, allt, segments=seg, **kw)
    _sync(ctx)
    t_bplan = time.perf_counter() - t0
    bp.launch(0.0, 1.0, alla)
    _sync(ctx)
    t_batched = []
    for _ in range(reps):
        _sync(ctx)
        t0 = time.perf_counter()
        bp.launch(0.0, 1.0, alla)
        _sync(ctx)
        t_batched.append(time.perf_counter() - t0)
    # the batched results are the per-substep results
    same = all(torch.equal(bp.ax[seg[s]:seg[s + 1]], plans[s].ax) for s in range(nsub))
    total = int(seg[-1])
    return {
        "workload": f"halo+disk N={n}, {nsub} substeps of {n_levels} power-of-two bins",
        "precision": precision, "group_size": group_size, "learned": model is not None,
        "active_per_substep": counts,
        "eager_ms_per_substep": (eager.min(0) * 1e3).round(3).tolist(),
        "walk_only_ms_per_substep": (walk_only.min(0) * 1e3).round(3).tolist(),
        "eager_evals_per_s": total / eager.min(0).sum(),
        "graph_ms_total": min(t_graph) * 1e3,
        "graph_evals_per_s": total / min(t_graph),
        "batched_plan_ms": t_bplan * 1e3,
        "batched_ms": min(t_batched) * 1e3,
        "batched_evals_per_s": total / min(t_batched),
        "batched_equals_per_substep": bool(same),
    }


def substeps_reuse(n=1 << 23, n_levels=5, precision="fp32", group_size=32, model=None, dt=1e-4,
                   sigma_v=0.3, seed=11):
    """The 16 substeps with one tree for all of them, as the reference runs between rebuilds
    (mfm.cc:679-905, 3788-3811): at substep s every particle drifts by v dt, the tree's sorted
    copy is re-read (update_sources), the walk nodes are refreshed from the incremental moments
    (refresh_walk_nodes), the active set is walked (a plan formed once per active list: the leaves
    do not change while the tree is reused), and its velocity changes are fed to the tree
    (kick_nodes). Velocities are Gaussian with sigma_v per axis; the kick adds sigma_v / 10."""
    x, y, z, m, soft = ics.halo_disk(n=n)
    rng = np.random.default_rng(seed)
    v = [torch.as_tensor(rng.normal(scale=sigma_v, size=n), device="cuda:0") for _ in range(3)]
    P = gz.Particles(x, y, z, m, soft)
    T = gz.build(P, vel=v)
    ctx = T.ctx
    tg = T.device_orderbuf()
    boot = gz.accel_grouped(T, P, tg, THETA_BOOT, 1.0, 0)
    amag = torch.empty(P.n, dtype=torch.float64, device=P.device)
    amag[tg.view(torch.int32).long()] = torch.linalg.norm(torch.stack(boot, 1), dim=1)
    level = ics.timestep_bins(x, y, z, n_levels=n_levels)
    nsub = 2 ** (n_levels - 1)
    lists = [gz.order_targets(T, ics.active_for_substep(level, s, n_levels)) for s in range(nsub)]
    aolds = [ALPHA_REF * amag[t.view(torch.int32).long()] for t in lists]
    plans = [gz.WalkPlan(T, P, t, precision=precision, group_size=group_size, model=model)
             for t in lists]
    vlast = [a.clone() for a in v]
    g = torch.Generator(device="cuda:0").manual_seed(seed)
    for a, va in zip((P.x, P.y, P.z), v):  # load the drift's kernels before timing
        a.add_(va, alpha=0.0)
    parts = np.zeros((nsub, 5))
    tau = 0.0
    for s in range(nsub):
        _sync(ctx)
        t0 = time.perf_counter()
        for a, va in zip((P.x, P.y, P.z), v):
            a.add_(va, alpha=dt)
        tau += dt
        _sync(ctx)
        t1 = time.perf_counter()
        T.update_sources(P)
        t2 = time.perf_counter()
        T.refresh_walk_nodes(tau)
        t3 = time.perf_counter()
        plans[s].launch(0.0, 1.0, aolds[s])
        _sync(ctx)
        t4 = time.perf_counter()
        act = lists[s].view(torch.int32).long()
        for va in v:
            va[act] += sigma_v / 10 * torch.randn(act.shape[0], generator=g, device="cuda:0",
                                                  dtype=torch.float64)
        T.kick_nodes(lists[s], v, None, vlast, tau)
        _sync(ctx)
        t5 = time.perf_counter()
        parts[s] = (t1 - t0, t2 - t1, t3 - t2, t4 - t3, t5 - t4)
    total = sum(int(t.shape[0]) for t in lists)
    return {
        "workload": f"halo+disk N={n}, {nsub} substeps on one reused tree (drift dt={dt}, "
                    f"sigma_v={sigma_v})",
        "precision": precision, "group_size": group_size, "learned": model is not None,
        "ms_per_substep": (parts.sum(1) * 1e3).round(3).tolist(),
        "walk_ms_per_substep": (parts[:, 3] * 1e3).round(3).tolist(),
        "mean_ms": {k: float(parts[:, i].mean() * 1e3) for i, k in
                    enumerate(("drift", "update_sources", "refresh", "walk", "kick_feed"))},
        "evals_per_s_incl_reuse": total / parts.sum(),
        "plan_status": plans[0].status(),
    }


def _p(t):
    import ctypes as C
    return C.c_void_p(t.data_ptr())


class Schedule:
    """One cycle of 2^(n_levels-1) substeps of a kick-drift-kick integrator on one reused tree, as
    the reference runs between rebuilds (mfm.cc mfm_step, :3440-3870): at substep s everyone
    drifts by dt_sub (gz_drift on the particle arrays, gz_tree_drift on the tree's sorted copy),
    the walk nodes are refreshed to the tree's age (gz_tree_refresh_walk_nodes), and the active
    set of substep s gets aold from its previous accelerations (gz_aold_from_accel), its walk (a
    plan formed once, with the launch parameters gz_walk_autotune picked for its size), the
    write-back (gz_scatter_accel), the kick with dt from its bin (gz_kick) and the kick feed
    (gz_tree_kick_nodes). Nothing in a cycle synchronises with the host, so a cycle can be
    captured into one hipGraph and replayed."""

    def __init__(self, n=1 << 23, n_levels=5, precision="fp32", model=None, dt_sub=1e-4,
                 sigma_v=0.01, seed=11, tune=True, group_size=32):
        from . import _lib as L
        self.L = L
        x, y, z, m, soft = ics.halo_disk(n=n)
        rng = np.random.default_rng(seed)
        dev = "cuda:0"
        self.v = [torch.as_tensor(rng.normal(scale=sigma_v, size=n), device=dev)
                  for _ in range(3)]
        self.P = P = gz.Particles(x, y, z, m, soft)
        self.T = T = gz.build(P, vel=self.v)
        self.ctx = T.ctx
        self.vlast = [a.clone() for a in self.v]
        tg = T.device_orderbuf()
        boot = gz.accel_grouped(T, P, tg, THETA_BOOT, 1.0, 0)
        self.a_grav = torch.zeros((n, 3), dtype=torch.float64, device=dev)
        gz.scatter_accel(tg, *boot, self.a_grav)
        del boot
        level = ics.timestep_bins(x, y, z, n_levels=n_levels)
        self.nsub = 2 ** (n_levels - 1)
        self.dt_sub = dt_sub
        # a particle of level b steps 2^(n_levels - 1 - b) substeps
        self.dt = torch.as_tensor(dt_sub * 2.0 ** (n_levels - 1 - level.astype(np.float64)),
                                  device=dev)
        self.pending = 0.5 * self.dt.clone()
        self.lists = [gz.order_targets(T, ics.active_for_substep(level, s, n_levels))
                      for s in range(self.nsub)]
        self.aolds = [torch.empty(int(t.shape[0]), dtype=torch.float64, device=dev)
                      for t in self.lists]
        self.out = [gz._outputs(int(t.shape[0]), P.device) for t in self.lists]
        self.choices = []
        self.plans = []
        for s, t in enumerate(self.lists):
            aold = gz.aold_from_accel(T, self.a_grav, t, ALPHA_REF, 1.0)
            if tune:
                ch, rep = gz.autotune(T, P, t, 0.0, 1.0, aold, precision=precision, model=model)
            else:
                ch, rep = {"group_size": group_size, "stack_limit": 0, "waves_per_cu": 0}, None
            self.choices.append(ch)
            self.plans.append(gz.WalkPlan(T, P, t, precision=precision, model=model, **ch))
        self.counts = [int(t.shape[0]) for t in self.lists]

    def enqueue_cycle(self):
        """The cycle's calls on the library stream, with no host synchronisation."""
        L, lib, h = self.L, self.L.load(), self.ctx.handle
        P, T = self.P, self.T
        for s in range(self.nsub):
            L.check(lib.gz_drift(h, P.n, self.dt_sub, *(_p(a) for a in self.v), _p(P.m),
                                 _p(P.x), _p(P.y), _p(P.z)), "gz_drift")
            # the tree's age advances on the device (gz_tree_drift); tau = -1 reads it there, so
            # every replay of a captured cycle refreshes at the right age
            L.check(lib.gz_tree_drift(h, T.handle, self.dt_sub), "gz_tree_drift")
            L.check(lib.gz_tree_refresh_walk_nodes(h, T.handle, -1.0),
                    "gz_tree_refresh_walk_nodes")
            t, aold, (ax, ay, az) = self.lists[s], self.aolds[s], self.out[s]
            nt = int(t.shape[0])
            L.check(lib.gz_aold_from_accel(h, _p(self.a_grav), _p(t), nt, ALPHA_REF, 1.0,
                                           _p(aold)), "gz_aold_from_accel")
            L.check(lib.gz_walk_plan_launch(h, self.plans[s].handle, 0.0, 1.0, _p(aold),
                                            _p(ax), _p(ay), _p(az)), "gz_walk_plan_launch")
            L.check(lib.gz_scatter_accel(h, _p(t), nt, _p(ax), _p(ay), _p(az),
                                         _p(self.a_grav)), "gz_scatter_accel")
            L.check(lib.gz_kick(h, _p(t), nt, _p(self.a_grav), None, _p(self.dt),
                                _p(self.pending), *(_p(a) for a in self.v)), "gz_kick")
            L.check(lib.gz_tree_kick_nodes(h, T.handle, _p(t), nt, *(_p(a) for a in self.v),
                                           _p(P.m), *(_p(a) for a in self.vlast), -1.0),
                    "gz_tree_kick_nodes")

    def state(self):
        return [a.clone() for a in (self.P.x, self.P.y, self.P.z, *self.v, self.a_grav,
                                    self.pending)]


def substeps_schedule(n=1 << 23, n_levels=5, precision="fp32", model=None, cycles=3):
    """Config 4a as an integrator runs it: each cycle's 16 substeps in sequence (drift, tree
    drift, refresh, walk, write-back, kick, kick feed), eager and as one captured hipGraph per
    cycle, alternately (the tree ages by one cycle each time, which widens its nodes a little:
    velocities of 0.01 per axis keep that to a few percent of the softening over the run).
    Evals/s count the active targets walked per cycle over the whole cycle's time."""
    S = Schedule(n=n, n_levels=n_levels, precision=precision, model=model)
    ctx = S.ctx
    total = sum(S.counts)
    S.enqueue_cycle()  # loads every kernel and reserves the scratch the capture reuses
    _sync(ctx)
    g = gz.Graph(ctx)
    g.begin()
    S.enqueue_cycle()  # captured, not run
    g.end()
    t_eager, t_graph = [], []
    for _ in range(cycles):
        _sync(ctx)
        t0 = time.perf_counter()
        S.enqueue_cycle()
        _sync(ctx)
        t_eager.append(time.perf_counter() - t0)
        _sync(ctx)
        t0 = time.perf_counter()
        g.launch()
        _sync(ctx)
        t_graph.append(time.perf_counter() - t0)
    return {
        "workload": f"halo+disk N={n}, {S.nsub} sequential substeps of {n_levels} power-of-two "
                    f"bins on one reused tree, kick-drift-kick",
        "precision": precision, "learned": model is not None,
        "tuned": S.choices, "active_per_substep": S.counts,
        "eager_cycle_ms": min(t_eager) * 1e3, "eager_evals_per_s": total / min(t_eager),
        "graph_cycle_ms": min(t_graph) * 1e3, "graph_evals_per_s": total / min(t_graph),
    }


def spheres(n_spheres=512, precision="fp32", group_size=32, theta=0.7, n_probe_spheres=16,
            reps=3):
    (x, y, z, m, soft), starts = ics.spheres(n_spheres=n_spheres)
    P = gz.Particles(x, y, z, m, soft)
    t0 = time.perf_counter()
    F = gz.build(P, clusters=starts)
    t_build = time.perf_counter() - t0
    ctx = F.ctx
    tg = F.device_orderbuf()
    plan = gz.WalkPlan(F, P, tg, precision=precision, group_size=group_size)
    g = gz.Graph(ctx)
    g.begin()
    plan.launch(theta, 1.0, None)
    g.end()
    g.launch()
    _sync(ctx)
    ts = []
    for _ in range(reps):
        _sync(ctx)
        t0 = time.perf_counter()
        g.launch()
        _sync(ctx)
        ts.append(time.perf_counter() - t0)
    # judge: each probed sphere's own direct sum
    order = tg.view(torch.int32).cpu().numpy()
    pos = np.empty(P.n, dtype=np.int64)
    pos[order] = np.arange(order.shape[0])
    rng = np.random.default_rng(7)
    errs = []
    for c in rng.choice(n_spheres, n_probe_spheres, replace=False):
        lo, hi = int(starts[c]), int(starts[c + 1])
        Pc = gz.Particles(x[lo:hi], y[lo:hi], z[lo:hi], m[lo:hi], soft[lo:hi])
        pr = np.sort(rng.choice(hi - lo, min(512, hi - lo), replace=False)).astype(np.uint32)
        ex = np.stack([v.cpu().numpy() for v in gz.accel_brute(Pc, pr)], 1)
        got = np.stack([v.cpu().numpy()[pos[pr + lo]] for v in (plan.ax, plan.ay, plan.az)], 1)
        errs.append(np.linalg.norm(got - ex, axis=1) / np.linalg.norm(ex, axis=1))
    e = np.concatenate(errs)
    return {
        "workload": f"{n_spheres} Plummer spheres, N={P.n} total, forest of one tree each",
        "precision": precision, "group_size": group_size, "theta": theta,
        "build_s": t_build, "graph_ms": min(ts) * 1e3, "evals_per_s": P.n / min(ts),
        "err": {"median": float(np.median(e)), "p99": float(np.percentile(e, 99)),
                "max": float(e.max())},
        "status": plan.status(),
    }


def main():
    import argparse
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--learned", action="store_true")
    args = ap.parse_args()
    model = None
    if args.learned:
        from . import learned
        model = learned.load_model()
    print(json.dumps(substeps_schedule(model=model)), flush=True)
    print(json.dumps(substeps(model=model)), flush=True)
    print(json.dumps(substeps_reuse(model=model)), flush=True)
    print(json.dumps(spheres()), flush=True)


if __name__ == "__main__":
    main()
