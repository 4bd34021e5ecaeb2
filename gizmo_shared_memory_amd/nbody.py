"""A gravity-only N-body integrator on the device, built from the library's calls: the
kick-drift-kick leapfrog with hierarchical power-of-two timesteps that the reference engine runs
(shmem/mfm.cc mfm_step, :3440-3870, gravity part), with the tree reused between rebuilds.

One step, as mfm_step orders it:

1. the particles due at the current clock are kicked: the half kick owed from their previous step
   closed and the next one opened, with dt from the forces at the current positions
   (assign_bins :3130-3160, the kick :3699-3767, gz_kick);
2. the kick is fed into the reused tree's speed bounds and incremental moments (:3793-3811,
   gz_tree_kick_nodes); a kick of more than TreeDomainUpdateFrequency of the particles marks the
   tree for a rebuild instead;
3. everyone drifts to the next sync point, the end of the deepest active bin's step (:3823-3853,
   gz_drift);
4. forces for the particles due at the new clock: the tree is rebuilt when the forces walked since
   the last build exceed TreeDomainUpdateFrequency of the particles or when its speed bound times
   its age passes a quarter of the softening (:3326-3344); otherwise the sorted copy is
   re-read (gz_tree_update_sources) and the walk nodes refreshed to the tree's age
   (gz_tree_refresh_walk_nodes). Targets in Morton order (gz_order_targets), aold from the
   previous accelerations (gz_aold_from_accel), the grouped walk with the relative criterion, and
   the write-back into a_grav (gz_scatter_accel), as compute_gravity does (:679-905).

The run starts with a bootstrap walk at theta = 0.7 (no previous accelerations), as the reference
initialises before its main loop (:3444-3452). Time is an integer clock of dt_max / 2^30 ticks
(mfm.h:276-314), so every bin's sync points are exact.

Energy uses the potential walk (gz_tree_potential) on a fresh tree, with velocities at a sync point
of every particle: the pending half kick closed (the state between the closing and the opening
kick is the synchronised one, mfm.cc:3704-3708).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
import torch

from . import tree as gz

MAX_BINS = 30
TICKS = 1 << MAX_BINS  # ticks per dt_max


@dataclass
class Params:
    """Run parameters, named after the params-file entries they stand for."""

    G: float = 1.0
    dt_max: float = 0.1                 # MaxSizeTimestep
    eta: float = 0.005                  # ErrTolIntAccuracy
    err_tol_force_acc: float = 0.0025   # ErrTolForceAcc
    theta_boot: float = 0.7             # the bootstrap walk's opening angle
    tree_update_freq: float = 0.1       # TreeDomainUpdateFrequency
    tree_pad_frac: float = 0.25         # rebuild once vmax * age exceeds this many softenings
    precision: str = "fp64"             # the walk: "fp64" (parity) or "fp32"
    group_size: int = 0                 # 0: the reference's leaf groups


@dataclass
class Stats:
    syncs: int = 0
    rebuilds: int = 0
    forces: int = 0
    walk_s: float = 0.0
    bins: dict = field(default_factory=dict)


class GravitySim:
    """Particles, velocities and timestep state resident on one device."""

    def __init__(self, x, y, z, vx, vy, vz, m, soft, params: Params | None = None,
                 device="cuda:0"):
        self.p = params or Params()
        self.P = gz.Particles(x, y, z, m, soft, device=device)
        dev = self.P.device
        f64 = dict(dtype=torch.float64, device=dev)
        self.v = [torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device=dev)
                  for a in (vx, vy, vz)]
        self.n = self.P.n
        self.a_grav = torch.zeros((self.n, 3), **f64)
        self.pending = torch.zeros(self.n, **f64)
        self.dt = torch.zeros(self.n, **f64)
        self.bin = torch.zeros(self.n, dtype=torch.int64, device=dev)
        self.vlast = [a.clone() for a in self.v]
        self.clock = 0
        self.T = None
        self.tree_valid = False
        self.t_since_build = 0.0
        self.forces_since_build = 0
        self.stats = Stats()
        self._all = torch.arange(self.n, dtype=torch.int32, device=dev).view(torch.uint32)
        self._forces(self._all, bootstrap=True)

    # -- time ---------------------------------------------------------------------------------
    @property
    def time(self) -> float:
        return self.p.dt_max * self.clock / TICKS

    @staticmethod
    def _ticks_in_bin(b):
        return 1 << (MAX_BINS - b)

    def _due(self) -> torch.Tensor:
        """Particles whose bin is aligned with the clock (mfm.cc:3252-3263)."""
        ticks = torch.bitwise_left_shift(torch.ones_like(self.bin), MAX_BINS - self.bin)
        due = torch.bitwise_and(torch.full_like(self.bin, self.clock), ticks - 1) == 0
        due &= self.P.m > 0
        return torch.nonzero(due).flatten().to(torch.int32).view(torch.uint32)

    def _assign_bins(self, active: torch.Tensor):
        """assign_bins (mfm.cc:3130-3160): dt = min(dt_max, sqrt(eta soft / |a|)) (desired_dt,
        :2974-2980) to the bin of the largest power-of-two step not above it; deepening is always
        allowed, coarsening one bin at a time while the coarser step is aligned with the clock."""
        ai = active.view(torch.int32).long()
        a = torch.linalg.norm(self.a_grav[ai], dim=1)
        soft = self.P.soft[ai] if self.P.soft is not None else torch.zeros_like(a)
        want_dt = torch.where(a > 0, torch.sqrt(self.p.eta * soft / torch.where(a > 0, a, 1.0)),
                              torch.full_like(a, math.inf))
        want_dt = torch.clamp(want_dt, max=self.p.dt_max)
        # bin_for_dt: the shallowest b with dt_max / 2^b <= want_dt
        ratio = torch.clamp(self.p.dt_max / want_dt, min=1.0)
        want = torch.clamp(torch.ceil(torch.log2(ratio) - 1e-12).long(), 0, MAX_BINS)
        cur = self.bin[ai]
        b = torch.where(want >= cur, want, cur)
        coarsen = want < cur
        if bool(coarsen.any()):
            bb = cur.clone()
            for _ in range(MAX_BINS):
                can = coarsen & (bb > want) & (bb > 0)
                if not bool(can.any()):
                    break
                step = torch.bitwise_left_shift(torch.ones_like(bb), MAX_BINS - (bb - 1))
                aligned = torch.bitwise_and(torch.full_like(bb, self.clock), step - 1) == 0
                go = can & aligned
                if not bool(go.any()):
                    break
                bb = torch.where(go, bb - 1, bb)
                coarsen = coarsen & go
            b = torch.where(want < cur, bb, b)
        self.bin[ai] = b
        self.dt[ai] = self.p.dt_max / torch.pow(2.0, b.to(torch.float64))

    # -- forces -------------------------------------------------------------------------------
    def _rebuild(self):
        self.T = gz.build(self.P, vel=self.v)
        for a, b in zip(self.vlast, self.v):
            a.copy_(b)
        self.t_since_build = 0.0
        self.forces_since_build = 0
        self.tree_valid = True
        self.stats.rebuilds += 1

    def _tree_for_walk(self, n_active: int):
        """evaluate_forces' tree choice (mfm.cc:3326-3344)."""
        typical_h = float(self.P.soft[0]) if self.P.soft is not None else 0.0
        rebuild = not self.tree_valid or self.T is None or typical_h <= 0.0
        if not rebuild:
            if self.forces_since_build + n_active > self.p.tree_update_freq * self.n:
                rebuild = True
            else:
                vmax_root = float(self.T.download_reuse()["vmax"][self.T.root])
                rebuild = vmax_root * self.t_since_build > self.p.tree_pad_frac * typical_h
        if rebuild:
            self._rebuild()
        else:
            self.T.update_sources()
            self.T.refresh_walk_nodes(self.t_since_build)
        self.forces_since_build += n_active

    def _forces(self, active: torch.Tensor, bootstrap: bool = False):
        if int(active.shape[0]) == 0:
            return
        self._tree_for_walk(int(active.shape[0]))
        T = self.T
        tg = gz.order_targets(T, None if int(active.shape[0]) == self.n else active)
        if bootstrap:
            theta, aold = self.p.theta_boot, None
        else:
            theta = 0.0
            aold = gz.aold_from_accel(T, self.a_grav, tg, self.p.err_tol_force_acc, self.p.G)
        import time
        t0 = time.perf_counter()
        ax, ay, az = gz.accel_grouped(T, self.P, tg, theta, self.p.G, 0, aold,
                                      precision=self.p.precision, group_size=self.p.group_size)
        gz.scatter_accel(tg, ax, ay, az, self.a_grav)
        self.stats.walk_s += time.perf_counter() - t0
        self.stats.forces += int(tg.shape[0])

    # -- one sync point -----------------------------------------------------------------------
    def step(self, t_end_ticks: int):
        active = self._due()
        na = int(active.shape[0])
        if na == 0:
            raise RuntimeError(f"empty active set at clock {self.clock}: the timestep hierarchy "
                               f"has desynchronised")
        self._assign_bins(active)
        gz.kick(active, self.a_grav, self.dt, self.pending, self.v)
        if na > self.p.tree_update_freq * self.n:
            self.tree_valid = False  # a big step rebuilds anyway (mfm.cc:3793-3796)
        if self.tree_valid:
            self.T.kick_nodes(active, self.v, vlast=self.vlast, tau=self.t_since_build)
        deepest = int(self.bin[active.view(torch.int32).long()].max())
        step_ticks = min(self._ticks_in_bin(deepest), t_end_ticks - self.clock)
        dt = self.p.dt_max * step_ticks / TICKS
        gz.drift(self.P, self.v, dt)
        self.t_since_build += dt
        self.clock += step_ticks
        self._forces(self._due())
        self.stats.syncs += 1

    def run_to(self, t_end: float):
        end = int(round(t_end / self.p.dt_max * TICKS))
        while self.clock < end:
            self.step(end)

    # -- diagnostics --------------------------------------------------------------------------
    def synchronised_velocities(self):
        """v with each particle's pending half kick closed at its current acceleration."""
        return [self.v[k] + self.a_grav[:, k] * self.pending for k in range(3)]

    def energy(self, theta: float = 0.5):
        """(E, KE, PE) at the current positions with synchronised velocities; the potential from
        the per-target potential walk on a fresh tree (tree.cc:899-944)."""
        v = self.synchronised_velocities()
        m = self.P.m
        ke = 0.5 * float(torch.sum(m * (v[0] ** 2 + v[1] ** 2 + v[2] ** 2)))
        T = gz.build(self.P)
        phi = gz.potential(T, self.P, T.device_orderbuf(), theta, self.p.G)
        order = T.device_orderbuf().view(torch.int32).long()
        pe = 0.5 * float(torch.sum(m[order] * phi))
        return ke + pe, ke, pe

    def com_velocity(self):
        v = self.synchronised_velocities()
        m = self.P.m
        M = float(torch.sum(m))
        return np.array([float(torch.sum(m * a)) / M for a in v])

    def positions(self):
        return np.stack([self.P.x.cpu().numpy(), self.P.y.cpu().numpy(),
                         self.P.z.cpu().numpy()], 1)


def lagrange_radii(pos, m, fractions=(0.1, 0.5, 0.9)):
    r = np.sqrt(np.sum(pos ** 2, axis=1))
    order = np.argsort(r)
    cum = np.cumsum(m[order]) / m.sum()
    rs = r[order]
    return np.array([rs[np.searchsorted(cum, f)] for f in fractions])
