"""Synthetic particle sets for the tree-gravity workloads (host numpy, seeded).

Every generator returns float64 SoA arrays ``(x, y, z, m, soft)`` with total mass 1 and equal-mass
particles, plus what the workload needs beside them. Softening is the cubic-spline kernel extent
h (shmem/tree.h:52-67), given directly as ``eps``, as shmem/test_tree.cc:25 does.

Radii come from the closed-form inverse cumulative mass of each profile:

* Plummer   M(<r) = r^3 / (r^2 + a^2)^(3/2)   ->  r = a (u^(-2/3) - 1)^(-1/2)
* Hernquist M(<r) = r^2 / (r + a)^2           ->  r = a sqrt(u) / (1 - sqrt(u))
* exponential disk M(<R) = 1 - (1 + R/Rd) exp(-R/Rd), inverted by Newton's method

with u uniform on (0, M(<r_max)) so the profile is truncated at r_max.
"""
from __future__ import annotations

import math

import numpy as np

SEEDS = {
    "plummer_64k": 20260101,
    "hernquist_4m": 20260102,
    "clustered_16m": 20260103,
    "halo_disk_8m": 20260104,
    "spheres_512": 20260105,
    "probes": 20260106,
    "bins": 20260107,
}


def _isotropic(rng, r):
    ct = rng.uniform(-1.0, 1.0, r.shape[0])
    st = np.sqrt(np.maximum(0.0, 1.0 - ct * ct))
    ph = rng.uniform(0.0, 2.0 * math.pi, r.shape[0])
    return r * st * np.cos(ph), r * st * np.sin(ph), r * ct


def plummer_radii(rng, n, a=1.0, rmax_over_a=20.0):
    q = rmax_over_a
    umax = q ** 3 / (1.0 + q * q) ** 1.5
    u = rng.uniform(0.0, umax, n)
    u = np.maximum(u, 1e-300)
    return a / np.sqrt(u ** (-2.0 / 3.0) - 1.0)


def hernquist_radii(rng, n, a=1.0, rmax_over_a=50.0):
    q = rmax_over_a
    umax = (q / (q + 1.0)) ** 2
    s = np.sqrt(rng.uniform(0.0, umax, n))
    return a * s / (1.0 - s)


def exp_disk_radii(rng, n, rd=0.1, rmax_over_rd=30.0):
    q = rmax_over_rd
    umax = 1.0 - (1.0 + q) * math.exp(-q)
    u = rng.uniform(0.0, umax, n)
    x = np.full(n, 1.7)  # R / Rd, start near the median
    for _ in range(60):
        f = 1.0 - (1.0 + x) * np.exp(-x) - u
        fp = x * np.exp(-x)
        step = f / np.maximum(fp, 1e-300)
        x = np.clip(x - step, 1e-12, q)
        if np.max(np.abs(step)) < 1e-13:
            break
    return rd * x


def _pack(x, y, z, eps):
    n = x.shape[0]
    m = np.full(n, 1.0 / n)
    soft = np.full(n, float(eps))
    return (np.ascontiguousarray(x, dtype=np.float64), np.ascontiguousarray(y, dtype=np.float64),
            np.ascontiguousarray(z, dtype=np.float64), m, soft)


def plummer(n=65536, a=1.0, rmax_over_a=20.0, eps=0.01, seed=SEEDS["plummer_64k"]):
    """Plummer sphere, M=1, truncated at rmax_over_a * a (the 64k CPU-reference workload)."""
    rng = np.random.default_rng(seed)
    x, y, z = _isotropic(rng, plummer_radii(rng, n, a, rmax_over_a))
    return _pack(x, y, z, eps)


def plummer_equilibrium(n=32768, a=1.0, rmax_over_a=130.5, eps=0.28, seed=42):
    """A Plummer sphere in equilibrium, G = M = 1: positions as plummer() (truncated at
    rmax_over_a, 100 half-mass radii by default, as test/plummer/make_plummer_ics.py truncates),
    speeds from the isotropic distribution function. The speed in units of the local escape
    speed sqrt(2 / sqrt(r^2 + a^2)) has density proportional to q^2 (1 - q^2)^(7/2) on [0, 1],
    drawn by rejection under its maximum (about 0.092 at q^2 = 2/9; Aarseth, Henon & Wielen
    1974, A&A 37, 183). Returns (x, y, z, vx, vy, vz, m, soft) with the centre of mass and its
    velocity at the origin. eps is the kernel extent h (0.28 = 2.8 x the Plummer-equivalent 0.1
    of the reference's plummer.params)."""
    rng = np.random.default_rng(seed)
    r = plummer_radii(rng, n, a, rmax_over_a)
    x, y, z = _isotropic(rng, r)
    q = np.empty(n)
    todo = np.arange(n)
    while todo.size:
        qq = rng.uniform(0.0, 1.0, todo.size)
        yy = rng.uniform(0.0, 0.1, todo.size)
        ok = yy < qq * qq * (1.0 - qq * qq) ** 3.5
        q[todo[ok]] = qq[ok]
        todo = todo[~ok]
    vesc = np.sqrt(2.0 / np.sqrt(r * r + a * a))
    vx, vy, vz = _isotropic(rng, q * vesc)
    m = np.full(n, 1.0 / n)
    out = []
    for c in (x, y, z, vx, vy, vz):
        out.append(np.ascontiguousarray(c - np.sum(m * c)))
    return (*out, m, np.full(n, float(eps)))


def hernquist(n=4194304, a=1.0, rmax_over_a=50.0, eps=0.005, seed=SEEDS["hernquist_4m"]):
    """Hernquist sphere, M=1, truncated at rmax_over_a * a (the 4M plain-GPU workload)."""
    rng = np.random.default_rng(seed)
    x, y, z = _isotropic(rng, hernquist_radii(rng, n, a, rmax_over_a))
    return _pack(x, y, z, eps)


def clustered(n=16777216, n_clumps=256, f_uniform=0.3, eps=0.001, seed=SEEDS["clustered_16m"]):
    """Non-periodic unit box: f_uniform of the particles uniform in [0,1)^3, the rest in
    n_clumps Plummer clumps (centres uniform in the box, clump masses log-uniform over 2 dex,
    scale radii uniform in [0.002, 0.02], truncated at 20 scale radii)."""
    rng = np.random.default_rng(seed)
    n_u = int(round(f_uniform * n))
    n_c = n - n_u
    xs, ys, zs = [rng.uniform(0.0, 1.0, n_u)], [rng.uniform(0.0, 1.0, n_u)], [rng.uniform(0.0, 1.0, n_u)]
    cm = 10.0 ** rng.uniform(0.0, 2.0, n_clumps)
    share = cm / cm.sum() * n_c
    counts = np.floor(share).astype(np.int64)
    rest = n_c - int(counts.sum())
    counts[np.argsort(-(share - counts))[:rest]] += 1
    centres = rng.uniform(0.0, 1.0, (n_clumps, 3))
    rs = rng.uniform(0.002, 0.02, n_clumps)
    for k in range(n_clumps):
        if counts[k] == 0:
            continue
        r = plummer_radii(rng, int(counts[k]), rs[k], 20.0)
        x, y, z = _isotropic(rng, r)
        xs.append(x + centres[k, 0]); ys.append(y + centres[k, 1]); zs.append(z + centres[k, 2])
    x, y, z = np.concatenate(xs), np.concatenate(ys), np.concatenate(zs)
    perm = rng.permutation(n)
    return _pack(x[perm], y[perm], z[perm], eps)


def halo_disk(n=8388608, f_halo=0.7, a=1.0, rd=0.1, zd=0.01, eps=0.005,
              seed=SEEDS["halo_disk_8m"]):
    """Hernquist halo (a, truncated at 50a) holding f_halo of the particles plus an exponential
    disk (scale length rd, exponential vertical profile with scale height zd). Equal masses."""
    rng = np.random.default_rng(seed)
    n_h = int(round(f_halo * n))
    n_d = n - n_h
    hx, hy, hz = _isotropic(rng, hernquist_radii(rng, n_h, a, 50.0))
    R = exp_disk_radii(rng, n_d, rd)
    ph = rng.uniform(0.0, 2.0 * math.pi, n_d)
    zz = -zd * np.log(np.maximum(rng.uniform(0.0, 1.0, n_d), 1e-300))
    zz *= np.where(rng.uniform(0.0, 1.0, n_d) < 0.5, -1.0, 1.0)
    x = np.concatenate([hx, R * np.cos(ph)])
    y = np.concatenate([hy, R * np.sin(ph)])
    z = np.concatenate([hz, zz])
    perm = rng.permutation(n)
    return _pack(x[perm], y[perm], z[perm], eps)


def spheres(n_spheres=512, lo_log2=10.0, hi_log2=16.0, eps=0.01, spacing=100.0,
            seed=SEEDS["spheres_512"]):
    """n_spheres independent Plummer spheres (a=1, truncated at 20a), N_i = floor(2^U(lo,hi)),
    each centred in its own cell of a cubic lattice of the given spacing. Returns the
    concatenated particle set and the per-sphere start offsets (length n_spheres + 1); each
    sphere has total mass 1."""
    rng = np.random.default_rng(seed)
    sizes = np.floor(2.0 ** rng.uniform(lo_log2, hi_log2, n_spheres)).astype(np.int64)
    side = int(math.ceil(n_spheres ** (1.0 / 3.0)))
    xs, ys, zs, ms = [], [], [], []
    for k in range(n_spheres):
        i, j, l = k % side, (k // side) % side, k // (side * side)
        x, y, z = _isotropic(rng, plummer_radii(rng, int(sizes[k]), 1.0, 20.0))
        xs.append(x + i * spacing); ys.append(y + j * spacing); zs.append(z + l * spacing)
        ms.append(np.full(int(sizes[k]), 1.0 / sizes[k]))
    starts = np.concatenate([[0], np.cumsum(sizes)])
    x, y, z, m = (np.concatenate(v) for v in (xs, ys, zs, ms))
    return (x, y, z, m, np.full(x.shape[0], float(eps))), starts


def probes(n, n_probe=65536, seed=SEEDS["probes"]):
    """Sorted, distinct random particle indices for the direct-sum error judge."""
    rng = np.random.default_rng(seed)
    if n_probe >= n:
        return np.arange(n, dtype=np.uint32)
    return np.sort(rng.choice(n, n_probe, replace=False)).astype(np.uint32)


def timestep_bins(x, y, z, n_levels=5, seed=SEEDS["bins"]):
    """Power-of-two timestep bins for a 16-substep hierarchy: level b holds particles whose
    timestep is 2^-b of the longest, with 4^-b of the particles at level >= b. Particles are
    ranked by a noisy dynamical time (r^(3/2), log-normal scatter of 0.5 dex) so short timesteps
    concentrate where the density is high, as they do in a real run."""
    rng = np.random.default_rng(seed)
    r = np.sqrt(x * x + y * y + z * z)
    tdyn = np.log10(np.maximum(r, 1e-6)) * 1.5 + rng.normal(0.0, 0.5, r.shape[0])
    order = np.argsort(tdyn)  # shortest first
    n = r.shape[0]
    level = np.zeros(n, dtype=np.int8)
    for b in range(1, n_levels):
        k = int(round(n * 4.0 ** (-b)))
        level[order[:k]] = b
    return level


def active_for_substep(level, s, n_levels=5):
    """Indices active at substep s of a 2^(n_levels-1)-substep cycle: everything at s = 0, and at
    s > 0 the levels b >= n_levels - 1 - tz(s), tz the number of trailing zero bits of s."""
    if s == 0:
        return np.arange(level.shape[0], dtype=np.uint32)
    tz = (s & -s).bit_length() - 1
    need = n_levels - 1 - tz
    return np.nonzero(level >= need)[0].astype(np.uint32)
