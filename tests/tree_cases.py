"""Small particle sets shared by the oracle and device parity tests, including the edge cases the
reference's own tests and ICs exercise (test/gravtree_collocated: bit-identical positions that
force MAX_LEVEL leaves; massless tombstones left out of the tree, tree.cc:355-359)."""
from __future__ import annotations

import numpy as np

from gizmo_shared_memory_amd import ics


def plummer_small(n=20000, seed=42, eps=1e-3):
    """Plummer a=1 with radii from u in [0, 0.999) as shmem/test_tree.cc:24-35 draws them."""
    rng = np.random.default_rng(seed)
    u = rng.uniform(0.0, 0.999, n)
    u = np.maximum(u, 1e-300)
    r = 1.0 / np.sqrt(u ** (-2.0 / 3.0) - 1.0)
    x, y, z = ics._isotropic(rng, r)
    return x, y, z, np.full(n, 1.0 / n), np.full(n, eps)


def collocated(seed=7, n_clusters=8, per=16, n_random=100, eps=0.01):
    """8 clusters of 16 bit-identical positions plus 100 random particles, after
    test/gravtree_collocated/make_gravtree_collocated_ics.py."""
    rng = np.random.default_rng(seed)
    c = rng.uniform(-1.0, 1.0, (n_clusters, 3))
    pts = np.repeat(c, per, axis=0)
    rnd = rng.uniform(-1.0, 1.0, (n_random, 3))
    allp = np.concatenate([pts, rnd])
    perm = rng.permutation(allp.shape[0])
    allp = allp[perm]
    n = allp.shape[0]
    return (allp[:, 0].copy(), allp[:, 1].copy(), allp[:, 2].copy(), np.full(n, 1.0 / n),
            np.full(n, eps))


def with_tombstones(seed=11, n=3000, frac=0.1, eps=0.02):
    """Uniform cube with a fraction of massless particles (tombstones)."""
    rng = np.random.default_rng(seed)
    x, y, z = (rng.uniform(0.0, 1.0, n) for _ in range(3))
    m = np.full(n, 1.0 / n)
    m[rng.uniform(0.0, 1.0, n) < frac] = 0.0
    soft = rng.uniform(0.5 * eps, eps, n)
    return x, y, z, m, soft


def close_pair_chain(seed=3, n=200, eps=1e-3):
    """A uniform cloud plus two particles 1e-9 apart: a single-child chain down to MAX_LEVEL."""
    rng = np.random.default_rng(seed)
    x, y, z = (rng.uniform(-1.0, 1.0, n) for _ in range(3))
    x[1], y[1], z[1] = x[0] + 1e-9, y[0], z[0]
    return x, y, z, np.full(n, 1.0 / n), np.full(n, eps)


def dense_cloud_in_box(seed=5, n_cloud=20000, n_ambient=50, eps=1e-4):
    """A cloud filling ~1/1000 of its bounding box (the MakeCloud case that outgrows the node
    arena's first estimate, tree.cc:428-437), kept in place by a few ambient particles."""
    rng = np.random.default_rng(seed)
    r = 0.05 * rng.uniform(0.0, 1.0, n_cloud) ** (1.0 / 3.0)
    x, y, z = ics._isotropic(rng, r)
    ax, ay, az = (rng.uniform(-1.0, 1.0, n_ambient) for _ in range(3))
    x, y, z = np.concatenate([x, ax]), np.concatenate([y, ay]), np.concatenate([z, az])
    n = x.shape[0]
    return x, y, z, np.full(n, 1.0 / n), np.full(n, eps)


def tiny(n, seed=1, eps=0.05):
    rng = np.random.default_rng(seed)
    x, y, z = (rng.uniform(-1.0, 1.0, n) for _ in range(3))
    return x, y, z, np.full(n, 1.0 / max(n, 1)), np.full(n, eps)


CASES = {
    "plummer20k": lambda: plummer_small(),
    "plummer_relaxed_eps": lambda: plummer_small(6000, seed=9, eps=0.05),
    "collocated": collocated,
    "tombstones": with_tombstones,
    "close_pair": close_pair_chain,
    "n1": lambda: tiny(1),
    "n2": lambda: tiny(2),
    "n16": lambda: tiny(16),
    "n17": lambda: tiny(17),
    "n300": lambda: tiny(300),
}
