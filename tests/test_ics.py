"""The synthetic initial conditions against their closed forms (CPU)."""
from __future__ import annotations

import math

import numpy as np

from gizmo_shared_memory_amd import ics


def test_plummer_equilibrium_is_virialised():
    x, y, z, vx, vy, vz, m, soft = ics.plummer_equilibrium(n=1 << 15, seed=3)
    ke = 0.5 * np.sum(m * (vx ** 2 + vy ** 2 + vz ** 2))
    # an untruncated Plummer sphere (G = M = a = 1) has KE = 3 pi / 64 = -W / 2
    assert abs(ke - 3 * math.pi / 64) < 0.01 * 3 * math.pi / 64
    r = np.sqrt(x * x + y * y + z * z)
    # half-mass radius (2^(2/3) - 1)^(-1/2) = 1.305
    assert abs(np.median(r) - 1.305) < 0.03
    # speeds never reach the local escape speed
    vesc = np.sqrt(2.0 / np.sqrt(r * r + 1.0))
    assert np.all(np.sqrt(vx ** 2 + vy ** 2 + vz ** 2) < vesc * 1.0000001 + 1e-9)
    # centre of mass and its velocity at the origin
    for c in (x, y, z, vx, vy, vz):
        assert abs(np.sum(m * c)) < 1e-12


def test_plummer_speed_distribution_moments():
    """The mean of q^2 under the density q^2 (1 - q^2)^(7/2) on [0, 1] is
    B(5/2, 9/2) / B(3/2, 9/2) = 1/4 (int q^(2k+2) (1 - q^2)^(7/2) dq = B(k + 3/2, 9/2) / 2)."""
    from scipy.special import beta
    x, y, z, vx, vy, vz, m, soft = ics.plummer_equilibrium(n=200000, seed=5)
    r = np.sqrt(x * x + y * y + z * z)
    q = np.sqrt(vx ** 2 + vy ** 2 + vz ** 2) / np.sqrt(2.0 / np.sqrt(r * r + 1.0))
    want = beta(2.5, 4.5) / beta(1.5, 4.5)
    assert abs(np.mean(q * q) - want) < 0.005
