"""CPU oracle for the learned opening-parameter network (test infrastructure only).

A float64 numpy restatement of the device MLP (gizmo_shared_memory_amd/csrc/gz_learned.hip):

    16 features -> standardise -> 64 (GELU) -> 64 (GELU) -> 1

with the exact GELU, x * (1 + erf(x / sqrt 2)) / 2, and the mean-squared-error loss
L = sum_b (y_b - t_b)^2 / B. The device keeps FP32 master weights, computes in FP32 from bf16
copies of the weight matrices (biases stay FP32), and accumulates in FP32; this oracle takes the
same rounded parameters (bf16_weights) and works in float64, so the two agree to FP32 rounding.

Parameter layout (5,313 floats, the device's): W1[64][16], b1[64], W2[64][64], b2[64], w3[64],
b3[1].

Pinning: the forward pass against an independent scalar loop (tests), the backward pass against
central finite differences of this forward pass, and the bf16 rounding against numpy's float32
bit pattern truncated with round-to-nearest-even.

group_features restates the 16 features of a target group (gz_features.h: group_features16 over
the statistics k_group_stats gathers). They are pinned by a group whose features are worked out
by hand, and by their invariance under a change of units and a translation of the system.
"""
from __future__ import annotations

import math

import numpy as np

N_IN, N_H = 16, 64
SIZES = {"W1": (N_H, N_IN), "b1": (N_H,), "W2": (N_H, N_H), "b2": (N_H,), "w3": (N_H,),
         "b3": (1,)}
N_PARAMS = sum(int(np.prod(s)) for s in SIZES.values())  # 5313

_erf = np.vectorize(math.erf)


def unpack(theta: np.ndarray) -> dict:
    out, o = {}, 0
    for k, s in SIZES.items():
        n = int(np.prod(s))
        out[k] = np.asarray(theta[o:o + n], dtype=np.float64).reshape(s)
        o += n
    return out


def pack(p: dict) -> np.ndarray:
    return np.concatenate([np.asarray(p[k], dtype=np.float64).ravel() for k in SIZES])


def bf16_round(x: np.ndarray) -> np.ndarray:
    """float32 -> bf16 -> float32 with round-to-nearest-even (NaN-free inputs)."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def gelu(x):
    return 0.5 * x * (1.0 + _erf(x / math.sqrt(2.0)))


def gelu_grad(x):
    return 0.5 * (1.0 + _erf(x / math.sqrt(2.0))) + x * np.exp(-0.5 * x * x) / math.sqrt(2 * math.pi)


def bf16_weights(theta):
    """The device's working copy: W1, W2 and w3 rounded to bf16, biases kept in FP32."""
    th = np.asarray(theta, dtype=np.float32).astype(np.float64)
    p = unpack(th)
    for k in ("W1", "W2", "w3"):
        p[k] = bf16_round(p[k].astype(np.float32)).astype(np.float64)
    return pack(p)


def forward(theta, X, mean=None, scale=None):
    """The network on inputs standardised as (X - mean) * scale (the device applies the same
    FP32 affine map before the first layer)."""
    p = unpack(theta)
    X = np.asarray(X, dtype=np.float64)
    if mean is not None:
        X = (X - np.asarray(mean, dtype=np.float64)) * np.asarray(scale, dtype=np.float64)
    z1 = X @ p["W1"].T + p["b1"]
    h1 = gelu(z1)
    z2 = h1 @ p["W2"].T + p["b2"]
    h2 = gelu(z2)
    y = h2 @ p["w3"] + p["b3"][0]
    return y, (X, z1, h1, z2, h2)


def loss_and_grad(theta, X, t, mean=None, scale=None, quantile=0.0):
    """Loss and its gradient with respect to every parameter (backpropagation). quantile 0 is
    the mean squared error; 0 < quantile < 1 the pinball loss of that quantile."""
    p = unpack(theta)
    y, (X, z1, h1, z2, h2) = forward(theta, X, mean, scale)
    B = X.shape[0]
    r = y - np.asarray(t, dtype=np.float64)
    if quantile > 0:
        loss = float(np.sum(np.where(r > 0, (1 - quantile) * r, -quantile * r)) / B)
        dy = (np.where(r > 0, 1.0, 0.0) - quantile) / B
    else:
        loss = float(np.sum(r * r) / B)
        dy = 2.0 * r / B
    g = {"w3": h2.T @ dy, "b3": np.array([dy.sum()])}
    dz2 = np.outer(dy, p["w3"]) * gelu_grad(z2)
    g["W2"] = dz2.T @ h1
    g["b2"] = dz2.sum(0)
    dh1 = dz2 @ p["W2"]
    dz1 = dh1 * gelu_grad(z1)
    g["W1"] = dz1.T @ X
    g["b1"] = dz1.sum(0)
    return loss, pack(g)


def adam_step(theta, m, v, grad, t, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam update (bias-corrected), returning new (theta, m, v)."""
    m = beta1 * m + (1 - beta1) * grad
    v = beta2 * v + (1 - beta2) * grad * grad
    mh = m / (1 - beta1 ** t)
    vh = v / (1 - beta2 ** t)
    return theta - lr * mh / (np.sqrt(vh) + eps), m, v


def init_params(seed=0):
    """He-style uniform initialisation; the device is seeded from these values."""
    rng = np.random.default_rng(seed)
    p = {"W1": rng.uniform(-1, 1, (N_H, N_IN)) * math.sqrt(3.0 / N_IN),
         "b1": np.zeros(N_H),
         "W2": rng.uniform(-1, 1, (N_H, N_H)) * math.sqrt(3.0 / N_H),
         "b2": np.zeros(N_H),
         "w3": rng.uniform(-1, 1, N_H) * math.sqrt(3.0 / N_H),
         "b3": np.zeros(1)}
    return pack(p)


def group_features(gstart, x, y, z, m, soft, amag, leaf_len, depth, root):
    """The 16 features of every group (float32, [ngroups, 16]).

    gstart: group starts in the target list, ngroups + 1 entries; x, y, z, m, soft, amag: per
    target in list order; leaf_len, depth: side and level of the leaf of each group's first
    target; root: (cx, cy, cz, side, mass) of the tree's root (centre of mass, side, mass)."""
    gstart = np.asarray(gstart, dtype=np.int64)
    ng = gstart.shape[0] - 1
    rcx, rcy, rcz, rside, rmass = (float(v) for v in root)
    side = max(rside, 1e-300)
    Mr = max(rmass, 1e-300)
    out = np.zeros((ng, N_IN), dtype=np.float32)

    def lg(v):
        return math.log10(max(v, 1e-300))

    for g in range(ng):
        s = slice(gstart[g], gstart[g + 1])
        nt = gstart[g + 1] - gstart[g]
        lo = [float(np.min(a[s])) for a in (x, y, z)]
        hi = [float(np.max(a[s])) for a in (x, y, z)]
        emax = max(float(np.max(soft[s])), 0.0)
        msum = float(np.sum(m[s]))
        amin, amax = float(np.min(amag[s])), max(float(np.max(amag[s])), 0.0)
        a2 = float(np.sum(amag[s] ** 2))
        ex, ey, ez = hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]
        ext = max(ex, ey, ez, 1e-12 * side)
        emn = max(min(ex, ey, ez), 1e-12 * side)
        cx = 0.5 * (lo[0] + hi[0]) - rcx
        cy = 0.5 * (lo[1] + hi[1]) - rcy
        cz = 0.5 * (lo[2] + hi[2]) - rcz
        dist = max(math.sqrt(cx * cx + cy * cy + cz * cz), 1e-12 * side)
        arms = math.sqrt(a2 / max(nt, 1))
        mg = max(msum, 1e-300)
        rho_g, rho_c = mg / ext ** 3, Mr / dist ** 3
        amn = max(amin, 1e-300)
        ll = float(leaf_len[g])
        out[g] = [lg(nt), lg(ext / ll), lg(amn * ext * ext / mg), lg(amax / amn), lg(emax / ext),
                  lg(emn / ext), lg(arms / amn), float(depth[g]) / 20.0, lg(ext / dist),
                  abs(cz) / dist, lg(Mr / (dist * dist * amn)), lg(mg / Mr), lg(rho_g / rho_c),
                  lg(ll / dist), lg(emax / dist), lg(emax / ll)]
    return out
