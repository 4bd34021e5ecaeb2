"""The MLP oracle (oracle/mlp_oracle.py) pinned by checks that share none of its code paths:
central finite differences for the backward pass, an explicit scalar loop for the forward pass,
torch's bfloat16 conversion for the rounding, and the closed form of Adam's first step."""
from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import mlp_oracle as mo


def test_param_count_matches_layout():
    assert mo.N_PARAMS == 5313
    th = mo.init_params(1)
    assert th.shape == (5313,)
    assert np.array_equal(mo.pack(mo.unpack(th)), th)


def test_forward_matches_scalar_loop():
    rng = np.random.default_rng(0)
    th = mo.init_params(2)
    X = rng.normal(size=(5, 16))
    y, _ = mo.forward(th, X)
    p = mo.unpack(th)
    for b in range(5):
        h1 = []
        for i in range(64):
            z = p["b1"][i] + sum(p["W1"][i, k] * X[b, k] for k in range(16))
            h1.append(0.5 * z * (1 + math.erf(z / math.sqrt(2))))
        out = p["b3"][0]
        for i in range(64):
            z = p["b2"][i] + sum(p["W2"][i, j] * h1[j] for j in range(64))
            out += p["w3"][i] * 0.5 * z * (1 + math.erf(z / math.sqrt(2)))
        assert y[b] == pytest.approx(out, rel=1e-12, abs=1e-12)


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(3)
    th = mo.init_params(4)
    X = rng.normal(size=(7, 16))
    t = rng.normal(size=7)
    _, g = mo.loss_and_grad(th, X, t)
    idx = rng.choice(mo.N_PARAMS, 60, replace=False)
    idx = np.concatenate([idx, [0, 1023, 1024, 5312, 5311, 5248]])  # one of each block
    h = 1e-6
    for i in idx:
        tp, tm = th.copy(), th.copy()
        tp[i] += h
        tm[i] -= h
        fd = (mo.loss_and_grad(tp, X, t)[0] - mo.loss_and_grad(tm, X, t)[0]) / (2 * h)
        assert g[i] == pytest.approx(fd, rel=1e-5, abs=1e-8), i


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(5)
    x = np.concatenate([rng.normal(size=1000).astype(np.float32) * 10.0 ** rng.uniform(-5, 5, 1000),
                        np.array([1.0, -2.5, 3.0e-3, 65504.0], dtype=np.float32)]).astype(np.float32)
    ours = mo.bf16_round(x)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(ours.view(np.uint32), ref.view(np.uint32))


def test_adam_first_step_closed_form():
    th = np.array([1.0, -2.0, 0.5])
    g = np.array([0.3, -4.0, 1e-3])
    new, m, v = mo.adam_step(th, np.zeros(3), np.zeros(3), g, 1, lr=0.01)
    # bias-corrected first step is lr * g / (|g| + eps)
    np.testing.assert_allclose(new, th - 0.01 * g / (np.abs(g) + 1e-8), rtol=1e-12)


def test_pinball_backward_matches_finite_differences():
    rng = np.random.default_rng(8)
    th = mo.init_params(9)
    X = rng.normal(size=(9, 16))
    t = rng.normal(size=9) * 3.0  # labels away from the outputs: no kink inside the stencil
    _, g = mo.loss_and_grad(th, X, t, quantile=0.2)
    h = 1e-6
    for i in rng.choice(mo.N_PARAMS, 40, replace=False):
        tp, tm = th.copy(), th.copy()
        tp[i] += h
        tm[i] -= h
        fd = (mo.loss_and_grad(tp, X, t, quantile=0.2)[0] -
              mo.loss_and_grad(tm, X, t, quantile=0.2)[0]) / (2 * h)
        assert g[i] == pytest.approx(fd, rel=1e-5, abs=1e-9), i


def test_pinball_fits_the_quantile():
    """A constant model trained with the pinball loss converges to the label quantile."""
    rng = np.random.default_rng(2)
    t = rng.normal(size=4000)
    y = 0.0
    for _ in range(4000):
        dy = np.mean(np.where(y - t > 0, 1.0, 0.0) - 0.1)
        y -= 0.05 * dy
    assert y == pytest.approx(np.quantile(t, 0.1), abs=0.03)


def _cube_group():
    """Eight targets on the corners of a cube of side 2 centred at (3, 4, 0), mass 0.5 each,
    |a_prev| = 1, softening 0.1, in a leaf of side 4 at level 5 of a tree whose root has its
    centre of mass at the origin, side 10 and mass 100."""
    c = np.array([[sx, sy, sz] for sx in (-1, 1) for sy in (-1, 1) for sz in (-1, 1)], float)
    c += [3.0, 4.0, 0.0]
    n = 8
    return (np.array([0, n]), c[:, 0], c[:, 1], c[:, 2], np.full(n, 0.5), np.full(n, 0.1),
            np.ones(n), np.array([4.0]), np.array([5]), (0.0, 0.0, 0.0, 10.0, 100.0))


def test_group_features_by_hand():
    f = mo.group_features(*_cube_group())[0].astype(np.float64)
    lg = math.log10
    # extent 2, group mass 4, distance 5 from the root's centre of mass
    want = [lg(8), lg(2 / 4), lg(1 * 4 / 4), 0.0, lg(0.1 / 2), 0.0, 0.0, 5 / 20, lg(2 / 5), 0.0,
            lg(100 / 25), lg(4 / 100), lg((4 / 8) / (100 / 125)), lg(4 / 5), lg(0.1 / 5),
            lg(0.1 / 4)]
    np.testing.assert_allclose(f, want, rtol=0, atol=1e-6)


def test_group_features_are_unit_and_translation_free():
    rng = np.random.default_rng(5)
    n = 40
    gstart = np.array([0, 7, 19, 20, 40])
    x, y, z = rng.normal(size=(3, n))
    m = rng.uniform(0.5, 2.0, n)
    soft = rng.uniform(0.01, 0.05, n)
    amag = rng.uniform(0.1, 3.0, n)
    leaf_len = rng.uniform(0.1, 1.0, 4)
    depth = np.array([3, 7, 9, 4])
    root = (0.1, -0.2, 0.3, 8.0, 50.0)
    f0 = mo.group_features(gstart, x, y, z, m, soft, amag, leaf_len, depth, root)
    # lengths x k, masses x q: accelerations scale as q / k^2
    k, q = 37.0, 1e-3
    f1 = mo.group_features(gstart, k * x, k * y, k * z, q * m, k * soft, q / k ** 2 * amag,
                           k * leaf_len, depth, (k * 0.1, k * -0.2, k * 0.3, k * 8.0, q * 50.0))
    np.testing.assert_allclose(f1, f0, rtol=0, atol=2e-6)
    # a translation of everything
    d = (5.0, -3.0, 11.0)
    f2 = mo.group_features(gstart, x + d[0], y + d[1], z + d[2], m, soft, amag, leaf_len, depth,
                           (0.1 + d[0], -0.2 + d[1], 0.3 + d[2], 8.0, 50.0))
    np.testing.assert_allclose(f2, f0, rtol=0, atol=2e-6)
    # and the group statistics do depend on what they should: doubling one group's masses
    m2 = m.copy()
    m2[7:19] *= 2
    f3 = mo.group_features(gstart, x, y, z, m2, soft, amag, leaf_len, depth, root)
    assert np.all(f3[1, [2, 11, 12]] != f0[1, [2, 11, 12]])
    np.testing.assert_array_equal(np.delete(f3, 1, 0), np.delete(f0, 1, 0))
