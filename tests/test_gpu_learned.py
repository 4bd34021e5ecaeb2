"""Device parity of the learned opening criterion: the network's forward pass, gradients and Adam
step against the float64 oracle (oracle/mlp_oracle.py) on the same bf16-rounded weights, and the
learned walk's fallback to the reference criterion."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import mlp_oracle as mo
from tests import tree_cases

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def gz():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from gizmo_shared_memory_amd import tree
    return tree


def _np(t):
    return t.detach().cpu().numpy()


def _bf(theta):
    return mo.bf16_weights(theta)


def test_forward_matches_oracle(gz):
    rng = np.random.default_rng(1)
    th = mo.init_params(7).astype(np.float32)
    th[1024:1088] = rng.normal(size=64)  # non-zero FP32 biases
    M = gz.Model(th)
    X = rng.normal(size=(1000, 16)).astype(np.float32)
    y = M.forward(torch.from_numpy(X).cuda()).cpu().numpy()
    yo, _ = mo.forward(_bf(th), X.astype(np.float64))
    np.testing.assert_allclose(y, yo, rtol=1e-4, atol=1e-4)
    # with an input map, as the trained model uses it
    mean = rng.normal(size=16).astype(np.float32)
    scale = rng.uniform(0.5, 3.0, 16).astype(np.float32)
    M.set_input_map(mean, scale)
    y = M.forward(torch.from_numpy(X).cuda()).cpu().numpy()
    yo, _ = mo.forward(_bf(th), X.astype(np.float64), mean, scale)
    np.testing.assert_allclose(y, yo, rtol=1e-4, atol=1e-4)


@pytest.mark.parametrize("B", [1, 63, 64, 200])
def test_gradient_matches_oracle(gz, B):
    rng = np.random.default_rng(B)
    th = mo.init_params(3).astype(np.float32)
    M = gz.Model(th)
    mean = rng.normal(size=16).astype(np.float32)
    scale = rng.uniform(0.5, 2.0, 16).astype(np.float32)
    M.set_input_map(mean, scale)
    X = rng.normal(size=(B, 16)).astype(np.float32)
    T = rng.normal(size=B).astype(np.float32)
    loss, g = M.grad(torch.from_numpy(X).cuda(), torch.from_numpy(T).cuda())
    lo, go = mo.loss_and_grad(_bf(th), X.astype(np.float64), T.astype(np.float64), mean, scale)
    g = g.cpu().numpy().astype(np.float64)
    assert loss == pytest.approx(lo, rel=1e-4)
    assert np.linalg.norm(g - go) <= 1e-4 * np.linalg.norm(go) + 1e-7
    # every parameter block is exercised
    for k, v in mo.unpack(go).items():
        gk = mo.unpack(g)[k]
        assert np.linalg.norm(gk - v) <= 2e-4 * (np.linalg.norm(v) + 1e-6), k


def test_pinball_gradient_matches_oracle(gz):
    rng = np.random.default_rng(11)
    th = mo.init_params(12).astype(np.float32)
    M = gz.Model(th)
    M.set_loss(0.1)
    X = rng.normal(size=(300, 16)).astype(np.float32)
    T = (rng.normal(size=300) * 2.0).astype(np.float32)
    loss, g = M.grad(torch.from_numpy(X).cuda(), torch.from_numpy(T).cuda())
    lo, go = mo.loss_and_grad(_bf(th), X.astype(np.float64), T.astype(np.float64), quantile=0.1)
    g = g.cpu().numpy().astype(np.float64)
    assert loss == pytest.approx(lo, rel=1e-4)
    assert np.linalg.norm(g - go) <= 1e-4 * np.linalg.norm(go) + 1e-7


def test_adam_step_matches_oracle(gz):
    rng = np.random.default_rng(4)
    th = mo.init_params(5).astype(np.float32)
    M = gz.Model(th)
    X = torch.from_numpy(rng.normal(size=(128, 16)).astype(np.float32)).cuda()
    T = torch.from_numpy(rng.normal(size=128).astype(np.float32)).cuda()
    _, g = M.grad(X, T)
    M.train_step(X, T, 1e-3)
    got = M.params().astype(np.float64)
    want, _, _ = mo.adam_step(th.astype(np.float64), np.zeros(mo.N_PARAMS),
                              np.zeros(mo.N_PARAMS), g.cpu().numpy().astype(np.float64), 1, 1e-3)
    np.testing.assert_allclose(got, want, rtol=1e-5, atol=1e-6)


def test_training_reduces_loss(gz):
    rng = np.random.default_rng(6)
    M = gz.Model(mo.init_params(6).astype(np.float32))
    X = rng.normal(size=(4096, 16)).astype(np.float32)
    T = (np.sin(X[:, 0]) + 0.5 * X[:, 1] * X[:, 2] - 0.3 * X[:, 3]).astype(np.float32)
    Xd, Td = torch.from_numpy(X).cuda(), torch.from_numpy(T).cuda()
    first = M.train_step(Xd, Td, 3e-3)
    for _ in range(300):
        last = M.train_step(Xd, Td, 3e-3)
    assert last < 0.2 * first, (first, last)


def _walk_setup(gz):
    case = tree_cases.plummer_small(20000, seed=17, eps=0.01)
    P = gz.Particles(*case)
    T = gz.build(P)
    tg = T.device_orderbuf()
    boot = gz.accel_grouped(T, P, tg, 0.7, 1.0, 0)
    aold = 0.0025 * torch.linalg.norm(torch.stack(boot, 1), dim=1)
    return P, T, tg, aold


def test_group_features_are_finite_and_match_walk_groups(gz):
    P, T, tg, aold = _walk_setup(gz)
    feats, gstart = gz.group_features(T, P, tg, aold, group_size=32)
    c = gz.accel_grouped(T, P, tg, 0.0, 1.0, 0, aold=aold, group_size=32, precision="fp32",
                         counts=True)[3]
    assert feats.shape == (c["groups"], 16)
    assert torch.isfinite(feats).all()
    gs = gstart.cpu().numpy()
    assert gs[0] == 0 and gs[-1] == tg.shape[0] and np.all(np.diff(gs) > 0)


@pytest.mark.parametrize("group_size", [0, 16, 32, 64])
def test_group_features_match_oracle(gz, group_size):
    """The device features of every walk group against the oracle's restatement, from the same
    groups, targets and tree (features are logarithms in FP32: 1e-5 absolute)."""
    P, T, tg, aold = _walk_setup(gz)
    feats, gstart = gz.group_features(T, P, tg, aold, group_size=group_size)
    feats, gstart = feats.cpu().numpy(), gstart.cpu().numpy()
    d = T.download()
    x, y, z, m, soft = tree_cases.plummer_small(20000, seed=17, eps=0.01)
    t = tg.cpu().numpy().astype(np.int64)
    leaf = d["leaf_of"][t[gstart[:-1]]]
    r = d["root"]
    root = (d["cx"][r], d["cy"][r], d["cz"][r], float(d["wn_len"][r]), d["mass"][r])
    want = mo.group_features(gstart, x[t], y[t], z[t], m[t], soft[t], aold.cpu().numpy(),
                             d["wn_len"][leaf].astype(np.float64), d["level"][leaf], root)
    assert feats.shape == want.shape
    np.testing.assert_allclose(feats, want, rtol=0, atol=1e-5)


def test_learned_walk_falls_back_to_reference(gz):
    """With an empty trusted label range every group keeps the reference tolerance, and the
    learned walk is the reference walk bit for bit."""
    P, T, tg, aold = _walk_setup(gz)
    M = gz.Model(mo.init_params(1).astype(np.float32))
    M.set_calibration(np.full(16, -1e30), np.full(16, 1e30), 5.0, 4.0, 0.0, np.log10(0.0025))
    ref = gz.accel_grouped(T, P, tg, 0.0, 1.0, 0, aold=aold, group_size=32, precision="fp32",
                           counts=True)
    lrn = gz.accel_grouped(T, P, tg, 0.0, 1.0, 0, aold=aold, group_size=32, precision="fp32",
                           counts=True, model=M)
    assert lrn[3]["fallback_groups"] == lrn[3]["groups"]
    for k in ("nodes_visited", "multipole", "pairs"):
        assert lrn[3][k] == ref[3][k]
    for u, v in zip(ref[:3], lrn[:3]):
        assert torch.equal(u, v)


def test_learned_walk_tolerance_scales_the_reference_test(gz):
    """A network that predicts a constant log10(alpha) opens exactly like the reference walk run
    with aold scaled by alpha / alpha_ref."""
    P, T, tg, aold = _walk_setup(gz)
    th = np.zeros(mo.N_PARAMS, dtype=np.float32)
    th[-1] = np.log10(0.01)  # b3: the prediction for every group
    M = gz.Model(th)
    pred = float(M.forward(torch.zeros((1, 16), device="cuda")).cpu()[0])  # bf16-rounded b3
    M.set_calibration(np.full(16, -1e30), np.full(16, 1e30), -10.0, 10.0, 0.0, np.log10(0.0025))
    lrn = gz.accel_grouped(T, P, tg, 0.0, 1.0, 0, aold=aold, group_size=32, precision="fp32",
                           counts=True, model=M)
    scale = np.float32(10.0) ** np.float32(pred - np.float32(np.log10(0.0025)))
    ref = gz.accel_grouped(T, P, tg, 0.0, 1.0, 0, aold=aold * float(scale), group_size=32,
                           precision="fp32", counts=True)
    assert lrn[3]["fallback_groups"] == 0
    # the scale is applied to the group minimum in FP64 in one case and to every aold in the
    # other; the interaction sets agree to within a handful of borderline nodes
    assert abs(lrn[3]["multipole"] - ref[3]["multipole"]) <= 1e-4 * ref[3]["multipole"]
    assert lrn[3]["multipole"] < 0.8 * gz.accel_grouped(T, P, tg, 0.0, 1.0, 0, aold=aold,
                                                       group_size=32, precision="fp32",
                                                       counts=True)[3]["multipole"]


def test_trained_network_against_the_method_without_it(gz):
    """What the trained criterion (learned_criterion.json) gains on a 1M-particle halo+disk, a
    distribution it never saw, judged by the FP64 direct sum on 8,192 probes: against the
    reference criterion (ErrTolForceAcc 0.0025) it does at least 30% fewer interactions at a p99
    error within 1e-2; against one global tolerance tuned to the same p99 (interpolated between
    the grid points that bracket it) it is no worse than 3%. Measured: 34% fewer than the
    reference, 3% fewer than the matched global tolerance, p99 9.3e-3 (6.7e-3 at 8M) -- the
    network adds little over a tuned constant."""
    from gizmo_shared_memory_amd import ics, learned
    x, y, z, m, soft = ics.halo_disk(n=1 << 20)
    P = gz.Particles(x, y, z, m, soft)
    T = gz.build(P)
    tg = T.device_orderbuf()
    boot = gz.accel_grouped(T, P, tg, 0.7, 1.0, 0)
    aold = 0.0025 * torch.linalg.norm(torch.stack(boot, 1), dim=1)
    probes = ics.probes(int(tg.shape[0]), 8192)
    pt = tg.view(torch.int32)[torch.as_tensor(probes.astype(np.int64), device="cuda:0")]
    ex = np.stack([_np(v) for v in gz.accel_brute(P, pt.view(torch.uint32), 1.0)], 1)

    def run(scale=1.0, model=None):
        r = gz.accel_grouped(T, P, tg, 0.0, 1.0, 0, aold=aold * scale, precision="fp32",
                             group_size=64, counts=True, model=model)
        g = np.stack([_np(v)[probes] for v in r[:3]], 1)
        e = np.linalg.norm(g - ex, axis=1) / np.linalg.norm(ex, axis=1)
        return (r[3]["multipole"] + r[3]["pairs"]) / tg.shape[0], float(np.percentile(e, 99))

    plain_n, plain_p99 = run()
    lrn_n, lrn_p99 = run(model=learned.load_model())
    grid = [run(2.0 ** (k / 2.0)) for k in range(2, 12)]
    print(f"reference: {plain_n:.0f} interactions/target, p99 {plain_p99:.3g}; learned: "
          f"{lrn_n:.0f}, p99 {lrn_p99:.3g}; global grid {[(round(a), round(b, 5)) for a, b in grid]}")
    assert lrn_p99 <= 1e-2
    assert lrn_n < 0.7 * plain_n
    lo = [g for g in grid if g[1] <= lrn_p99]
    hi = [g for g in grid if g[1] > lrn_p99]
    assert lo and hi, "the global grid does not bracket the learned p99"
    (n0, e0), (n1, e1) = lo[-1], hi[0]
    w = (np.log(lrn_p99) - np.log(e0)) / (np.log(e1) - np.log(e0))
    matched = float(np.exp(np.log(n0) + w * (np.log(n1) - np.log(n0))))
    assert lrn_n <= 1.03 * matched, (lrn_n, matched)
