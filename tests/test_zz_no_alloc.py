"""gz_order_targets_noalloc, the allocation-free variant of gz_order_targets, called through the C
ABI: its output equals gz_order_targets' on consecutive calls that share the fixed scratch (the
second on a smaller tree and active set, so it runs over what the first left), on the all-active
and empty cases, and it refuses a tree larger than its fixed scratch."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from tests import tree_cases

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def gz():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from gizmo_shared_memory_amd import tree
    return tree


def _noalloc(gz, T, active):
    from gizmo_shared_memory_amd import _lib as L
    out = torch.full((max(T.n_live, 1),), -7, dtype=torch.int32, device=T.P.device)
    n_out = C.c_size_t(0)
    T.ctx.enter()
    if active is None:
        st = L.load().gz_order_targets_noalloc(T.ctx.handle, T.handle, None, T.P.n,
                                               C.c_void_p(out.data_ptr()), C.byref(n_out))
    else:
        act = torch.as_tensor(np.ascontiguousarray(active.astype(np.uint32).view(np.int32)),
                              device=T.P.device)
        st = L.load().gz_order_targets_noalloc(T.ctx.handle, T.handle,
                                               C.c_void_p(act.data_ptr()), int(act.shape[0]),
                                               C.c_void_p(out.data_ptr()), C.byref(n_out))
    return st, out, int(n_out.value)


def _check(gz, T, active):
    st, out, n = _noalloc(gz, T, active)
    assert st == 0, gz_last_error()
    want = gz.order_targets(T, active)
    assert n == int(want.shape[0])
    assert torch.equal(out[:n].cpu(), want.view(torch.int32).cpu())


def gz_last_error():
    from gizmo_shared_memory_amd import _lib as L
    return L.load().gz_last_error().decode()


def test_noalloc_order_equals_allocating_order(gz):
    rng = np.random.default_rng(77)
    big = tree_cases.plummer_small(300000, seed=3, eps=0.01)
    small = tree_cases.plummer_small(20000, seed=4, eps=0.01)
    x, y, z, m, soft = big
    m = m.copy()
    m[::101] = 0.0  # tombstones: not in the tree, dropped from the ordered list
    print("[noalloc] call 1: n=300000, 90000 active", flush=True)
    T1 = gz.build(gz.Particles(x, y, z, m, soft))
    act1 = rng.permutation(300000)[:90000]
    _check(gz, T1, act1)
    print("[noalloc] call 2: n=20000, 3000 active (smaller, over the first call's scratch)",
          flush=True)
    T2 = gz.build(gz.Particles(*small))
    act2 = rng.permutation(20000)[:3000]
    _check(gz, T2, act2)
    print("[noalloc] all active, empty active set", flush=True)
    _check(gz, T2, None)
    _check(gz, T2, np.zeros(0, np.uint32))


def test_noalloc_refuses_trees_beyond_its_scratch(gz):
    n = (1 << 23) + 1
    rng = np.random.default_rng(5)
    print(f"[noalloc] refusal: n={n}", flush=True)
    x, y, z = (rng.uniform(-1, 1, n) for _ in range(3))
    m = np.full(n, 1.0 / n)
    T = gz.build(gz.Particles(x, y, z, m, np.full(n, 0.01)))
    st, out, k = _noalloc(gz, T, np.arange(10, dtype=np.uint32))
    assert st == 1 and k == 0  # GZ_ERR_INVALID
    assert "fixed scratch" in gz_last_error()
