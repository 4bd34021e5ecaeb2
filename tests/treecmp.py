"""Canonical comparison of two builds of the same tree.

Both trees are put in DFS preorder with children in octant order; then node i of one is node i of
the other, and every per-node field can be compared directly. The sorted particle order is
compared modulo the order inside runs of equal keys, which the reference's unstable sort leaves
unspecified (tree.cc:420); those runs always lie inside one leaf.
"""
from __future__ import annotations

import numpy as np

from oracle.bindings import preorder_from_children, preorder_from_threaded


def canonical(t, nchild=None):
    """Return (perm, fields) where perm[i] is the node id of the i-th node in preorder and fields
    maps field name to the array in preorder. Child links are renumbered to preorder ids."""
    if nchild is None:
        perm = preorder_from_threaded(t.first, t.next, t.root)
    else:
        perm = preorder_from_children(t.first, nchild, t.root)
    inv = np.full(t.nnodes, -1, dtype=np.int64)
    inv[perm] = np.arange(perm.shape[0])
    f = {}
    for name in ("plo", "phi", "size", "mass", "cx", "cy", "cz", "soft", "delta", "gcx", "gcy",
                 "gcz", "wn_s", "wn_soft", "wn_len"):
        f[name] = getattr(t, name)[perm]
    first = t.first[perm]
    f["first"] = np.where(first >= 0, inv[np.maximum(first, 0)], -1)
    parent = t.parent[perm]
    f["parent"] = np.where(parent >= 0, inv[np.maximum(parent, 0)], -1)
    nxt = t.next[perm]
    f["next"] = np.where(nxt >= 0, inv[np.maximum(nxt, 0)], -1)
    return perm, inv, f


BIT_FIELDS = ("plo", "phi", "first", "next", "parent", "size", "gcx", "gcy", "gcz", "mass", "cx",
              "cy", "cz", "soft", "delta", "wn_s", "wn_soft", "wn_len")


def assert_same_tree(a, b, a_nchild=None, b_nchild=None, fields=BIT_FIELDS, what=("a", "b")):
    """Assert that trees a and b have the same canonical topology and bit-identical fields."""
    assert a.nnodes == b.nnodes, f"node count {what[0]}={a.nnodes} {what[1]}={b.nnodes}"
    assert a.n_live == b.n_live
    _, ainv, fa = canonical(a, a_nchild)
    _, binv, fb = canonical(b, b_nchild)
    assert fa["plo"].shape[0] == a.nnodes, "preorder did not reach every node"
    for name in fields:
        x, y = fa[name], fb[name]
        if x.dtype.kind == "f":
            same = (x.view(np.uint64 if x.dtype == np.float64 else np.uint32) ==
                    y.view(np.uint64 if y.dtype == np.float64 else np.uint32))
        else:
            same = x == y
        if not np.all(same):
            i = int(np.nonzero(~same)[0][0])
            raise AssertionError(f"field {name} differs at preorder node {i}: "
                                 f"{what[0]}={x[i]!r} {what[1]}={y[i]!r} "
                                 f"({int((~same).sum())} nodes differ)")
    # leaf membership: leaf_of mapped to preorder ids must agree for every live particle
    la = np.where(a.leaf_of >= 0, ainv[np.maximum(a.leaf_of, 0)], -1)
    lb = np.where(b.leaf_of >= 0, binv[np.maximum(b.leaf_of, 0)], -1)
    assert np.array_equal(la, lb), "leaf_of differs"


def assert_same_order_modulo_ties(order_a, order_b, keys_sorted):
    """orderbuf equality except inside runs of equal sorted keys, where only the set matters."""
    assert order_a.shape == order_b.shape
    if order_a.shape[0] == 0:
        return
    brk = np.nonzero(np.diff(keys_sorted) != 0)[0] + 1
    starts = np.concatenate([[0], brk])
    ends = np.concatenate([brk, [keys_sorted.shape[0]]])
    single = (ends - starts) == 1
    idx = starts[single]
    assert np.array_equal(order_a[idx], order_b[idx]), "orderbuf differs outside equal-key runs"
    for s, e in zip(starts[~single], ends[~single]):
        assert np.array_equal(np.sort(order_a[s:e]), np.sort(order_b[s:e])), \
            f"equal-key run [{s},{e}) holds different particles"
