"""Pins the oracle's audit_tree restatement (gzo_audit_tree, shmem/tree.cc:973-1119): it finds
nothing wherever the reference's own audit finds nothing, on the oracle's trees and on the
reference's, and it catches each invariant when a mistake is planted in a copy of the tree arrays
(on the CPU). CPU only."""
from __future__ import annotations

import copy

import numpy as np
import pytest

from oracle.bindings import HostParticles
from tests import tree_cases


def _hp(case):
    return HostParticles(*case)


@pytest.mark.parametrize("name", sorted(tree_cases.CASES))
def test_audit_clean_where_reference_is_clean(oracle, reference, name):
    P = _hp(tree_cases.CASES[name]())
    to = oracle.build(P)
    rp = reference.particles(P)
    tr = reference.build(rp)
    try:
        bad, msg = reference.audit(tr, rp)
        assert bad == 0, msg
        bad, msg = oracle.audit_tree(oracle.tree(to, P), P)
        assert bad == 0, msg
        bad, msg = oracle.audit_tree(reference.tree(tr, P.n), P)
        assert bad == 0, msg
    finally:
        oracle.free(to)
        reference.free(tr)
        reference.free_particles(rp)


@pytest.fixture(scope="module")
def base(oracle):
    P = _hp(tree_cases.plummer_small(5000, seed=12, eps=0.01))
    to = oracle.build(P)
    T = oracle.tree(to, P)
    oracle.free(to)
    return P, T


def _interior_with_two_children(T):
    for X in range(T.nnodes):
        c = T.first[X]
        if c >= 0 and T.next[c] != T.next[X]:
            return X, c
    raise AssertionError("no interior node with two children")


def _planted(oracle, P, T, edit, **kw):
    T2 = copy.deepcopy(T)
    edit(T2)
    return oracle.audit_tree(T2, P, **kw)


def test_audit_catches_planted_mistakes(oracle, base):
    P, T = base
    assert oracle.audit_tree(T, P) == (0, "")
    X, c = _interior_with_two_children(T)
    leaf = int(T.leaf_of[0])

    def skip_child(t):
        t.next[c] = t.next[t.next[c]] if t.next[t.next[c]] != t.next[X] else t.next[X]
    bad, msg = _planted(oracle, P, T, skip_child)
    assert bad > 0

    def wrong_leaf_of(t):
        t.leaf_of[0] = t.parent[leaf]
    bad, msg = _planted(oracle, P, T, wrong_leaf_of)
    assert bad > 0 and "leaf_of[0]" in msg

    def wrong_parent(t):
        t.parent[c] = c
    bad, msg = _planted(oracle, P, T, wrong_parent)
    assert bad > 0 and "parent" in msg

    def duplicate_particle(t):
        t.orderbuf[t.plo[leaf] + 1 if t.phi[leaf] - t.plo[leaf] > 1 else t.plo[leaf]] = 0
        t.orderbuf[t.plo[leaf]] = 0
    bad, msg = _planted(oracle, P, T, duplicate_particle)
    assert bad > 0

    def heavier_node(t):
        t.mass[X] *= 1 + 1e-8
    bad, msg = _planted(oracle, P, T, heavier_node)
    assert bad > 0 and "mass" in msg

    def moved_com(t):
        t.cx[X] += 1e-4 * t.size[X]
    bad, msg = _planted(oracle, P, T, moved_com)
    assert bad > 0 and "COM" in msg
    # ...which a tree widened by a speed bound tolerates in proportion
    assert oracle.audit_tree(T, P, vmax=np.full(T.nnodes, 0.0, np.float32))[0] == 0

    def low_soft(t):
        t.wn_soft[leaf] = np.float32(P.soft[0] * 0.5)
    bad, msg = _planted(oracle, P, T, low_soft)
    assert bad > 0 and "soft" in msg


def test_audit_catches_particle_outside_box_and_slow_vmax(oracle, base):
    P, T = base
    leaf = int(T.leaf_of[7])
    x = P.x.copy()
    x[7] = T.gcx[leaf] + 0.6 * T.size[leaf]  # outside its leaf's box
    P2 = HostParticles(x, P.y, P.z, P.m, P.soft)
    bad, msg = oracle.audit_tree(T, P2, check_com=False)
    assert bad > 0 and "outside node" in msg
    # a speed bound that covers the motion over tau admits it
    v = np.zeros((3, P.n))
    v[0, 7] = 1.0
    vmax = np.full(T.nnodes, 1.0, np.float32)
    bad, msg = oracle.audit_tree(T, P2, vel=v, vmax=vmax, tau=0.2 * T.size[leaf],
                                 check_com=False)
    assert bad == 0, msg
    # and a bound below |v| is caught
    bad, msg = oracle.audit_tree(T, P, vel=v, vmax=np.full(T.nnodes, 0.5, np.float32))
    assert bad > 0 and "vmax" in msg
