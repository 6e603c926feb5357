"""Barnes-Hut collision pass (system::bh_s_coll, gregsys.hpp:631-686, with
bh_cube::nn_iterator, greg8tree.hpp:615-700): the oracle, on the CPU.

Pins:
  * the long-double oracle reproduces the compiled reference's own pass
    (oracle/_ref/ref_bhcoll calling the private method on the reference's
    tree) bit for bit on dense clusters (fixtures ref_bhcoll_*.npz);
  * the neighbour sequence equals, by brute force over the exported tree,
    every leaf whose closed box touches the own leaf's box, in DFS pre-order,
    own leaf excluded -- and the relation is symmetric;
  * the double form (the GPU's) stays within 1e-12 of the long-double form.
"""
import glob
import os

import numpy as np
import pytest

import oracle as O


def fixtures(golden):
    return sorted(glob.glob(os.path.join(golden, "ref_bhcoll_*.npz")))


def test_fixtures_present(golden):
    assert len(fixtures(golden)) >= 3


@pytest.mark.parametrize("name", ["ref_bhcoll_n64_seed9", "ref_bhcoll_n300_seed7", "ref_bhcoll_n2000_seed8"])
def test_oracle_ld_matches_reference(golden, name):
    f = np.load(os.path.join(golden, name + ".npz"))
    v = O.bh_s_coll(f["m"], f["r"], f["e"], f["pos"], f["vel"], dtype=O.LD)
    assert np.array_equal(v, f["vel_after"])
    moved = (f["vel_after"] != f["vel"].astype(O.LD)).any(axis=1)
    assert moved.sum() > len(moved) // 2  # the case is dense: most bodies bounce


def node_boxes(tree):
    """Per-node cube corners by replaying the octant digits from the root box
    (the same halving as set_corners)."""
    k = len(tree["depth"])
    lo = np.zeros((k, 3))
    hi = np.zeros((k, 3))
    lo[0], hi[0] = tree["box"][:3], tree["box"][3:]
    for p in range(1, k):
        q = tree["parent"][p]
        o = tree["octant"][p]
        for c in range(3):
            mid = (lo[q, c] + hi[q, c]) * 0.5
            if (o >> c) & 1:
                lo[p, c], hi[p, c] = mid, hi[q, c]
            else:
                lo[p, c], hi[p, c] = lo[q, c], mid
    return lo, hi


def test_neighbours_are_touching_leaves_in_preorder():
    rng = np.random.default_rng(3)
    n = 400
    pos = np.concatenate([rng.normal(0, 1, (n // 2, 3)), rng.uniform(-4, 4, (n - n // 2, 3))])
    m = rng.uniform(1, 2, n)
    tree = O.bh_tree(m, pos)
    lo, hi = node_boxes(tree)
    # the tree's own side lengths confirm the replayed boxes
    assert np.array_equal(hi[:, 0] - lo[:, 0], tree["side"])
    leaves = np.flatnonzero(tree["body"] >= 0)  # pre-order
    leaf_of = {int(tree["body"][p]): p for p in leaves}
    nb_sets = {}
    for b in range(0, n, 7):
        own = leaf_of[b]
        touch = ~((hi[own] < lo[leaves]).any(axis=1) | (lo[own] > hi[leaves]).any(axis=1))
        expect = [int(tree["body"][p]) for p, t in zip(leaves, touch) if t and p != own]
        got = O.bh_neighbours(m, pos, b)
        assert got.tolist() == expect
        nb_sets[b] = set(expect)
    for b, s in nb_sets.items():
        for j in s:
            if j in nb_sets:
                assert b in nb_sets[j]


def test_double_form_close_to_long_double(golden):
    f = np.load(os.path.join(golden, "ref_bhcoll_n2000_seed8.npz"))
    vd = O.bh_s_coll(f["m"], f["r"], f["e"], f["pos"], f["vel"], dtype=np.float64)
    vl = f["vel_after"].astype(np.float64)
    scale = np.abs(vl).max()
    assert np.max(np.abs(vd - vl)) <= 1e-12 * scale


def test_kdk_coll_without_radii_is_plain_bh():
    rng = np.random.default_rng(5)
    n = 200
    pos = rng.normal(0, 1, (n, 3))
    vel = rng.normal(0, 0.1, (n, 3))
    m = np.full(n, 1.0 / n)
    p1, v1, _ = O.bh_kdk(m, pos, vel, 1.0, 1e-3, 0.01, 0.5, 5)
    p2, v2, _ = O.bh_kdk_coll(m, np.zeros(n), np.ones(n), pos, vel, 1.0, 1e-3, 0.01, 0.5, 5)
    assert np.array_equal(p1, p2) and np.array_equal(v1, v2)
