"""The construction-time overlap check (system::check_overlap,
gregsys.hpp:307-342) in the oracle, pinned to the compiled reference.

Fixtures (tests/golden/ref_dropin_{merge,check}_*.npz) come from
examples/drop_in_driver.cpp built against the reference's headers: the
bodies that survive construction with mergeOverlappingBodies, and whether
the plain check throws.  The cases cover a merge whose grown body then
reaches a body before it (the reference erases the body after that one),
an outer body of zero radius, rubble clusters whose first overlapping outer
sits at several depths, and a cluster without any overlap.

Bars: the long-double restatement reproduces the surviving bodies exactly
(values as the driver printed them, rounded to double); the double
restatement keeps the same bodies within 1e-13 relative; the restatement
throws exactly when the reference does.
"""
import glob
import os

import numpy as np
import pytest

import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")
MERGE = sorted(glob.glob(os.path.join(GOLD, "ref_dropin_merge_*.npz")))
CHECK = sorted(glob.glob(os.path.join(GOLD, "ref_dropin_check_*.npz")))


def as_rows(res):
    k = res["n"]
    out = np.zeros((k, 9))
    out[:, 0], out[:, 1] = res["m"].astype(np.float64), res["r"].astype(np.float64)
    out[:, 2:5], out[:, 5:8] = res["pos"].astype(np.float64), res["vel"].astype(np.float64)
    out[:, 8] = res["e"].astype(np.float64)
    return out


@pytest.mark.parametrize("path", MERGE, ids=os.path.basename)
def test_merge_loop_matches_reference(path):
    f = np.load(path)
    ld = O.check_overlap(f["m"], f["r"], f["pos"], f["vel"], f["e"], merge=True)
    assert np.array_equal(as_rows(ld), f["bodies"])
    d = O.check_overlap(f["m"], f["r"], f["pos"], f["vel"], f["e"], merge=True, dtype=np.float64)
    assert d["n"] == ld["n"]
    scale = np.abs(f["bodies"]).max(axis=0) + 1e-300
    assert (np.abs(as_rows(d) - f["bodies"]) <= 1e-13 * scale).all()


@pytest.mark.parametrize("path", CHECK, ids=os.path.basename)
def test_check_throws_as_reference(path):
    f = np.load(path)
    first = O.check_overlap(f["m"], f["r"], f["pos"], f["vel"], f["e"], merge=False)
    assert (first is not None) == bool(f["throws"])


def test_fixtures_exercise_the_loop():
    """The cases reach a prefix merge, a zero-radius outer, truncation deep in
    the list and the no-overlap path."""
    counts = {os.path.basename(p): (len(np.load(p)["m"]), len(np.load(p)["bodies"])) for p in MERGE}
    assert counts["ref_dropin_merge_prefix.npz"] == (6, 2)
    assert counts["ref_dropin_merge_rubble11.npz"] == (60, 60)
    assert any(0 < b < n - 1 and b > 5 for n, b in counts.values())


def test_first_pair_brute_force():
    """The first overlapping pair in (outer, inner > outer) order against a
    numpy brute force over all pairs."""
    rng = np.random.default_rng(5)
    for trial in range(20):
        n = 80
        pos = rng.uniform(-1, 1, (n, 3))
        r = rng.uniform(0, 0.08, n)
        m = np.ones(n)
        d = np.sqrt(((pos[None, :, :] - pos[:, None, :]) ** 2).sum(-1))
        ov = (d < r[:, None] + r[None, :]) & np.triu(np.ones((n, n), bool), 1)
        want = None
        if ov.any():
            i = int(np.where(ov.any(1))[0][0])
            want = (i, int(np.where(ov[i])[0][0]))
        assert O.check_overlap(m, r, pos, np.zeros((n, 3)), merge=False, dtype=np.float64) == want
