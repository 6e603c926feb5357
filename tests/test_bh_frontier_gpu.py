"""The FP32 Barnes-Hut frontier walk (tuning knob bh_walk = 1, the default)
against the per-wave depth-first walk (bh_walk = 2), through the C ABI.

Both walks accept the same cells for each 64-body group (a cell is a
monopole only when every body of the group passes the opening test), so:
  * the interaction counts are identical, for every bucket size and shape;
  * the accelerations differ only by FP32 summation order (relative
    difference below 1e-4 per body);
  * a frontier stack too small for the tree (bh_stack = 64) makes the walk
    fall back to the depth-first walk for the whole step: the result is then
    bit-identical to bh_walk = 2;
  * repeated walks are bit-identical (the heaviest-first group order changes
    between calls; each group's sums do not depend on it);
  * theta = 0 accepts no cell: both walks reduce to the direct sum.
"""
import numpy as np
import pytest

from gregpackagescpp_amd import capi

pytestmark = pytest.mark.gpu


def system(m, pos, eps=0.01, theta=0.5, walk=1, bucket=0, stack=0):
    s = capi.System(0)
    s.set_G(1.0)
    s.set_softening(eps)
    s.set_bh_opening_angle(theta)
    s.set_bodies(m, pos, np.zeros_like(pos))
    s.set_precision(capi.PREC_MIXED)
    if walk:
        s.set_tuning("bh_walk", walk)
    if bucket:
        s.set_tuning("bh_bucket", bucket)
    if stack:
        s.set_tuning("bh_stack", stack)
    return s


def rel_diff(a, b):
    return np.linalg.norm(a - b, axis=1) / np.maximum(np.linalg.norm(b, axis=1), 1e-300)


@pytest.mark.parametrize("bucket", [1, 8, 16, 32])
def test_frontier_same_cells_as_wave_walk(bucket):
    print(f"case frontier vs wave walk bucket={bucket}", flush=True)
    m, pos, _ = capi.ic_two_plummer(41, 65536)
    a, ca = system(m, pos, bucket=bucket).bh_accelerations()
    b, cb = system(m, pos, walk=2, bucket=bucket).bh_accelerations()
    assert ca == cb
    assert rel_diff(a, b).max() < 1e-4


def test_frontier_far_outliers_and_ragged_group():
    """Bodies far from the cluster make 64-body groups whose box spans the
    domain (the box test rejects everything; the per-body test decides), and
    n = 5000 leaves a partial last group and XCD ranges of unequal size."""
    print("case frontier outliers", flush=True)
    m, pos, _ = capi.ic_plummer(42, 5000)
    rng = np.random.default_rng(42)
    far = rng.choice(5000, 40, replace=False)
    pos[far] = rng.uniform(-500, 500, (40, 3))
    a, ca = system(m, pos).bh_accelerations()
    b, cb = system(m, pos, walk=2).bh_accelerations()
    assert ca == cb
    assert rel_diff(a, b).max() < 1e-4


def test_frontier_unsoftened():
    print("case frontier eps=0", flush=True)
    m, pos, _ = capi.ic_two_plummer(43, 16384)
    a, ca = system(m, pos, eps=0.0).bh_accelerations()
    b, cb = system(m, pos, eps=0.0, walk=2).bh_accelerations()
    assert ca == cb
    assert np.all(np.isfinite(a))
    assert rel_diff(a, b).max() < 1e-4


def test_frontier_small_systems():
    print("case frontier small", flush=True)
    for n in (2, 63, 65, 700):
        m, pos, _ = capi.ic_plummer(44 + n, n)
        a, ca = system(m, pos).bh_accelerations()
        b, cb = system(m, pos, walk=2).bh_accelerations()
        assert ca == cb, n
        assert rel_diff(a, b).max() < 1e-4, n


def test_frontier_stack_overflow_falls_back():
    print("case frontier stack overflow", flush=True)
    m, pos, _ = capi.ic_two_plummer(45, 65536)
    a, ca = system(m, pos, stack=64).bh_accelerations()
    b, cb = system(m, pos, walk=2).bh_accelerations()
    assert ca == cb
    assert np.array_equal(a, b)


def test_frontier_repeatable_and_evolves():
    print("case frontier repeat", flush=True)
    m, pos, vel = capi.ic_two_plummer(46, 65536)
    s = system(m, pos)
    first, c1 = s.bh_accelerations()
    for _ in range(3):
        again, c2 = s.bh_accelerations()
        assert c1 == c2
        assert np.array_equal(first, again)
    # a few KDK steps with each walk stay close (same cells, FP32 order only)
    out = []
    for walk in (1, 2):
        t = system(m, pos, walk=walk)
        t.set_bodies(m, pos, vel)
        t.set_timestep(1e-3)
        t.set_iterations(5)
        t.evolve(capi.LEAPFROG_KDK | capi.BARNES_HUT)
        out.append(t.state()["vel"])
    dv = np.linalg.norm(out[0] - out[1], axis=1) / np.linalg.norm(out[1] - vel, axis=1)
    assert np.median(dv) < 1e-4


def test_frontier_theta_zero_is_direct_sum():
    """theta = 0 accepts no cell: every body is summed body by body (counts
    n*n with the own body) by the frontier walk itself, not its fallback."""
    print("case frontier theta=0", flush=True)
    n = 4096
    m, pos, _ = capi.ic_plummer(47, n)
    a, ca = system(m, pos, theta=0.0).bh_accelerations()
    b, cb = system(m, pos, theta=0.0, walk=2).bh_accelerations()
    assert ca == cb == n * n
    d = pos[None, :, :] - pos[:, None, :]
    r2 = (d * d).sum(2) + 0.01 ** 2
    ref = (m[None, :, None] * d / (r2 * np.sqrt(r2))[:, :, None]).sum(1)
    assert rel_diff(a, ref).max() < 1e-4
    assert rel_diff(b, ref).max() < 1e-4
    # the frontier walk's own sums were used: a fallback to the depth-first
    # walk would return b bit for bit, while the frontier walk adds the same
    # bodies in another order (bucket ranges in pop order, not DFS order)
    assert not np.array_equal(a, b)
