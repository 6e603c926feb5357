"""Construction-time overlap check on the GPU (nb_set_bodies with
check_overlap, nb_merge_overlapping) against the oracle's restatement of
system::check_overlap (gregsys.hpp:307-342), itself pinned to the compiled
reference (tests/test_overlap.py).

Bars: the reported pair is the oracle's first (outer, inner > outer) pair;
after nb_merge_overlapping the bodies (m, r, restitution, x, v) and labels
equal the oracle's double restatement bit for bit; a context that fails the
check holds no bodies.
"""
import glob
import os

import numpy as np
import pytest

import oracle as O
from gregpackagescpp_amd import capi

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")
CASES = sorted(glob.glob(os.path.join(GOLD, "ref_dropin_merge_*.npz")))


def ctx(m, r, e, pos, vel, check=False):
    s = capi.System(0)
    s.set_bodies(m, pos, vel, r=r, rest_c=e, check_overlap=check)
    return s


@pytest.mark.parametrize("path", CASES, ids=os.path.basename)
def test_merge_overlapping_bit_exact(path):
    f = np.load(path)
    print(f"case merge {os.path.basename(path)}", flush=True)
    s = ctx(f["m"], f["r"], f["e"], f["pos"], f["vel"])
    removed = s.merge_overlapping()
    ref = O.check_overlap(f["m"], f["r"], f["pos"], f["vel"], f["e"], merge=True, dtype=np.float64)
    st = s.state()
    assert removed == len(f["m"]) - ref["n"]
    assert s.num_bodies() == ref["n"]
    for key in ("m", "r", "pos", "vel"):
        assert np.array_equal(st[key], ref[key]), key
    assert np.array_equal(s.ids(), ref["labels"])


@pytest.mark.parametrize("path", CASES, ids=os.path.basename)
def test_check_reports_first_pair(path):
    f = np.load(path)
    print(f"case check {os.path.basename(path)}", flush=True)
    want = O.check_overlap(f["m"], f["r"], f["pos"], f["vel"], f["e"], merge=False, dtype=np.float64)
    if want is None:
        s = ctx(f["m"], f["r"], f["e"], f["pos"], f["vel"], check=True)
        assert s.num_bodies() == len(f["m"])
        return
    s = capi.System(0)
    with pytest.raises(capi.NBodyError) as e:
        s.set_bodies(f["m"], f["pos"], f["vel"], r=f["r"], rest_c=f["e"], check_overlap=True)
    assert e.value.code == capi.NB_ERR_OVERLAPPING_BODIES
    assert f"id={want[0]} and id={want[1]} " in str(e.value)
    assert s.num_bodies() == 0


def test_check_large_n_planted_pairs():
    """200,000 bodies in a unit cube with radii 1e-7 (no natural overlap: the
    expected number of overlapping pairs is ~1e-9) and three planted
    overlapping pairs; the check reports the first in (outer, inner) order,
    after the pass has covered the ~2e10 pairs.  Without the planted pairs the
    bodies pass."""
    n = 200000
    rng = np.random.default_rng(9)
    pos = rng.uniform(0, 1, (n, 3))
    r = np.full(n, 1e-7)
    m = np.ones(n)
    vel = np.zeros((n, 3))
    print("case large check clean", flush=True)
    s = capi.System(0)
    s.set_bodies(m, pos, vel, r=r, check_overlap=True)
    assert s.num_bodies() == n
    planted = pos.copy()
    for (i, j) in ((150001, 199990), (150001, 160000), (170000, 170001)):
        planted[j] = planted[i] + 5e-8
    print("case large check planted", flush=True)
    with pytest.raises(capi.NBodyError) as e:
        s.set_bodies(m, planted, vel, r=r, check_overlap=True)
    assert "id=150001 and id=160000 " in str(e.value)


def test_merge_then_evolve():
    """A system built by merging steps on like one built from the survivors."""
    f = np.load(os.path.join(GOLD, "ref_dropin_merge_rubble20.npz"))
    s = ctx(f["m"], f["r"], f["e"], f["pos"], f["vel"])
    s.merge_overlapping()
    st = s.state()
    t = ctx(st["m"], st["r"], np.full(len(st["m"]), 0.5), st["pos"], st["vel"])
    for x in (s, t):
        x.set_G(6.6743e-11)
        x.set_timestep(5.0)
        x.set_iterations(20)
        x.evolve()
    assert np.array_equal(s.state()["pos"], t.state()["pos"])
