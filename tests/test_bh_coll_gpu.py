"""Barnes-Hut KDK with collisions on the GPU (bh_s_coll after each step's
kick, on that step's tree), through the C ABI, against the oracle.

Bars:
  * FP64 path: positions and velocities bit-identical to the double oracle
    (orc_bh_kdk_coll_d), whose long-double form reproduces the compiled
    reference's pass bit for bit (tests/test_bh_coll.py).  Masses are 1/N, a
    power of two, so node COMs are exact as in the plain BH test;
  * mergers on: the same result (the reference's BH path only bounces);
  * mixed precision (FP32 walk): the pass itself is FP64; velocities track
    the FP64 run (median relative difference of the velocity change < 1e-2).
"""
import numpy as np
import pytest

import oracle as O
from gregpackagescpp_amd import capi

pytestmark = pytest.mark.gpu


def cluster(seed, n, radius):
    m, pos, vel = capi.ic_two_plummer(seed, n)
    rng = np.random.default_rng(seed)
    r = np.full(n, radius)
    e = rng.uniform(0.1, 0.9, n)
    return m, r, e, pos, vel


def run(m, r, e, pos, vel, steps, prec=capi.PREC_FP64, mergers=False, coll=True):
    s = capi.System(0)
    s.set_G(1.0)
    s.set_softening(0.01)
    s.set_bh_opening_angle(0.5)
    s.set_timestep(1e-3)
    s.set_precision(prec)
    if coll:
        s.set_collisions(capi.OVERLAP_COLL_CHECK, mergers=mergers, min_tot_com_mom=1e10)
    s.set_bodies(m, pos, vel, r=r, rest_c=e)
    s.set_iterations(steps)
    s.evolve(capi.LEAPFROG_KDK | capi.BARNES_HUT)
    return s.state()


@pytest.mark.parametrize("n,radius", [(4096, 0.03), (16384, 0.02)])
def test_bh_coll_fp64_bit_identical(n, radius):
    print(f"case bh coll fp64 n={n} r={radius}", flush=True)
    m, r, e, pos, vel = cluster(21, n, radius)
    one = O.bh_s_coll(m, r, e, pos, vel, dtype=np.float64)
    bounced = int((one != vel).any(axis=1).sum())
    print(f"  bodies one pass on the initial state bounces: {bounced}", flush=True)
    assert bounced > 100
    st = run(m, r, e, pos, vel, 5)
    rp, rv, _ = O.bh_kdk_coll(m, r, e, pos, vel, 1.0, 1e-3, 0.01, 0.5, 5)
    assert np.array_equal(st["pos"], rp)
    assert np.array_equal(st["vel"], rv)


def test_bh_coll_after_steps_without_collisions():
    """BH steps without collisions (radii and restitutions are then not moved
    into tree order), then with them on the same context: bit-identical to
    the double oracle run in the same two parts."""
    print("case bh coll after plain steps", flush=True)
    m, r, e, pos, vel = cluster(24, 4096, 0.03)
    s = capi.System(0)
    s.set_G(1.0)
    s.set_softening(0.01)
    s.set_bh_opening_angle(0.5)
    s.set_timestep(1e-3)
    s.set_precision(capi.PREC_FP64)
    s.set_bodies(m, pos, vel, r=r, rest_c=e)
    s.set_iterations(3)
    s.evolve(capi.LEAPFROG_KDK | capi.BARNES_HUT)
    s.set_collisions(capi.OVERLAP_COLL_CHECK, mergers=False, min_tot_com_mom=1e10)
    s.evolve(capi.LEAPFROG_KDK | capi.BARNES_HUT)
    st = s.state()
    s.close()
    p1, v1, _ = O.bh_kdk(m, pos, vel, 1.0, 1e-3, 0.01, 0.5, 3)
    rp, rv, _ = O.bh_kdk_coll(m, r, e, p1, v1, 1.0, 1e-3, 0.01, 0.5, 3)
    assert np.array_equal(st["pos"], rp)
    assert np.array_equal(st["vel"], rv)


def test_bh_coll_mergers_flag_only_bounces():
    print("case bh coll mergers", flush=True)
    m, r, e, pos, vel = cluster(22, 4096, 0.03)
    a = run(m, r, e, pos, vel, 3, mergers=False)
    b = run(m, r, e, pos, vel, 3, mergers=True)
    assert len(b["m"]) == len(m)
    assert np.array_equal(a["vel"], b["vel"]) and np.array_equal(a["pos"], b["pos"])


def test_bh_coll_mixed_tracks_fp64():
    print("case bh coll mixed", flush=True)
    m, r, e, pos, vel = cluster(23, 16384, 0.02)
    out = [run(m, r, e, pos, vel, 5, prec=p) for p in (capi.PREC_MIXED, capi.PREC_FP64)]
    dv = out[0]["vel"] - vel
    dv_ref = out[1]["vel"] - vel
    rel = np.linalg.norm(dv - dv_ref, axis=1) / np.linalg.norm(dv_ref, axis=1)
    assert np.median(rel) < 1e-2
    assert np.all(np.isfinite(out[0]["vel"]))


def test_bh_coll_signed_zero_and_one_large_radius():
    """Radii of 0 and -0.0 everywhere but a few bodies of radius 0.2: the
    walk's reach comes from the largest radius, which a -0.0 must not hide."""
    print("case bh coll signed zero", flush=True)
    m, pos, vel = capi.ic_two_plummer(24, 4096)
    r = np.zeros(4096)
    r[1::2] = -0.0
    r[::100] = 0.2
    e = np.full(4096, 0.7)
    one = O.bh_s_coll(m, r, e, pos, vel, dtype=np.float64)
    assert (one != vel).any(axis=1).sum() > 30
    st = run(m, r, e, pos, vel, 3)
    rp, rv, _ = O.bh_kdk_coll(m, r, e, pos, vel, 1.0, 1e-3, 0.01, 0.5, 3)
    assert np.array_equal(st["pos"], rp) and np.array_equal(st["vel"], rv)
