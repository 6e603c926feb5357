"""body_tracker reductions on the GPU (nb_track), against the oracle.

Bar: all nine outputs (mass, COM, COM velocity, mean distance to a point,
mean separation) bit-identical to the double oracle, on the initial state and
after 20 FP64 KDK steps, for a subset picked by a predicate (x > 0) and for
the whole system; an empty subset asking for the separation is an error, as
the reference's mean_sep() throws.
"""
import numpy as np
import pytest

import oracle as O
from gregpackagescpp_amd import capi

pytestmark = pytest.mark.gpu
G_SI = float(np.longdouble(66743) / np.longdouble(10 ** 15))


def vals(t):
    return np.concatenate([[t["mass"]], t["com"], t["com_vel"], [t["mean_dist"], t["mean_sep"]]])


@pytest.mark.parametrize("n", [1, 2, 1000, 20000])
def test_track_bit_exact(n):
    print(f"case track n={n}", flush=True)
    m, pos, vel = capi.ic_uniform_ball(41, n, 1e20, 1e22, 1e9, 1e3)
    s = capi.System(0)
    s.set_G(G_SI)
    s.set_timestep(10.0)
    s.set_iterations(20)
    s.set_bodies(m, pos, vel)
    point = np.array(O.TRACK_POINT)
    for evolved in (False, True):
        if evolved:
            s.evolve(capi.LEAPFROG_KDK)
        st = s.state()
        for idx in (np.flatnonzero(st["pos"][:, 0] > 0), np.arange(n)):
            if len(idx) == 0:
                continue
            got = vals(s.track(idx, point))
            ref = vals(O.track(st["m"], st["pos"], st["vel"], idx, point, dtype=np.float64))
            assert np.array_equal(got, ref)


def test_track_empty_subset():
    m, pos, vel = capi.ic_uniform_ball(42, 10, 1e20, 1e22, 1e9, 1e3)
    s = capi.System(0)
    s.set_bodies(m, pos, vel)
    with pytest.raises(capi.NBodyError):
        s.track(np.zeros(0, np.int64))
    with pytest.raises(capi.NBodyError):
        s.track(np.array([0, 10]))  # slot 10 of a 10-body system
