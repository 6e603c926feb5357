"""Contexts on separate host threads at the same time (one context per host
thread, each with its own stream): results equal those of the same runs one
after another, bit for bit."""
import threading

import numpy as np
import pytest

from gregpackagescpp_amd import capi

pytestmark = pytest.mark.gpu
G_SI = float(np.longdouble(66743) / np.longdouble(10 ** 15))


def direct_run(seed, out, key):
    m, pos, vel = capi.ic_uniform_ball(seed, 2048, 1e20, 1e22, 1e9, 1e3)
    s = capi.System(0)
    s.set_G(G_SI)
    s.set_timestep(10.0)
    s.set_iterations(50)
    s.set_bodies(m, pos, vel)
    s.evolve(capi.LEAPFROG_KDK)
    out[key] = s.state()["pos"]
    s.close()


def bh_run(seed, out, key):
    m, pos, vel = capi.ic_two_plummer(seed, 65536)
    s = capi.System(0)
    s.set_G(1.0)
    s.set_softening(0.01)
    s.set_bh_opening_angle(0.5)
    s.set_timestep(1e-3)
    s.set_iterations(5)
    s.set_precision(capi.PREC_MIXED)
    s.set_bodies(m, pos, vel)
    s.evolve(capi.LEAPFROG_KDK | capi.BARNES_HUT)
    out[key] = s.state()["pos"]
    s.close()


def test_threads_match_serial():
    print("case concurrency", flush=True)
    jobs = [(direct_run, 51, "d1"), (direct_run, 52, "d2"), (bh_run, 53, "b1")]
    serial, par = {}, {}
    for f, seed, key in jobs:
        f(seed, serial, key)
    threads = [threading.Thread(target=f, args=(seed, par, key)) for f, seed, key in jobs]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    assert all(not t.is_alive() for t in threads)
    for _, _, key in jobs:
        assert np.array_equal(serial[key], par[key]), key
