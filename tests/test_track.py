"""body_tracker reductions (gregsys.hpp:3492-3611): the oracle on the CPU.

Pins: the long-double oracle reproduces the compiled reference's tracker bit
for bit (mass, COM, COM velocity, mean distance to a point, mean separation;
fixtures ref_track_*.npz, bodies with x > 0 of uniform balls); the double form
(rows of the separation summed first, the GPU's association) stays within
1e-14 of it; closed forms: a single body has zero separation and its own
position as COM, and two bodies' separation is their distance.
"""
import glob
import os

import numpy as np
import pytest

import oracle as O


@pytest.mark.parametrize("name", ["ref_track_n1024_seed0", "ref_track_n3000_seed6"])
def test_oracle_ld_matches_reference(golden, name):
    f = np.load(os.path.join(golden, name + ".npz"))
    idx = np.flatnonzero(f["pos"][:, 0] > 0)
    assert len(idx) == int(f["count"])
    t = O.track(f["m"], f["pos"], f["vel"], idx, f["point"], dtype=O.LD)
    got = np.concatenate([[t["mass"]], t["com"], t["com_vel"], [t["mean_dist"], t["mean_sep"]]])
    assert np.array_equal(got, f["vals"])
    d = O.track(f["m"], f["pos"], f["vel"], idx, f["point"], dtype=np.float64)
    dv = np.concatenate([[d["mass"]], d["com"], d["com_vel"], [d["mean_dist"], d["mean_sep"]]])
    ref = f["vals"].astype(np.float64)
    assert np.all(np.abs(dv - ref) <= 1e-14 * np.abs(ref) + 1e-300)


def test_closed_forms():
    m = np.array([2.0, 3.0])
    pos = np.array([[1.0, 2.0, 3.0], [4.0, 6.0, 3.0]])
    vel = np.array([[1.0, 0.0, 0.0], [0.0, 1.0, 0.0]])
    one = O.track(m, pos, vel, [1], (0.0, 0.0, 0.0), dtype=np.float64)
    assert one["mean_sep"] == 0 and np.array_equal(one["com"], pos[1]) and one["mass"] == 3.0
    assert one["mean_dist"] == np.sqrt(16 + 36 + 9)
    two = O.track(m, pos, vel, [0, 1], (0.0, 0.0, 0.0), dtype=np.float64)
    assert two["mean_sep"] == 5.0
    assert np.allclose(two["com"], (2 * pos[0] + 3 * pos[1]) / 5, rtol=0, atol=1e-15)
    assert np.allclose(two["com_vel"], [0.4, 0.6, 0.0], rtol=0, atol=1e-15)
