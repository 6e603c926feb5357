"""Generate .nsys snapshot fixtures with the COMPILED REFERENCE (oracle/_ref).

Runs only where /root/reference is mounted.  The reference's own writer,
system<T,T,T>::to_nsys (gregsys.hpp:3214-3258), snapshots a 5-body system
built from the engine's seeded generator (uniform ball, seed 41; radii
1e3*(i+1) m, restitution 0.5+0.1*i; dt = 10 s; units "M3:2,D5:7,T11:13"):

  nsys_ld_n5_t0.nsys  T = long double, not evolved (time elapsed 0)
  nsys_ld_n5_t3.nsys  T = long double, after evolve(leapfrog_kdk) of 3 steps
  nsys_d_n5_t0.nsys   T = double, not evolved

Usage: python tests/golden/make_nsys_fixtures.py
"""
import os
import shutil
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle as O  # noqa: E402
from gregpackagescpp_amd import capi  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))
N, DT = 5, 10.0


def bodies():
    m, pos, vel = capi.ic_uniform_ball(41, N, 1e20, 1e22, 1e9, 1e3)
    r = 1e3 * np.arange(1, N + 1, dtype=np.float64)
    e = 0.5 + 0.1 * np.arange(N, dtype=np.float64)
    return m, r, e, pos, vel


def main():
    if not O.ref_available():
        sys.exit("oracle/_ref is not built (needs /root/reference): nothing generated")
    tmp = tempfile.mkdtemp()
    inp = os.path.join(tmp, "in.bin")
    m, r, e, pos, vel = bodies()
    for mode, steps, name in (("nsys_ld", 0, "nsys_ld_n5_t0"), ("nsys_ld", 3, "nsys_ld_n5_t3"),
                              ("nsys_d", 0, "nsys_d_n5_t0")):
        O.write_ref_input(inp, m, pos, vel, r=r, e=e, steps=steps, dt=DT)
        out = os.path.join(tmp, name + ".nsys")
        O.run_ref(mode, inp, out)
        shutil.copyfile(out, os.path.join(OUT, name + ".nsys"))
        print(name, os.path.getsize(out), "bytes")
    shutil.rmtree(tmp)


if __name__ == "__main__":
    main()
