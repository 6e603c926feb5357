"""Generate golden vectors from the COMPILED REFERENCE (oracle/_ref).

Runs only where /root/reference is mounted (the harness is built from its
headers by oracle/Makefile).  Inputs come from the engine's seeded IC
generators; outputs are the reference's own results, stored as .npz:

  ref_kdk_n{N}_s{S}.npz  gtd::system<ld,...>::evolve(leapfrog_kdk), G = SI,
                         dt = 10 s, uniform-ball ICs (seed 0): inputs (double)
                         and final pos/vel (x87 long double).
  ref_tree_plummer_n{N}.npz  bh_tree<double> of a Plummer sphere (seed 5):
                         DFS pre-order depth, octant, body index, node mass.
  ref_coll_{bounce,merge}_seed{K}.npz  system<ld,...,collisions=3> without /
                         with -DGREGSYS_MERGERS: every body after every step.
  ref_integ_{method}_n64_s100.npz  evolve(euler | modified_euler | midpoint |
                         leapfrog_dkd), uniform ball (seed 0), dt = 10 s.
  ref_icoll_{method}_{bounce,merge}_seed1.npz  the same integrators with
                         collisions = 3, rubble cluster: every body after
                         every 25th step, and the final body count.
  ref_track_n{N}_seed{K}.npz  body_tracker over the bodies with x > 0 (uniform
                         ball): count, mass, com, com_vel, mean distance to
                         oracle.TRACK_POINT, mean separation (long double).
  ref_bhcoll_n{N}_seed{K}.npz  one bh_s_coll pass (oracle/ref_bhcoll, the
                         reference's own private method on its own tree) on
                         a dense rubble cluster: inputs and the velocities
                         after the pass (x87 long double).

  ref_dropin_{cor,cor_pred}_seed1.npz  examples/drop_in_driver.cpp built
                         against the reference (oracle/_ref/drop_in_driver_ref):
                         long double system with collisions = 3 whose
                         for_all_bods rewrites every restitution coefficient
                         the way random_comet does; every body after every
                         step (rubble cluster, seed 1, 300 steps).
  ref_dropin_merge_{case}.npz / ref_dropin_check_{case}.npz  the same driver's
                         construction with mergeOverlappingBodies (surviving
                         bodies) and the plain overlap check (throws or not).

Usage: python tests/golden/make_fixtures.py [integ|bhcoll|track|dropin]   (only
       the integrator / bh_s_coll / body_tracker / drop-in driver fixtures)
"""
import subprocess
import os
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle as O  # noqa: E402
from gregpackagescpp_amd import capi  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))

KDK_CASES = [(2, 1), (2, 1000), (3, 10), (64, 10), (64, 1000), (1024, 10), (1024, 1000)]
TREE_CASES = [2, 3, 17, 1000, 16384]
COLL_CASES = [(24, 300, s) for s in range(4)]
INTEG_METHODS = ["euler", "modeuler", "midpoint", "dkd"]
BHCOLL_CASES = [(300, 7, 2.5), (2000, 8, 1.5), (64, 9, 4.0)]  # (n, seed, radius factor)


def coll_ics(n, seed):
    """Rubble cluster: bodies of 1e11-1e13 kg at 500 kg/m^3 in a 12 km box."""
    rng = np.random.default_rng(seed)
    m = 10 ** rng.uniform(11, 13, n)
    r = (3 * m / (4 * np.pi * 500)) ** (1 / 3)
    pos = rng.uniform(-6e3, 6e3, (n, 3))
    vel = rng.normal(0, 0.5, (n, 3))
    e = rng.uniform(0.1, 0.9, n)
    return m, r, e, pos, vel


def integ_fixtures(inp, out):
    for meth in INTEG_METHODS:
        m, pos, vel = capi.ic_uniform_ball(0, 64, 1e20, 1e22, 1e9, 1e3)
        O.write_ref_input(inp, m, pos, vel, steps=100, dt=10.0)
        O.run_ref(meth, inp, out)
        res = O.read_ref_kdk(out, 64)
        np.savez_compressed(os.path.join(OUT, f"ref_integ_{meth}_n64_s100.npz"), m=m, pos=pos, vel=vel, dt=10.0,
                            steps=100, pos_T=res[:, :3], vel_T=res[:, 3:])
        print("integ", meth)
        m, r, e, pos, vel = coll_ics(24, 1)  # seed 1: bounces and three mergers under KDK
        for merge in (False, True):
            O.write_ref_input(inp, m, pos, vel, r=r, e=e, steps=300, dt=5.0)
            O.run_ref("coll_" + meth, inp, out, merge=merge)
            ids0, st, ids, vals = O.read_ref_coll(out)
            tag = "merge" if merge else "bounce"
            keep = st % 25 == 0  # every 25th step's bodies (the full trace is 7200 records)
            np.savez_compressed(os.path.join(OUT, f"ref_icoll_{meth}_{tag}_seed1.npz"), m=m, r=r, e=e, pos=pos,
                                vel=vel, dt=5.0, steps=300, ids0=ids0, step=st[keep], ids=ids[keep],
                                vals=vals[keep], n_final=int((st == st.max()).sum()))
            print("icoll", meth, tag, len(st))


def bhcoll_ics(n, seed, rfac):
    """Rubble cluster with radii inflated by rfac: many overlapping pairs."""
    m, r, e, pos, vel = coll_ics(n, seed)
    return m, r * rfac, e, pos, vel


def bhcoll_fixtures(tmp):
    for n, seed, rfac in BHCOLL_CASES:
        m, r, e, pos, vel = bhcoll_ics(n, seed, rfac)
        v = O.ref_bhcoll(m, r, e, pos, vel, tmp)
        np.savez_compressed(os.path.join(OUT, f"ref_bhcoll_n{n}_seed{seed}.npz"), m=m, r=r, e=e, pos=pos, vel=vel,
                            vel_after=v)
        print("bhcoll", n, seed, int((v != vel.astype(O.LD)).any(axis=1).sum()), "bodies bounced")


TRACK_CASES = [(1024, 0), (3000, 6)]


def track_fixtures(tmp):
    for n, seed in TRACK_CASES:
        m, pos, vel = capi.ic_uniform_ball(seed, n, 1e20, 1e22, 1e9, 1e3)
        cnt, vals = O.ref_track(m, pos, vel, tmp)
        np.savez_compressed(os.path.join(OUT, f"ref_track_n{n}_seed{seed}.npz"), m=m, pos=pos, vel=vel, count=cnt,
                            vals=vals, point=np.array(O.TRACK_POINT))
        print("track", n, seed, cnt)


def overlap_cases():
    """Inputs of the construction-time overlap check: (name, m, r, e, pos, vel)."""
    cases = []
    # a merge whose grown body then reaches a prefix body (q <= o - 2): the
    # reference erases the body after q, shifts the outer index down, and
    # erases every body after the outer through self-additions
    m = np.array([1.0, 1.0, 1.0, 1.0, 2.0, 3.0])
    r = np.ones(6)
    pos = np.array([[0.0, 0, 0], [100.0, 0, 0], [2.5, 0, 0], [2.0, 0, 0], [50.0, 0, 0], [70.0, 0, 0]])
    vel = np.arange(18, dtype=float).reshape(6, 3) * 0.1
    cases.append(("prefix", m, r, np.full(6, 0.5), pos, vel))
    # an outer of zero radius: its self-test fails, so the loop keeps testing
    # (one element behind) the bodies after it
    m = np.array([1.0, 2.0, 1.0, 1.0, 1.0, 1.0, 4.0])
    r = np.array([1.0, 0.0, 1.0, 1.0, 1.0, 1.0, 1.0])
    pos = np.array([[0.0, 0, 0], [10.0, 0, 0], [10.5, 0, 0], [30.0, 0, 0], [31.0, 0, 0], [60.0, 0, 0],
                    [61.5, 0, 0]])
    vel = np.arange(21, dtype=float).reshape(7, 3) * -0.2
    cases.append(("zero_radius", m, r, np.linspace(0.2, 0.8, 7), pos, vel))
    # rubble clusters with inflated radii (several overlapping pairs) and one
    # without any overlap
    for seed, rfac, n in ((11, 0.3, 60), (12, 0.3, 60), (20, 0.5, 60), (26, 0.3, 60), (13, 3.0, 200)):
        m, r, e, pos, vel = coll_ics(n, seed)
        cases.append((f"rubble{seed}", m, r * rfac, e, pos, vel))
    return cases


def dropin_fixtures(tmp):
    exe = os.path.join(O.REF_DIR, "drop_in_driver_ref")
    inp, out = os.path.join(tmp, "d_in.bin"), os.path.join(tmp, "d_out.bin")
    m, r, e, pos, vel = coll_ics(24, 1)
    for mode in ("cor", "cor_pred"):
        O.write_ref_input(inp, m, pos, vel, r=r, e=e, steps=300, dt=5.0)
        subprocess.run([exe, mode, inp, out], check=True, timeout=600)
        rec = np.fromfile(out, dtype=np.float64).reshape(-1, 11)
        rec = rec[rec[:, 0] % 25 == 0]  # every body at every 25th step
        np.savez_compressed(os.path.join(OUT, f"ref_dropin_{mode}_seed1.npz"), m=m, r=r, e=e, pos=pos, vel=vel,
                            dt=5.0, steps=300, rec=rec)
        print("dropin", mode, len(rec))
    for name, m, r, e, pos, vel in overlap_cases():
        O.write_ref_input(inp, m, pos, vel, r=r, e=e, steps=1, dt=1.0)
        subprocess.run([exe, "merge", inp, out], check=True, timeout=600)
        res = np.fromfile(out, dtype=np.float64)
        k = int(res[0])
        np.savez_compressed(os.path.join(OUT, f"ref_dropin_merge_{name}.npz"), m=m, r=r, e=e, pos=pos, vel=vel,
                            bodies=res[1:].reshape(k, 9))
        subprocess.run([exe, "check", inp, out], check=True, timeout=600, capture_output=True)
        flag = np.fromfile(out, dtype=np.float64)[0]
        np.savez_compressed(os.path.join(OUT, f"ref_dropin_check_{name}.npz"), m=m, r=r, e=e, pos=pos, vel=vel,
                            throws=flag)
        print("dropin merge", name, len(m), "->", k, "check throws", flag)


def main():
    if not O.ref_available():
        sys.exit("oracle/_ref is not built (needs /root/reference): nothing generated")
    tmp = tempfile.mkdtemp()
    inp, out = os.path.join(tmp, "in.bin"), os.path.join(tmp, "out.bin")
    if sys.argv[1:] == ["integ"]:
        integ_fixtures(inp, out)
        return
    if sys.argv[1:] == ["bhcoll"]:
        bhcoll_fixtures(tmp)
        return
    if sys.argv[1:] == ["track"]:
        track_fixtures(tmp)
        return
    if sys.argv[1:] == ["dropin"]:
        dropin_fixtures(tmp)
        return
    bhcoll_fixtures(tmp)
    track_fixtures(tmp)
    dropin_fixtures(tmp)
    integ_fixtures(inp, out)
    for n, steps in KDK_CASES:
        m, pos, vel = capi.ic_uniform_ball(0, n, 1e20, 1e22, 1e9, 1e3)
        O.write_ref_input(inp, m, pos, vel, steps=steps, dt=10.0)
        O.run_ref("kdk", inp, out)
        res = O.read_ref_kdk(out, n)
        np.savez_compressed(os.path.join(OUT, f"ref_kdk_n{n}_s{steps}.npz"), m=m, pos=pos, vel=vel, dt=10.0, steps=steps,
                 pos_T=res[:, :3], vel_T=res[:, 3:])
        print("kdk", n, steps)
    for n in TREE_CASES:
        m, pos, vel = capi.ic_plummer(5, n)
        O.write_ref_input(inp, m, pos, vel)
        O.run_ref("tree", inp, out)
        rec = O.read_ref_tree(out)
        np.savez_compressed(os.path.join(OUT, f"ref_tree_plummer_n{n}.npz"), m=m, pos=pos, depth=rec["d"], octant=rec["o"],
                 body=rec["b"], mass=rec["m"])
        print("tree", n, len(rec))
    for n, steps, seed in COLL_CASES:
        m, r, e, pos, vel = coll_ics(n, seed)
        for merge in (False, True):
            O.write_ref_input(inp, m, pos, vel, r=r, e=e, steps=steps, dt=5.0)
            O.run_ref("coll", inp, out, merge=merge)
            ids0, st, ids, vals = O.read_ref_coll(out)
            tag = "merge" if merge else "bounce"
            np.savez_compressed(os.path.join(OUT, f"ref_coll_{tag}_seed{seed}.npz"), m=m, r=r, e=e, pos=pos, vel=vel, dt=5.0,
                     steps=steps, ids0=ids0, step=st, ids=ids, vals=vals)
            print("coll", tag, seed, len(st))


if __name__ == "__main__":
    main()
