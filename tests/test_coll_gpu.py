"""Collisions (s_coll bounce and merge) on the GPU, single systems and batched
ensembles, against the oracle and the compiled reference's fixtures.

Bars:
  * FP64 path: states and merge records bit-identical to the double
    restatement (oracle *_d), which uses the reference's operation order and
    the same fixed Newton cube root for merged radii;
  * FP64 path vs the compiled reference's long-double fixtures: the same
    bodies survive, merged in the same steps, states within 1e-9 relative;
  * batched ensemble (one workgroup per system, many steps per launch):
    FP64 batch bit-identical to the oracle per system (also over the full
    20,000-step comet workload for a stratified sample of 16 systems, n from
    8 to 512 fragments: merge log, ids, m, r, x, v); the mixed (FP32 pair
    forces) batch reproduces >= 98% of the FP64 batch's (system, outer,
    absorbed) merge records and its merge count within 0.1% -- FP32 forces
    shift the contact step of slow fragment approaches, so merge lists are
    bit-exact only on the FP64 path.
"""
import glob
import os

import numpy as np
import pytest

import oracle as O
from gregpackagescpp_amd import capi

pytestmark = pytest.mark.gpu
G_SI = float(np.longdouble(66743) / np.longdouble(10 ** 15))


def rubble(n, seed):
    rng = np.random.default_rng(seed)
    m = 10 ** rng.uniform(11, 13, n)
    r = (3 * m / (4 * np.pi * 500)) ** (1 / 3)
    pos = rng.uniform(-6e3, 6e3, (n, 3))
    vel = rng.normal(0, 0.5, (n, 3))
    e = rng.uniform(0.1, 0.9, n)
    return m, r, e, pos, vel


def run_single(m, r, e, pos, vel, steps, merge, prec=capi.PREC_FP64, dt=5.0, method=capi.LEAPFROG_KDK):
    s = capi.System(0)
    s.set_G(G_SI)
    s.set_timestep(dt)
    s.set_iterations(steps)
    s.set_precision(prec)
    s.set_collisions(capi.OVERLAP_COLL_CHECK, mergers=merge, min_tot_com_mom=1e10)
    s.set_bodies(m, pos, vel, r=r, rest_c=e)
    s.evolve(method)
    return s.state(), s.ids(), s.merge_log()


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
@pytest.mark.parametrize("merge", [False, True])
def test_single_bit_exact_vs_double_oracle(seed, merge):
    print(f"case single seed={seed} merge={merge}", flush=True)
    m, r, e, pos, vel = rubble(24, seed)
    st, ids, log = run_single(m, r, e, pos, vel, 300, merge)
    ref = O.kdk(m, pos, vel, G_SI, 5.0, 300, coll=2 if merge else 1, r=r, e=e, dtype=np.float64)
    assert len(ids) == ref["n"]
    assert np.array_equal(ids, ref["labels"])
    for k_gpu, k_ref in (("m", "m"), ("r", "r"), ("pos", "pos"), ("vel", "vel")):
        assert np.array_equal(st[k_gpu], ref[k_ref]), k_gpu
    assert np.array_equal(log, ref["log"])


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden",
                                                               "ref_coll_*.npz"))), ids=os.path.basename)
def test_single_vs_reference_fixture(path):
    f = np.load(path)
    merge = "merge" in os.path.basename(path)
    steps = int(f["steps"])
    st, ids, log = run_single(f["m"], f["r"], f["e"], f["pos"], f["vel"], steps, merge)
    last = f["step"] == steps
    vals = f["vals"][last].astype(np.float64)
    assert len(ids) == vals.shape[0]
    assert np.abs(st["m"] - vals[:, 0]).max() / vals[:, 0].max() < 1e-12
    scale = np.abs(vals[:, 2:5]).max()
    assert np.abs(st["pos"] - vals[:, 2:5]).max() / scale < 1e-9
    vs = np.abs(vals[:, 5:8]).max()
    assert np.abs(st["vel"] - vals[:, 5:8]).max() / vs < 1e-7


def make_batch(systems, prec, merge=True, dt=5.0, G=G_SI):
    offsets = np.concatenate([[0], np.cumsum([len(s[0]) for s in systems])]).astype(np.int64)
    cat = [np.concatenate([s[k] for s in systems]) for k in range(5)]
    m, r, e, pos, vel = cat
    b = capi.Batch(offsets, m, r, pos, vel, e)
    b.set_params(G, dt, 0.0, capi.OVERLAP_COLL_CHECK, merge, 1e10, prec)
    return b, offsets


def test_batch_fp64_bit_exact_vs_oracle():
    systems = [rubble(n, 100 + k) for k, n in enumerate([8, 13, 24, 40, 24, 64, 9, 31])]
    b, offsets = make_batch(systems, capi.PREC_FP64)
    b.evolve(400)
    st = b.state()
    log = b.merge_log()
    k = 0
    for s, (m, r, e, pos, vel) in enumerate(systems):
        ref = O.kdk(m, pos, vel, G_SI, 5.0, 400, coll=2, r=r, e=e, dtype=np.float64)
        c = int(st["counts"][s])
        assert c == ref["n"]
        assert np.array_equal(st["pos"][k:k + c], ref["pos"])
        assert np.array_equal(st["vel"][k:k + c], ref["vel"])
        assert np.array_equal(st["m"][k:k + c], ref["m"]) and np.array_equal(st["r"][k:k + c], ref["r"])
        assert np.array_equal(st["ids"][k:k + c], ref["labels"])
        assert np.array_equal(log[log[:, 0] == s][:, 1:], ref["log"])
        k += c
    assert len(log) > 0


@pytest.mark.parametrize("sizes", [[300, 520, 9], [800, 64], [1, 2, 65, 128, 129, 257, 3]])
def test_batch_fp64_large_systems_bit_exact(sizes):
    """Systems whose lanes own two or three 64-body blocks (up to 768 bodies)
    and, with an 800-body system in the batch, the four-block kernel; sizes at
    the edges (1 and 2 bodies, one past a block): states, ids and merge
    records bit-identical to the double restatement."""
    print(f"case batch fp64 sizes={sizes}", flush=True)
    systems = [rubble(n, 300 + k) for k, n in enumerate(sizes)]
    b, _ = make_batch(systems, capi.PREC_FP64)
    b.evolve(12)
    st = b.state()
    log = b.merge_log()
    k = 0
    for s, (m, r, e, pos, vel) in enumerate(systems):
        ref = O.kdk(m, pos, vel, G_SI, 5.0, 12, coll=2, r=r, e=e, dtype=np.float64)
        c = int(st["counts"][s])
        assert c == ref["n"]
        assert np.array_equal(st["pos"][k:k + c], ref["pos"])
        assert np.array_equal(st["vel"][k:k + c], ref["vel"])
        assert np.array_equal(st["m"][k:k + c], ref["m"]) and np.array_equal(st["r"][k:k + c], ref["r"])
        assert np.array_equal(st["ids"][k:k + c], ref["labels"])
        assert np.array_equal(log[log[:, 0] == s][:, 1:], ref["log"])
        k += c
    assert len(log) > 0


def test_batch_matches_single_system_api():
    systems = [rubble(n, 200 + k) for k, n in enumerate([12, 30])]
    b, _ = make_batch(systems, capi.PREC_FP64)
    b.evolve(250)
    st = b.state()
    k = 0
    for s, (m, r, e, pos, vel) in enumerate(systems):
        one, ids, _ = run_single(m, r, e, pos, vel, 250, True)
        c = int(st["counts"][s])
        assert np.array_equal(st["pos"][k:k + c], one["pos"])
        k += c


COMET_STEPS = 20000


@pytest.fixture(scope="module")
def comet_runs():
    """Config 3: 4096 comet-string systems (seed 3; central body + 8..512
    fragments), dt = 5 s, 20,000 steps, merging on, run by the batch in the
    mixed precision (the one the benchmark times) and in FP64."""
    offsets, m, r, pos, vel, e = capi.ic_comet_ensemble(3, 4096, 8, 512)
    logs, states = {}, {}
    for prec in (capi.PREC_MIXED, capi.PREC_FP64):
        print(f"case comet ensemble prec={prec}", flush=True)
        b = capi.Batch(offsets, m, r, pos, vel, e)
        b.set_params(G_SI, 5.0, 0.0, capi.OVERLAP_COLL_CHECK, True, 1e10, prec)
        ns = b.evolve(COMET_STEPS)
        print(f"  {ns / 1e9:.2f} s", flush=True)
        logs[prec] = b.merge_log()
        states[prec] = b.state()
        b.close()
    return dict(ics=(offsets, m, r, pos, vel, e), logs=logs, states=states)


def test_comet_ensemble_mixed_merge_list_matches_fp64(comet_runs):
    """The mixed batch's merge list against the FP64 batch's: the same events
    (who absorbs whom) to >= 98% and the same merge count to 0.1%."""
    logs = comet_runs["logs"]
    a, b = logs[capi.PREC_MIXED], logs[capi.PREC_FP64]
    same = sum(np.array_equal(a[a[:, 0] == s], b[b[:, 0] == s]) for s in range(4096))
    pairs_a = {tuple(r) for r in a[:, [0, 2, 3]]}
    pairs_b = {tuple(r) for r in b[:, [0, 2, 3]]}
    common = len(pairs_a & pairs_b) / max(len(pairs_b), 1)
    print(f"merges: fp64 {len(b)}, mixed {len(a)}; systems with identical lists {same}/4096; "
          f"(system, outer, absorbed) records in common {common:.5f}")
    assert len(b) > 0
    # Most merges here close slow (cm/s) gravitational approaches between
    # fragments, whose contact step moves when the pair force carries FP32
    # rounding; the events themselves (who absorbs whom) must still agree.
    assert abs(len(a) - len(b)) <= 1e-3 * len(b)
    assert common >= 0.98


def comet_sample(sizes, k=16):
    """k systems at evenly spaced ranks of the size order: the smallest, the
    median region and the largest (n = 512 fragments) are all included."""
    order = np.argsort(sizes, kind="stable")
    ranks = np.unique(np.round(np.linspace(0, len(sizes) - 1, k)).astype(int))
    return [int(order[q]) for q in ranks]


def test_comet_ensemble_fp64_bit_exact_vs_oracle(comet_runs):
    """The FP64 batch against the double restatement over all 20,000 steps,
    for a stratified sample of 16 systems (smallest to largest n): the full
    merge log (step, outer, absorbed, merged labels), the surviving ids and
    every m, r, position and velocity, bit for bit."""
    from concurrent.futures import ThreadPoolExecutor

    offsets, m, r, pos, vel, e = comet_runs["ics"]
    sizes = np.diff(offsets)
    pick = comet_sample(sizes)
    assert sizes[pick].max() == sizes.max() and sizes[pick].min() == sizes.min()
    st = comet_runs["states"][capi.PREC_FP64]
    log = comet_runs["logs"][capi.PREC_FP64]
    starts = np.concatenate([[0], np.cumsum(st["counts"])])

    def oracle_run(s):
        a, bnd = offsets[s], offsets[s + 1]
        return O.kdk(m[a:bnd], pos[a:bnd], vel[a:bnd], G_SI, 5.0, COMET_STEPS, coll=2, r=r[a:bnd], e=e[a:bnd],
                     dtype=np.float64)

    print(f"case comet fp64 vs oracle systems={pick}", flush=True)
    with ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 4)) as ex:
        refs = list(ex.map(oracle_run, pick))
    merges = 0
    for s, ref in zip(pick, refs):
        k0, k1 = starts[s], starts[s + 1]
        print(f"  system {s}: n0={sizes[s]} survivors={ref['n']} merges={len(ref['log'])}", flush=True)
        assert k1 - k0 == ref["n"]
        assert np.array_equal(st["ids"][k0:k1], ref["labels"])
        for key in ("m", "r", "pos", "vel"):
            assert np.array_equal(st[key][k0:k1], ref[key]), (s, key)
        assert np.array_equal(log[log[:, 0] == s][:, 1:], ref["log"]), s
        merges += len(ref["log"])
    assert merges > 0


@pytest.mark.parametrize("method", [capi.EULER, capi.MODIFIED_EULER, capi.MIDPOINT, capi.LEAPFROG_DKD])
@pytest.mark.parametrize("merge", [False, True])
def test_other_integrators_collisions_bit_exact(method, merge):
    """Euler, modified Euler, midpoint and leapfrog DKD with s_coll after every
    step (rubble cluster, seed 1: bounces and, with mergers, three merges)
    against the oracle's double restatement (pinned to the reference's
    ref_icoll_* fixtures in long double)."""
    print(f"case other integrators coll method={method} merge={merge}", flush=True)
    m, r, e, pos, vel = rubble(24, 1)
    st, ids, log = run_single(m, r, e, pos, vel, 300, merge, method=method)
    ref = O.kdk(m, pos, vel, G_SI, 5.0, 300, coll=2 if merge else 1, r=r, e=e, dtype=np.float64, method=method)
    assert len(ids) == ref["n"]
    assert np.array_equal(ids, ref["labels"])
    for k in ("m", "r", "pos", "vel"):
        assert np.array_equal(st[k], ref[k]), k
    assert np.array_equal(log, ref["log"])
