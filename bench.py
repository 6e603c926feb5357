"""Benchmark: direct-summation KDK step, Plummer N = 65,536 (mixed precision).

python bench.py [--gpus 1] [--steps K] [--warmup W]

A step is one leapfrog-KDK step of the whole system (force pass + FP64
kick/drift) on one MI355X, inputs resident in HBM.  value = interactions/s =
N(N-1)*K / t over the K timed steps.  The JSON line also carries:
  roofline     dominant kernel (the FP32 pair-force kernel): achieved FP32
               FLOP/s at 20 FLOP per interaction (N(N-1) per launch) over its
               average launch duration, timed with HIP events on the
               library's stream;
  cpu_baseline the compiled reference's own evolve(leapfrog_kdk) (oracle/_ref,
               long double, one thread) on a bounded sample, else the oracle;
  cpu_baseline_allcore  an FP64 restatement over all host threads (OpenMP);
  bh           Barnes-Hut KDK at 1,048,576 bodies (I_alg interactions/s), with
               its own single-thread and all-core CPU baselines;
  adjoint      backward/forward wall-time ratio of the adjoint at N = 4096;
  batch        4096-system comet ensemble with merging (interactions/s), the
               mixed batch and, beside it, the FP64 batch (bit-exact merge
               list), each with its kernel roofline.
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N_BODIES = 65536
EPS, DT = 0.01, 1e-3
FLOP_PER_INTERACTION = 20
PEAK_FP32_TFLOPS = 157.3
# rocprofv3 FETCH_SIZE / WRITE_SIZE (KiB per dispatch) of the FP32 force kernel
# at 65,536 bodies per launch configuration (direct_kernel, direct_ipl,
# direct_waves), from `bench.py --profile-direct --kernel K --ipl I --waves W`;
# the counters are scaled by what they read for a copy of known size
# (`bench.py --hbm-copy`): FETCH_SIZE x 2.0, WRITE_SIZE x 1.0.
PROFILED_FORCE_TRAFFIC = {
    (2, 1, 16): (4115.8, 24576.0),
    (2, 2, 16): (4123.2, 49152.0),
    (2, 4, 16): (4145.3, 98304.0),
    (1, 2, 16): (7715.4, 49152.0),
    (1, 4, 16): (7827.2, 98304.0),
}
FETCH_FACTOR, WRITE_FACTOR = 2.0, 1.0


def cpu_baseline(m, pos, vel, allow_port=False):
    """Time the reference CPU path on a bounded sample of the same workload:
    the compiled reference (oracle/_ref, built from the reference's sources
    beside this repository).  Without it the run stops, unless allow_port
    asks for the oracle's restatement instead (labelled kind "port")."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    ns_sample, steps = 8192, 2
    mm, pp, vv = m[:ns_sample].copy(), pos[:ns_sample].copy(), vel[:ns_sample].copy()
    pairs_per_pass = ns_sample * (ns_sample - 1)
    if not O.ref_available() and not allow_port:
        raise SystemExit("bench.py: oracle/_ref/ref_harness (the compiled reference) is missing; build it with "
                         "__graft_entry__.build() where the reference sources are mounted, or pass "
                         "--cpu-baseline-port to time the oracle's restatement instead")
    if O.ref_available():
        with tempfile.TemporaryDirectory() as td:
            inp, out = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
            O.write_ref_input(inp, mm, pp, vv, steps=steps, dt=DT)
            exe = os.path.join(O.REF_DIR, "ref_harness")
            res = subprocess.run([exe, "kdk", inp, out], check=True, capture_output=True, text=True, timeout=600)
            ns = int(res.stdout.split("evolve_ns")[1].split()[0])
        t = ns * 1e-9
        kind = "reference"
        what = "compiled reference gtd::system<long double>::evolve(leapfrog_kdk)"
    else:
        t0 = time.perf_counter()
        O.kdk(mm, pp, vv, 1.0, DT, steps, eps=EPS, form=O.FORM_SOFT)
        t = time.perf_counter() - t0
        kind = "port"
        what = "oracle long-double restatement of the reference KDK"
    # evolve does steps + 1 force passes (initial a(x0)); each pass is N(N-1) interactions
    value = (steps + 1) * pairs_per_pass / t
    return {"value": value, "unit": "interactions/s", "cores": 1, "kind": kind,
            "sample": f"{what}, 1 thread, first {ns_sample} bodies of the N={N_BODIES} Plummer IC, {steps} steps "
                      f"({steps + 1} force passes, {t:.1f} s)"}


def hbm_copy():
    """Plain device copy of known size (256 MiB read + 256 MiB written per
    launch, 3 launches): the calibration run for FETCH_SIZE / WRITE_SIZE."""
    import torch

    n = 64 * 1024 * 1024
    x = torch.ones(n, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"hbm_copy_bytes_per_launch": 4 * n, "launches": 3}))


BH_BODIES = 1 << 20
BH_THETA = 0.5


PEAK_HBM_GBS = 8000.0
PEAK_FP64_TFLOPS = 78.6
# rocprofv3 FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0 summed over the build kernels
# of one BH step at 1,048,576 bodies (profiles/bh_walk_build_two_plummer_1M.md),
# and the same for the frontier walk launch; None when not profiled.
PROFILED_BH_BUILD_BYTES = 1.101e9
PROFILED_BH_WALK_BYTES = 1.935e8
# the same for one 1000-step launch of the mixed batch kernel on the 4096-system
# comet ensemble (profiles/batch_comet_ensemble4096.md)
PROFILED_BATCH_LAUNCH_BYTES = 2.535e8


def build_bytes(n, nodes, passes, kicks=0.0):
    """Algorithmic HBM bytes of one octree build of a KDK step: the step's
    drift, which the build's first pass applies (x and v read, x written:
    72 B per body), and the previous step's kick in a fraction `kicks` of
    the builds (a read, v written: 48 B per body); per body 32 B read (x, y,
    z, m) and 12 B written (key, index); per radix pass 24 B per body (key
    and index read and written); 48 B per node written; 32 B per body for
    the reorder of the bodies into tree order."""
    return n * (72 + 48 * kicks) + n * (32 + 12) + passes * n * 24 + nodes * 48 + n * 32


def bench_bh(capi, steps):
    """Barnes-Hut KDK on two 524,288-body Plummer spheres (theta = 0.5, eps =
    0.01, dt = 1e-3, full rebuild every step).  Interactions are I_alg: the
    cells accepted under the reference's per-body MAC, counted by the FP64
    per-lane walk on the same positions.  Rooflines: the build against HBM
    (algorithmic bytes over its HIP-event time), the walk against FP32 VALU
    at 20 FLOP per I_alg interaction."""
    m, pos, vel = capi.ic_two_plummer(2, BH_BODIES)
    s = capi.System(0)
    s.set_G(1.0)
    s.set_timestep(DT)
    s.set_softening(EPS)
    s.set_bh_opening_angle(BH_THETA)
    s.set_bodies(m, pos, vel)
    s.set_precision(capi.PREC_FP64)
    _, i_alg = s.bh_accelerations()  # reference per-body MAC count on the initial positions
    s.set_precision(capi.PREC_MIXED)
    s.autotune(capi.LEAPFROG_KDK | capi.BARNES_HUT)  # project autotuner: bucket size of the FP32 walk
    s.set_iterations(1)
    s.evolve(capi.LEAPFROG_KDK | capi.BARNES_HUT)  # warm-up
    t0 = time.perf_counter()
    s.advance(steps)
    t = (time.perf_counter() - t0) / steps
    s.profile(True)
    s.advance(2)  # two builds: the second one also applies the first step's kick
    b_ms, nb = s.profile_read("bh_build")
    w_ms, nw = s.profile_read("bh_walk")
    s.profile(False)
    _, walk_cnt = s.bh_accelerations()
    nodes = s.tree_count()
    passes = s.tuning("bh_sort_passes")  # 9-bit radix passes the engine's sort takes in these builds
    bucket = s.tuning("bh_bucket")
    s.close()
    build_ms, walk_ms = b_ms / max(nb, 1), w_ms / max(nw, 1)
    bb = build_bytes(BH_BODIES, nodes, passes, kicks=0.5)
    build_gbs = bb / (build_ms * 1e-3) / 1e9
    walk_tf = FLOP_PER_INTERACTION * i_alg / (walk_ms * 1e-3) / 1e12
    return {"value": i_alg / t, "unit": "interactions/s", "bodies": BH_BODIES, "theta": BH_THETA,
            "ms_per_step": t * 1e3, "interactions_per_step": i_alg, "walk_interactions": walk_cnt,
            "tuned": {"bh_bucket": bucket}, "nodes": nodes,
            "build_ms": build_ms, "walk_ms": walk_ms, "steps": steps,
            "build_roofline": {"bound": "hbm", "achieved": build_gbs, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                               "frac": build_gbs / PEAK_HBM_GBS, "algorithmic_bytes": bb,
                               "radix_passes": passes, "traffic": PROFILED_BH_BUILD_BYTES},
            "walk_roofline": {"bound": "valu", "achieved": walk_tf, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                              "frac": walk_tf / PEAK_FP32_TFLOPS, "flop_per_interaction": FLOP_PER_INTERACTION,
                              "evaluated_tflops": FLOP_PER_INTERACTION * walk_cnt / (walk_ms * 1e-3) / 1e12,
                              "traffic": PROFILED_BH_WALK_BYTES},
            "workload": "BH leapfrog KDK, two Plummer spheres of 524288 at x=+-10, vrel=0.5, theta=0.5, "
                        "eps=0.01, dt=1e-3, FP32 frontier walk, FP64 state and node COM"}


def host_info():
    """The host the CPU baselines ran on: CPU model and logical CPU count."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count()}


def bench_bh_coll(capi, steps, radius):
    """The BH step with bh_s_coll (touching-cell overlaps, bounces) on the same
    1M-body two-Plummer workload with every radius `radius` (0.005: about 48k
    bodies bounce in the first pass; 0.002: about 3k): the pass's own time per
    step from HIP events, and the whole step."""
    m, pos, vel = capi.ic_two_plummer(2, BH_BODIES)
    r = np.full(BH_BODIES, radius)
    e = np.full(BH_BODIES, 0.5)
    s = capi.System(0)
    s.set_G(1.0)
    s.set_timestep(DT)
    s.set_softening(EPS)
    s.set_bh_opening_angle(BH_THETA)
    s.set_precision(capi.PREC_MIXED)
    s.set_collisions(capi.OVERLAP_COLL_CHECK, mergers=False, min_tot_com_mom=1e10)
    s.set_bodies(m, pos, vel, r=r, rest_c=e)
    s.set_iterations(1)
    s.evolve(capi.LEAPFROG_KDK | capi.BARNES_HUT)  # warm-up
    t0 = time.perf_counter()
    s.advance(steps)
    t = (time.perf_counter() - t0) / steps
    s.profile(True)
    s.advance(2)
    c_ms, nc = s.profile_read("bh_coll")
    s.profile(False)
    s.close()
    pass_ms = c_ms / max(nc, 1)
    return {"bodies": BH_BODIES, "radius": radius, "ms_per_step": t * 1e3, "pass_ms": pass_ms,
            "bodies_per_s": BH_BODIES / (pass_ms * 1e-3), "steps": steps,
            "workload": f"BH KDK + bh_s_coll, two Plummer spheres of 524288, r={radius}, e=0.5, FP32 walk, FP64 pass"}


def bench_integrators(capi, steps=20):
    """ms per step of the other direct-sum integrators on the 65,536-body
    Plummer workload, mixed precision (Euler, modified Euler and midpoint
    evaluate forces once, twice and twice per step; DKD once)."""
    m, pos, vel = capi.ic_plummer(1, N_BODIES)
    out = {}
    for name, meth in (("euler", capi.EULER), ("modified_euler", capi.MODIFIED_EULER),
                       ("midpoint", capi.MIDPOINT), ("leapfrog_dkd", capi.LEAPFROG_DKD)):
        s = capi.System(0)
        s.set_G(1.0)
        s.set_timestep(DT)
        s.set_softening(EPS)
        s.set_precision(capi.PREC_MIXED)
        s.set_bodies(m, pos, vel)
        s.set_iterations(1)
        s.evolve(meth)  # warm-up
        t0 = time.perf_counter()
        s.advance(steps)
        out[name] = (time.perf_counter() - t0) / steps * 1e3
        s.close()
    return {"unit": "ms_per_step", "bodies": N_BODIES, **out}


def bench_adjoint(capi):
    """Adjoint workload: N = 4096 uniform ball (seed 4), eps = 1e6 m, dt = 10 s,
    2000 KDK steps, checkpoints every 64.  ratio = backward wall time (segment
    recompute + reverse sweep) / forward wall time (with checkpoint saves)."""
    n, steps, ckpt = 4096, 2000, 64
    m, x, v = capi.ic_uniform_ball(4, n, 1e20, 1e22, 1e9, 1e3)
    s = capi.System(0)
    s.set_G(6.6743e-11)
    s.set_timestep(10.0)
    s.set_softening(1e6)
    s.set_precision(capi.PREC_FP64_FAST)
    s.set_bodies(m, x, v)
    rng = np.random.default_rng(0)
    gx = rng.normal(size=(n, 3))
    gv = np.zeros((n, 3))
    s.adjoint(64, 64, gx, gv)  # warm-up (allocations, code objects)
    s.set_bodies(m, x, v)
    _, _, _, fns, bns = s.adjoint(steps, ckpt, gx, gv)
    s.close()
    return {"value": bns / fns, "unit": "backward/forward wall-time ratio", "higher_is_better": False,
            "forward_ms": fns / 1e6, "backward_ms": bns / 1e6, "bodies": n, "steps": steps, "ckpt_every": ckpt,
            "workload": "adjoint of direct KDK, uniform ball N=4096 (seed 4), eps=1e6 m, dt=10 s, 2000 steps, "
                        "FP64 split kernels"}


def cpu_bh_baseline(m, pos):
    """Oracle (corrected Barnes-Hut restatement, FP64, one thread) on a 65,536-body sample."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    t0 = time.perf_counter()
    _, cnt = O.bh_acc(m, pos, 1.0, EPS, BH_THETA)
    t = time.perf_counter() - t0
    return {"value": cnt / t, "unit": "interactions/s", "cores": 1, "kind": "port",
            "sample": f"oracle insertion octree + per-body MAC walk, 1 thread, {len(m)}-body two-Plummer sample, "
                      f"one force evaluation ({cnt} interactions, {t:.1f} s)"}


def _oracle():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    return O


def cpu_direct_allcore(m, pos):
    """All-core FP64 restatement (OpenMP over i, every i summing all j) on the
    full 65,536-body IC: one force evaluation."""
    O = _oracle()
    threads = O.max_threads()
    t0 = time.perf_counter()
    O.direct_acc_mt(m, pos, 1.0, EPS, threads)
    t = time.perf_counter() - t0
    n = len(m)
    return {"value": n * (n - 1) / t, "unit": "interactions/s", "cores": threads, "kind": "port",
            "sample": f"FP64 softened direct sum, OpenMP over {threads} host threads, full N={n} Plummer IC, "
                      f"one force evaluation ({t:.2f} s)"}


def cpu_bh_allcore(m, pos):
    """All-core corrected BH: serial insertion tree, OpenMP walk, full 1M bodies."""
    O = _oracle()
    threads = O.max_threads()
    t0 = time.perf_counter()
    _, cnt = O.bh_acc_mt(m, pos, 1.0, EPS, BH_THETA, threads)
    t = time.perf_counter() - t0
    return {"value": cnt / t, "unit": "interactions/s", "cores": threads, "kind": "port",
            "sample": f"oracle insertion octree (serial build) + per-body MAC walk over {threads} host threads, "
                      f"full {len(m)}-body two-Plummer IC, one force evaluation ({cnt} interactions, {t:.1f} s)"}


def bench_batch(capi, steps, precision=None):
    """Batched ensemble: 4096 comet-string systems (central body of 1.9e27 kg
    plus n ~ U{8..512} fragments, seed 3), dt = 5 s, merging on, one
    persistent launch per 1000 steps; mixed precision (FP32 pair forces) by
    default, or FP64 (the precision whose merge list is bit-exact against
    the double restatement).  Interactions are the force pairs actually
    evaluated, sum over steps and systems of n(n-1) with n reduced by each
    merge from the step after it; the collision scan's n(n-1)/2 overlap tests
    per step are reported beside them.  The roofline takes the stepping
    launches' HIP-event time, at 20 FLOP per interaction, against the FP32
    (mixed) or FP64 VALU peak."""
    G_SI = 6.6743e-11
    if precision is None:
        precision = capi.PREC_MIXED
    offsets, m, r, pos, vel, e = capi.ic_comet_ensemble(3, 4096, 8, 512)
    sizes = np.diff(offsets).astype(np.int64)
    b = capi.Batch(offsets, m, r, pos, vel, e)
    b.set_params(G_SI, 5.0, 0.0, capi.OVERLAP_COLL_CHECK, True, 1e10, precision)
    ns = b.evolve(steps)
    k_ms, k_launches = b.kernel_time()
    log = b.merge_log()
    b.close()
    t = ns * 1e-9
    # n_s(t) for the force pass of step t: merges at steps < t have removed bodies
    inter, tests = 0, 0
    counts = np.zeros(len(sizes), np.int64)
    if len(log):
        np.add.at(counts, log[:, 0], 1)
    for sys_id in range(len(sizes)):
        n0 = int(sizes[sys_id])
        msteps = np.sort(log[log[:, 0] == sys_id, 1]) if counts[sys_id] else np.zeros(0, np.int64)
        # piecewise-constant n over [1, steps]
        bounds = np.concatenate([[1], msteps + 1, [steps + 1]])
        for k in range(len(bounds) - 1):
            nk = n0 - k
            span = int(bounds[k + 1] - bounds[k])
            inter += span * nk * (nk - 1)
            tests += span * nk * (nk - 1) // 2
    mixed = precision == capi.PREC_MIXED
    peak = PEAK_FP32_TFLOPS if mixed else PEAK_FP64_TFLOPS
    achieved = FLOP_PER_INTERACTION * inter / (k_ms * 1e-3) / 1e12
    return {"value": inter / t, "unit": "interactions/s", "systems": len(sizes), "bodies": int(offsets[-1]),
            "steps": steps, "wall_s": t, "interactions": inter, "overlap_tests_per_s": tests / t,
            "merges": int(len(log)), "kernel_ms": k_ms, "launches": k_launches,
            "roofline": {"bound": "valu", "kernel": "k_batch_steps", "achieved": achieved, "peak": peak,
                         "unit": "TFLOP/s", "frac": achieved / peak, "flop_per_interaction": FLOP_PER_INTERACTION,
                         "avg_launch_ms": k_ms / max(k_launches, 1),
                         "traffic": PROFILED_BATCH_LAUNCH_BYTES if mixed else None},
            "workload": "4096 comet-string systems (seed 3): 1.9e27 kg central body + U{8..512} fragments, "
                        "dt=5 s, inelastic merging on, "
                        + ("FP32 pair forces from LDS" if mixed else "FP64 pair forces in the reference's order")
                        + ", FP64 state, exact FP64 collision replay; persistent launch, one workgroup per system"}


def profile_run(capi, args):
    """Fixed-configuration runs for counter collection (no autotuning, no HIP
    events): the kernels seen are exactly those of the timed step."""
    s = capi.System(0)
    s.set_G(1.0)
    s.set_timestep(DT)
    s.set_softening(EPS)
    s.set_precision(capi.PREC_MIXED)
    if args.profile_direct:
        m, pos, vel = capi.ic_plummer(1, N_BODIES)
        s.set_bodies(m, pos, vel)
        s.set_tuning("direct_kernel", args.kernel)
        s.set_tuning("direct_ipl", args.ipl)
        s.set_tuning("direct_waves", args.waves)
        s.set_iterations(args.steps)
        s.evolve(capi.LEAPFROG_KDK)
    else:
        m, pos, vel = capi.ic_two_plummer(2, BH_BODIES)
        s.set_bodies(m, pos, vel)
        s.set_bh_opening_angle(BH_THETA)
        s.set_tuning("bh_bucket", 16)  # the autotuner's usual choice at this size
        s.set_iterations(args.bh_steps)
        s.evolve(capi.LEAPFROG_KDK | capi.BARNES_HUT)
    s.close()
    print("profile run done", flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-bh", action="store_true")
    ap.add_argument("--bh-steps", type=int, default=5)
    ap.add_argument("--batch-steps", type=int, default=2000)
    ap.add_argument("--batch-fp64-steps", type=int, default=1000)
    ap.add_argument("--cpu-baseline-port", action="store_true",
                    help="time the oracle's restatement when the compiled reference (oracle/_ref) is absent")
    ap.add_argument("--hbm-copy", action="store_true", help="run only the HBM copy calibration kernel")
    ap.add_argument("--profile-direct", action="store_true",
                    help="for rocprofv3: the 65,536-body mixed direct step with fixed tuning, no events, no JSON")
    ap.add_argument("--kernel", type=int, default=1, help="--profile-direct: direct_kernel knob")
    ap.add_argument("--ipl", type=int, default=2, help="--profile-direct: direct_ipl knob")
    ap.add_argument("--waves", type=int, default=16, help="--profile-direct: direct_waves knob")
    ap.add_argument("--profile-bh-coll", action="store_true",
                    help="for rocprofv3: 3 BH steps with bh_s_coll on the 1M-body workload, radius 0.005")
    ap.add_argument("--profile-batch", action="store_true",
                    help="for rocprofv3: the 4096-system ensemble for --batch-steps steps, no JSON")
    ap.add_argument("--profile-bh", action="store_true",
                    help="for rocprofv3: the 1M-body BH step (mixed, bucket 16), no events, no JSON")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this engine runs one system per GPU and measures one MI355X; --gpus {args.gpus} is not "
              f"supported (multi-GPU sharding is a design note only)", file=sys.stderr)
        sys.exit(2)
    if args.hbm_copy:
        hbm_copy()
        return

    from gregpackagescpp_amd import capi

    if args.profile_batch:
        print(json.dumps(bench_batch(capi, args.batch_steps)), flush=True)
        return
    if args.profile_bh_coll:
        print(json.dumps(bench_bh_coll(capi, 3, 0.005)), flush=True)
        return
    if args.profile_direct or args.profile_bh:
        profile_run(capi, args)
        return

    n = N_BODIES
    m, pos, vel = capi.ic_plummer(1, n)
    s = capi.System(0)
    s.set_G(1.0)
    s.set_timestep(DT)
    s.set_softening(EPS)
    s.set_precision(capi.PREC_MIXED)
    s.set_bodies(m, pos, vel)
    s.autotune(capi.LEAPFROG_KDK)  # project autotuner: i-bodies per lane, j-split width
    tuned = {k: s.tuning(k) for k in ("direct_kernel", "direct_ipl", "direct_waves")}
    s.set_iterations(max(1, args.warmup))
    s.evolve(capi.LEAPFROG_KDK)  # warm-up: a(x0) + W steps

    t0 = time.perf_counter()
    s.advance(args.steps)  # synchronises the library's stream before returning
    t = time.perf_counter() - t0
    interactions = n * (n - 1) * args.steps
    value = interactions / t

    # per-kernel timing of the dominant kernel, HIP events on the library's stream
    s.profile(True)
    prof_steps = min(args.steps, 50)
    s.advance(prof_steps)
    force_ms, launches = s.profile_read("force")
    upd_ms, _ = s.profile_read("update")
    s.profile(False)
    avg_force_s = force_ms / launches * 1e-3
    achieved = FLOP_PER_INTERACTION * n * (n - 1) / avg_force_s / 1e12

    # HBM bytes per launch of k_force_mixed from rocprofv3 FETCH_SIZE / WRITE_SIZE
    # (profiles/k_force_mixed_plummer65536.md), each scaled by what the counter
    # reads for a copy kernel of known size; valid for the configuration it was
    # measured at (its write traffic scales with the j-split count).
    traffic = None
    key = (tuned["direct_kernel"], tuned["direct_ipl"], tuned["direct_waves"])
    if n == N_BODIES and key in PROFILED_FORCE_TRAFFIC:
        fetch_kib, write_kib = PROFILED_FORCE_TRAFFIC[key]
        traffic = (fetch_kib * FETCH_FACTOR + write_kib * WRITE_FACTOR) * 1024

    line = {
        "metric": "interactions/s at N=64K direct & 1M BH; adjoint/forward time ratio",
        "value": value,
        "unit": "interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 pair forces (Kahan-compensated sums), fp64 state",
        "data": "synthetic: seeded Plummer sphere (G=M=a=1, equal masses), no external data",
        "config": {"workload": "direct-sum leapfrog KDK, Plummer N=65536, eps=0.01, dt=1e-3, mixed precision",
                   "bodies": n, "global_batch": 1, "parallelism": "single GPU"},
        "roofline": {"bound": "valu", "kernel": "k_force_mixed_lds" if tuned["direct_kernel"] == 2 else "k_force_mixed",
                     "achieved": achieved, "peak": PEAK_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS, "traffic": traffic,
                     "avg_launch_ms": avg_force_s * 1e3, "update_kernel_avg_ms": upd_ms / max(launches, 1),
                     "flop_per_interaction": FLOP_PER_INTERACTION},
        "tuned": tuned,
    }
    s.close()
    if not args.no_bh:
        line["bh"] = bench_bh(capi, args.bh_steps)
        line["adjoint"] = bench_adjoint(capi)
        line["batch"] = bench_batch(capi, args.batch_steps)
        line["batch"]["fp64"] = bench_batch(capi, args.batch_fp64_steps, capi.PREC_FP64)
        line["bh_coll"] = [bench_bh_coll(capi, 3, 0.002), bench_bh_coll(capi, 3, 0.005)]
        line["integrators"] = bench_integrators(capi)
    if not args.no_cpu_baseline:
        line["host"] = host_info()
        line["cpu_baseline"] = cpu_baseline(m, pos, vel, allow_port=args.cpu_baseline_port)
        line["cpu_baseline_allcore"] = cpu_direct_allcore(m, pos)
        if not args.no_bh:
            bm, bp, _ = capi.ic_two_plummer(2, 65536)
            line["bh"]["cpu_baseline"] = cpu_bh_baseline(bm, bp)
            bm, bp, _ = capi.ic_two_plummer(2, BH_BODIES)
            line["bh"]["cpu_baseline_allcore"] = cpu_bh_allcore(bm, bp)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
