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
               long double, one thread) on a bounded sample, else the oracle.
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
PROFILED_FORCE_TRAFFIC = {"bodies": 65536, "ipl": 2, "waves": 16, "fetch_kib": 7709.8, "write_kib": 49152.0,
                          "fetch_factor": 2.0, "write_factor": 1.0}


def cpu_baseline(m, pos, vel):
    """Time the reference CPU path on a bounded sample of the same workload."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    ns_sample, steps = 8192, 2
    mm, pp, vv = m[:ns_sample].copy(), pos[:ns_sample].copy(), vel[:ns_sample].copy()
    pairs_per_pass = ns_sample * (ns_sample - 1)
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


def bench_bh(capi, steps):
    """Barnes-Hut KDK on two 524,288-body Plummer spheres (theta = 0.5, eps =
    0.01, dt = 1e-3, full rebuild every step).  Interactions are I_alg: the
    cells accepted under the reference's per-body MAC, counted by the FP64
    per-lane walk on the same positions."""
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
    s.autotune(capi.LEAPFROG_KDK | capi.BARNES_HUT)  # project autotuner: bucket size of the wave walk
    s.set_iterations(1)
    s.evolve(capi.LEAPFROG_KDK | capi.BARNES_HUT)  # warm-up
    t0 = time.perf_counter()
    s.advance(steps)
    t = (time.perf_counter() - t0) / steps
    s.profile(True)
    s.advance(2)
    b_ms, nb = s.profile_read("bh_build")
    w_ms, nw = s.profile_read("bh_walk")
    s.profile(False)
    _, wave_cnt = s.bh_accelerations()
    bucket = s.tuning("bh_bucket")
    s.close()
    return {"value": i_alg / t, "unit": "interactions/s", "bodies": BH_BODIES, "theta": BH_THETA,
            "ms_per_step": t * 1e3, "interactions_per_step": i_alg, "wave_walk_interactions": wave_cnt,
            "tuned": {"bh_bucket": bucket},
            "build_ms": b_ms / max(nb, 1), "walk_ms": w_ms / max(nw, 1), "steps": steps,
            "workload": "BH leapfrog KDK, two Plummer spheres of 524288 at x=+-10, vrel=0.5, theta=0.5, "
                        "eps=0.01, dt=1e-3, FP32 wave walk, FP64 state and node COM"}


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
        s.set_tuning("direct_ipl", 2)
        s.set_tuning("direct_waves", 16)
        s.set_iterations(args.steps)
        s.evolve(capi.LEAPFROG_KDK)
    else:
        m, pos, vel = capi.ic_two_plummer(2, BH_BODIES)
        s.set_bodies(m, pos, vel)
        s.set_bh_opening_angle(BH_THETA)
        s.set_tuning("bh_bucket", 32)
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
    ap.add_argument("--hbm-copy", action="store_true", help="run only the HBM copy calibration kernel")
    ap.add_argument("--profile-direct", action="store_true",
                    help="for rocprofv3: the 65,536-body mixed direct step with fixed tuning, no events, no JSON")
    ap.add_argument("--profile-bh", action="store_true",
                    help="for rocprofv3: the 1M-body BH step (mixed, bucket 32), no events, no JSON")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this engine runs one system per GPU and measures one MI355X; --gpus {args.gpus} is not "
              f"supported (multi-GPU sharding is a design note only)", file=sys.stderr)
        sys.exit(2)
    if args.hbm_copy:
        hbm_copy()
        return

    from gregpackagescpp_amd import capi

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
    tuned = {"direct_ipl": s.tuning("direct_ipl"), "direct_waves": s.tuning("direct_waves")}
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
    prof = PROFILED_FORCE_TRAFFIC
    if n == prof["bodies"] and tuned["direct_ipl"] == prof["ipl"] and tuned["direct_waves"] == prof["waves"]:
        traffic = (prof["fetch_kib"] * prof["fetch_factor"] + prof["write_kib"] * prof["write_factor"]) * 1024

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
        "roofline": {"bound": "valu", "kernel": "k_force_mixed", "achieved": achieved, "peak": PEAK_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS, "traffic": traffic,
                     "avg_launch_ms": avg_force_s * 1e3, "update_kernel_avg_ms": upd_ms / max(launches, 1),
                     "flop_per_interaction": FLOP_PER_INTERACTION},
        "tuned": tuned,
    }
    s.close()
    if not args.no_bh:
        line["bh"] = bench_bh(capi, args.bh_steps)
        line["adjoint"] = bench_adjoint(capi)
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(m, pos, vel)
        if not args.no_bh:
            bm, bp, _ = capi.ic_two_plummer(2, 65536)
            line["bh"]["cpu_baseline"] = cpu_bh_baseline(bm, bp)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
