"""Direct-summation KDK on the GPU through the C ABI, against the oracle.

Bars:
  * FP64 path (NB_PREC_FP64): bit-identical to the IEEE-double restatement of
    the reference's operation order (oracle *_d), for both pair forms.
  * FP64 path vs the compiled reference's long-double KDK fixtures: positions
    within 1e-12 relative, velocities within 1e-9 of the largest increment or
    8 FP64 ulps of |v| (only the arithmetic type differs).
  * mixed path (FP32 forces, Kahan-compensated FP32 sums, FP64 state): the
    velocity increment v_T - v_0 within 1e-5 relative of the long-double
    oracle's (FP32 pair forces carry ~1e-7 relative error per term).
"""
import glob
import os

import numpy as np
import pytest

import oracle as O
from gregpackagescpp_amd import capi

pytestmark = pytest.mark.gpu
G_SI = 6.6743e-11
LD = np.longdouble


def make_system(m, pos, vel, dt, steps, eps=0.0, prec=capi.PREC_FP64, G=G_SI):
    s = capi.System(0)
    s.set_G(G)
    s.set_timestep(dt)
    s.set_iterations(steps)
    s.set_softening(eps)
    s.set_precision(prec)
    s.set_bodies(m, pos, vel)
    return s


@pytest.mark.parametrize("n", [1, 2, 3, 64, 257])
@pytest.mark.parametrize("eps", [0.0, 1e6])
def test_fp64_bit_exact_vs_double_oracle(n, eps):
    print(f"case fp64 n={n} eps={eps}", flush=True)
    m, pos, vel = capi.ic_uniform_ball(11, n, 1e20, 1e22, 1e9, 1e3)
    steps, dt = 10, 10.0
    s = make_system(m, pos, vel, dt, steps, eps)
    s.evolve(capi.LEAPFROG_KDK)
    st = s.state()
    ref = O.kdk(m, pos, vel, G_SI, dt, steps, eps=eps, form=O.FORM_SOFT if eps > 0 else O.FORM_NEWTON,
                dtype=np.float64)
    assert np.array_equal(st["pos"], ref["pos"])
    assert np.array_equal(st["vel"], ref["vel"])
    assert np.array_equal(st["acc"], ref["acc"])


def test_fp64_energy_bit_exact():
    m, pos, vel = capi.ic_uniform_ball(12, 300, 1e20, 1e22, 1e9, 1e3)
    for eps, form in ((0.0, O.FORM_NEWTON), (1e6, O.FORM_SOFT)):
        s = make_system(m, pos, vel, 10.0, 5, eps)
        ke, pe = s.energy()
        rke, rpe = O.energy(m, pos, vel, G_SI, eps, form, dtype=np.float64)
        assert ke == rke and pe == rpe


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden",
                                                               "ref_kdk_*.npz"))))
def test_fp64_vs_reference_fixture(path):
    f = np.load(path)
    steps = int(f["steps"])
    print(f"case {os.path.basename(path)}", flush=True)
    s = make_system(f["m"], f["pos"], f["vel"], float(f["dt"]), steps)
    s.evolve(capi.LEAPFROG_KDK)
    st = s.state()
    pos_T, vel_T = f["pos_T"].astype(LD), f["vel_T"].astype(LD)
    scale = np.abs(f["pos"]).max()
    assert np.abs(st["pos"] - pos_T).max() / scale < 1e-12
    dv_ref = vel_T - f["vel"]
    # velocities: 1e-9 of the largest increment, or a few FP64 ulps of |v|
    tol = max(1e-9 * float(np.abs(dv_ref).max()), 8 * np.finfo(np.float64).eps * float(np.abs(vel_T).max()))
    assert float(np.abs(st["vel"] - vel_T).max()) < tol


@pytest.mark.parametrize("n,eps", [(257, 1e6), (1024, 1e6), (130, 0.0)])
def test_mixed_vs_long_double_oracle(n, eps):
    print(f"case mixed n={n} eps={eps}", flush=True)
    m, pos, vel = capi.ic_uniform_ball(0, n, 1e20, 1e22, 1e9, 1e3)
    steps, dt = 50, 10.0
    s = make_system(m, pos, vel, dt, steps, eps, prec=capi.PREC_MIXED)
    s.evolve(capi.LEAPFROG_KDK)
    st = s.state()
    ref = O.kdk(m, pos, vel, G_SI, dt, steps, eps=eps, form=O.FORM_SOFT if eps > 0 else O.FORM_NEWTON)
    dv_ref = (ref["vel"] - vel).astype(np.float64)
    dv = st["vel"] - vel
    err = np.abs(dv - dv_ref).max() / np.abs(dv_ref).max()
    print(f"mixed velocity-increment error {err:.2e}")
    assert err < 1e-5
    dx_ref = (ref["pos"] - pos).astype(np.float64)
    assert np.abs((st["pos"] - pos) - dx_ref).max() / np.abs(dx_ref).max() < 1e-5


def test_config0_full_run():
    """The stated config-0 workload: N = 1024 uniform ball (seed 0), masses
    U[1e20, 1e22] kg, R = 1e9 m, sigma = 1e3 m/s, eps = 1e6 m, dt = 10 s,
    1000 KDK steps; outputs final positions, velocities and total energy.

    Bars:
      * FP64 path: positions, velocities, accelerations and the (KE, PE) of
        the final state bit-identical to the double restatement's;
      * FP64 path vs the long-double oracle (only the arithmetic type
        differs): displacement x_T - x_0 and velocity increment v_T - v_0
        within 1e-11 relative (max-norm), energy within 1e-14 of |E|
        (measured on MI355X: 1.1e-13, 3.0e-13, 1.7e-16);
      * mixed path (FP32 pair forces, Kahan sums) vs the long-double oracle:
        displacement within 1e-7, velocity increment within 1e-5 relative
        (FP32 terms carry ~1e-7 relative error; 1000 steps add them up), and
        total energy within 1e-9 of |E| (the velocities are ~1e3 m/s against
        increments of a few m/s, so increment errors move KE very little;
        measured: 6.9e-9, 4.4e-7, 2.1e-11)."""
    n, steps, dt, eps = 1024, 1000, 10.0, 1e6
    m, pos, vel = capi.ic_uniform_ball(0, n, 1e20, 1e22, 1e9, 1e3)
    print("case config0 fp64 vs double oracle", flush=True)
    s = make_system(m, pos, vel, dt, steps, eps)
    s.evolve(capi.LEAPFROG_KDK)
    st = s.state()
    ke, pe = s.energy()
    ref = O.kdk(m, pos, vel, G_SI, dt, steps, eps=eps, form=O.FORM_SOFT, dtype=np.float64)
    for k in ("pos", "vel", "acc"):
        assert np.array_equal(st[k], ref[k]), k
    rke, rpe = O.energy(m, ref["pos"], ref["vel"], G_SI, eps, O.FORM_SOFT, dtype=np.float64)
    assert ke == rke and pe == rpe
    print("case config0 long-double oracle", flush=True)
    ld = O.kdk(m, pos, vel, G_SI, dt, steps, eps=eps, form=O.FORM_SOFT)
    lke, lpe = O.energy(m, ld["pos"], ld["vel"], G_SI, eps, O.FORM_SOFT)
    E_ld = float(lke + lpe)
    dx_ref = (ld["pos"] - pos.astype(LD)).astype(np.float64)
    dv_ref = (ld["vel"] - vel.astype(LD)).astype(np.float64)

    def errs(state, E):
        ex = np.abs((state["pos"] - pos) - dx_ref).max() / np.abs(dx_ref).max()
        ev = np.abs((state["vel"] - vel) - dv_ref).max() / np.abs(dv_ref).max()
        return ex, ev, abs(E - E_ld) / abs(E_ld)

    ex, ev, ee = errs(st, ke + pe)
    print(f"fp64 vs long double: x {ex:.2e} v {ev:.2e} E {ee:.2e}", flush=True)
    assert ex < 1e-11 and ev < 1e-11 and ee < 1e-14
    print("case config0 mixed", flush=True)
    s = make_system(m, pos, vel, dt, steps, eps, prec=capi.PREC_MIXED)
    s.evolve(capi.LEAPFROG_KDK)
    ex, ev, ee = errs(s.state(), sum(s.energy()))
    print(f"mixed vs long double: x {ex:.2e} v {ev:.2e} E {ee:.2e}", flush=True)
    assert ex < 1e-7 and ev < 1e-5 and ee < 1e-9


def test_advance_continues_evolve():
    m, pos, vel = capi.ic_uniform_ball(3, 100, 1e20, 1e22, 1e9, 1e3)
    a = make_system(m, pos, vel, 10.0, 20, 1e6)
    a.evolve(capi.LEAPFROG_KDK)
    b = make_system(m, pos, vel, 10.0, 10, 1e6)
    b.evolve(capi.LEAPFROG_KDK)
    b.advance(10)
    sa, sb = a.state(), b.state()
    assert np.array_equal(sa["pos"], sb["pos"]) and np.array_equal(sa["vel"], sb["vel"])


def test_mixed_advance_matches_evolve():
    m, pos, vel = capi.ic_plummer(1, 4096)
    a = make_system(m, pos, vel, 1e-3, 6, 0.01, capi.PREC_MIXED, G=1.0)
    a.evolve(capi.LEAPFROG_KDK)
    b = make_system(m, pos, vel, 1e-3, 3, 0.01, capi.PREC_MIXED, G=1.0)
    b.evolve(capi.LEAPFROG_KDK)
    b.advance(3)
    assert np.array_equal(a.state()["pos"], b.state()["pos"])


def test_empty_and_single_body():
    s = capi.System(0)
    s.set_bodies(np.zeros(0), np.zeros((0, 3)), np.zeros((0, 3)))
    assert s.evolve() == 0  # evolve returns 0 ns for an empty system
    m, pos, vel = np.array([5.0]), np.array([[1.0, 2.0, 3.0]]), np.array([[0.5, -1.0, 2.0]])
    for prec in (capi.PREC_FP64, capi.PREC_MIXED):
        s = make_system(m, pos, vel, 0.5, 4, prec=prec)
        s.evolve()
        st = s.state()
        assert np.array_equal(st["pos"], pos + 4 * (0.5 * vel))  # exact: free drift in binary steps
        assert np.array_equal(st["vel"], vel)


def test_config1_energy_drift_mixed_vs_fp64():
    """Plummer N = 65,536, eps = 0.01, dt = 1e-3, 1000 steps: the mixed path's
    energy stays within 1e-6 (relative to |E0|) of the FP64 path's."""
    n, steps = 65536, 1000
    m, pos, vel = capi.ic_plummer(1, n)
    out = {}
    for prec in (capi.PREC_MIXED, capi.PREC_FP64):
        print(f"case config1 prec={prec}", flush=True)
        s = make_system(m, pos, vel, 1e-3, steps, 0.01, prec, G=1.0)
        e0 = sum(s.energy())
        s.evolve()
        out[prec] = (e0, sum(s.energy()))
    e0 = out[capi.PREC_FP64][0]
    d_mixed = (out[capi.PREC_MIXED][1] - e0) / abs(e0)
    d_fp64 = (out[capi.PREC_FP64][1] - e0) / abs(e0)
    print(f"relative energy drift: mixed {d_mixed:.3e}, fp64 {d_fp64:.3e}")
    assert abs(d_mixed - d_fp64) < 1e-6


def test_error_paths():
    s = capi.System(0)
    with pytest.raises(capi.NBodyError) as e:
        s.set_timestep(0.0)
    assert e.value.code == capi.NB_ERR_INVALID_ARGUMENT
    with pytest.raises(capi.NBodyError) as e:
        s.set_bodies(np.array([1.0, -1.0]), np.zeros((2, 3)), np.zeros((2, 3)))
    assert e.value.code == capi.NB_ERR_NEGATIVE_MASS
    with pytest.raises(capi.NBodyError) as e:
        s.set_bodies(np.ones(2), np.array([[0, 0, 0], [1.0, 0, 0]]), np.zeros((2, 3)), r=np.ones(2),
                     check_overlap=True)
    assert e.value.code == capi.NB_ERR_OVERLAPPING_BODIES
    s.set_bodies(np.ones(3), np.eye(3), np.zeros((3, 3)))
    with pytest.raises(capi.NBodyError) as e:
        s.evolve(capi.TWO_BODY)
    assert e.value.code == capi.NB_ERR_TWO_BODY
    with pytest.raises(capi.NBodyError) as e:
        s.advance(1)
    assert e.value.code == capi.NB_ERR_NO_EVOLUTION
    with pytest.raises(capi.NBodyError):
        s.set_units("M1:0,D1:1,T1:1")
    s.set_units("M1:1,D1:1,T1:1")
    assert s.G == float(LD(66743) / LD(10 ** 15))


def test_tuning_knobs_keep_results():
    """Every (i-bodies per lane, waves per SIMD) setting of the FP32 kernel and
    the autotuner's pick give the same dynamics to FP32 summation accuracy."""
    m, pos, vel = capi.ic_plummer(2, 8192)
    ref = None
    for ipl in (1, 2, 4):
        for waves in (4, 16):
            s = make_system(m, pos, vel, 1e-3, 5, 0.01, capi.PREC_MIXED, G=1.0)
            s.set_tuning("direct_ipl", ipl)
            s.set_tuning("direct_waves", waves)
            s.evolve()
            dv = s.state()["vel"] - vel
            if ref is None:
                ref = dv
            assert np.abs(dv - ref).max() / np.abs(ref).max() < 1e-6, (ipl, waves)
    s = make_system(m, pos, vel, 1e-3, 5, 0.01, capi.PREC_MIXED, G=1.0)
    s.autotune(capi.LEAPFROG_KDK)
    assert s.tuning("direct_ipl") in (1, 2, 4) and s.tuning("direct_waves") in (4, 8, 16)
    s.evolve()
    assert np.abs((s.state()["vel"] - vel) - ref).max() / np.abs(ref).max() < 1e-6
    with pytest.raises(capi.NBodyError):
        s.set_tuning("direct_ipl", 3)


@pytest.mark.parametrize("prec", [capi.PREC_FP64, capi.PREC_MIXED, capi.PREC_FP64_FAST])
def test_step_graphs_match_eager(prec):
    """Interior steps of small systems replay a captured HIP graph of 32 steps;
    with graphs off every launch is issued directly.  Same kernels in the same
    order: bit-identical states.  The FP64 run is also held bit-exact to the
    double restatement over 100 steps (three graph replays)."""
    print(f"case step graphs prec={prec}", flush=True)
    m, pos, vel = capi.ic_uniform_ball(13, 300, 1e20, 1e22, 1e9, 1e3)
    outs = []
    for graphs in (1, 2):
        s = make_system(m, pos, vel, 10.0, 100, 1e6, prec)
        s.set_tuning("step_graphs", graphs)
        s.evolve(capi.LEAPFROG_KDK)
        s.advance(70)  # a second call reuses the captured graph
        outs.append(s.state())
    for k in ("pos", "vel", "acc"):
        assert np.array_equal(outs[0][k], outs[1][k])
    if prec == capi.PREC_FP64:
        s = make_system(m, pos, vel, 10.0, 100, 1e6, prec)
        s.set_tuning("step_graphs", 1)
        s.evolve(capi.LEAPFROG_KDK)
        ref = O.kdk(m, pos, vel, G_SI, 10.0, 100, eps=1e6, form=O.FORM_SOFT, dtype=np.float64)
        assert np.array_equal(s.state()["pos"], ref["pos"])
        assert np.array_equal(s.state()["vel"], ref["vel"])


def test_fp64_one_lane_kernel_above_term_path():
    """Above 16,384 bodies the FP64 reference-order path runs one lane per i
    (below, the pair terms are evaluated in parallel and summed in order by a
    second kernel); both must equal the double restatement bit for bit."""
    n = 16400
    print(f"case fp64 one-lane n={n}", flush=True)
    m, pos, vel = capi.ic_uniform_ball(14, n, 1e20, 1e22, 1e9, 1e3)
    s = make_system(m, pos, vel, 10.0, 1, 1e6)
    s.evolve(capi.LEAPFROG_KDK)
    st = s.state()
    ref = O.kdk(m, pos, vel, G_SI, 10.0, 1, eps=1e6, form=O.FORM_SOFT, dtype=np.float64)
    assert np.array_equal(st["pos"], ref["pos"])
    assert np.array_equal(st["acc"], ref["acc"])


OTHER_METHODS = [capi.EULER, capi.MODIFIED_EULER, capi.MIDPOINT, capi.LEAPFROG_DKD]


@pytest.mark.parametrize("method", OTHER_METHODS)
@pytest.mark.parametrize("eps", [0.0, 1e6])
def test_other_integrators_fp64_bit_exact(method, eps):
    """evolve(euler | modified_euler | midpoint | leapfrog_dkd) in FP64: the
    reference's operation order, bit-identical to the double restatement
    (whose long-double form is pinned to the compiled reference); nb_advance
    continues the same sequence."""
    print(f"case other integrator method={method} eps={eps}", flush=True)
    m, pos, vel = capi.ic_uniform_ball(15, 257, 1e20, 1e22, 1e9, 1e3)
    s = make_system(m, pos, vel, 10.0, 12, eps)
    s.evolve(method)
    s.advance(8)
    st = s.state()
    form = O.FORM_SOFT if eps > 0 else O.FORM_NEWTON
    ref = O.kdk(m, pos, vel, G_SI, 10.0, 20, eps=eps, form=form, dtype=np.float64, method=method)
    assert np.array_equal(st["pos"], ref["pos"])
    assert np.array_equal(st["vel"], ref["vel"])
    assert np.array_equal(st["acc"], ref["acc"])
    assert s.time_elapsed() == 200.0


@pytest.mark.parametrize("method", OTHER_METHODS)
@pytest.mark.parametrize("prec", [capi.PREC_MIXED, capi.PREC_FP64_FAST])
def test_other_integrators_fast_precisions(method, prec):
    """The same integrators with FP32 pair forces (Kahan sums) or the FP64
    split kernels: velocity increments within 1e-5 (mixed) / 1e-10 (FP64
    split: rsq + 2 Newton steps and a split summation order, whose ~1e-16
    per-term errors are amplified by cancellation among the 1023 terms of
    each acceleration) relative of the long-double oracle."""
    print(f"case other integrator method={method} prec={prec}", flush=True)
    m, pos, vel = capi.ic_uniform_ball(16, 1024, 1e20, 1e22, 1e9, 1e3)
    s = make_system(m, pos, vel, 10.0, 10, 1e6, prec)
    s.evolve(method)
    st = s.state()
    ref = O.kdk(m, pos, vel, G_SI, 10.0, 10, eps=1e6, form=O.FORM_SOFT, method=method)
    dv = st["vel"] - vel
    dv_ref = (ref["vel"] - vel.astype(LD)).astype(np.float64)
    tol = 1e-5 if prec == capi.PREC_MIXED else 1e-10
    assert np.abs(dv - dv_ref).max() <= tol * np.abs(dv_ref).max()
