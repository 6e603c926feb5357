"""Adjoint of the direct KDK on the GPU (nb_adjoint), against the oracle's
FP64 CPU adjoint and central finite differences.

Bars:
  * GPU adjoint vs the oracle's FP64 reverse-mode gradient (itself pinned by
    finite differences in tests/test_oracle.py): every component of dL/dx0,
    dL/dv0, dL/dm within 1e-9 relative (max-norm).
  * checkpoint interval does not change the result (bit-identical).
  * the 4096-body workload on a 128-step horizon: all three gradients
    (dL/dm included) within 1e-9 of the oracle adjoint.
  * the 4096-body, 2000-step workload (checkpoints every 64): dL/dx0 and
    dL/dv0 components within 1e-5 (relative to the gradient's max-norm) of
    central differences with h = 1e-7 relative.  (dL/dm is not checked by
    FP64 differences at this scale: a
    1e-7 relative mass change moves x_T by ~1e-3 m while FP64 rounding on
    |x| ~ 1e9 m over 2000 steps contributes ~1e-6 m per body, which buries the
    difference quotient; dL/dm is pinned by the oracle adjoint instead.)
"""
import numpy as np
import pytest

import oracle as O
from gregpackagescpp_amd import capi

pytestmark = pytest.mark.gpu
G_SI = 6.6743e-11


def sys_fast(m, x, v, dt=10.0, eps=1e6, steps=1, G=G_SI):
    s = capi.System(0)
    s.set_G(G)
    s.set_timestep(dt)
    s.set_softening(eps)
    s.set_precision(capi.PREC_FP64_FAST)
    s.set_iterations(steps)
    s.set_bodies(m, x, v)
    return s


def rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.mark.parametrize("n,steps,ckpt", [(64, 100, 16), (200, 37, 10)])
def test_adjoint_vs_oracle(n, steps, ckpt):
    print(f"case adjoint vs oracle n={n} steps={steps}", flush=True)
    m, x, v = capi.ic_uniform_ball(21, n, 1e20, 1e22, 1e9, 1e3)
    rng = np.random.default_rng(1)
    gxT, gvT = rng.normal(size=(n, 3)), rng.normal(size=(n, 3)) * 1e3
    s = sys_fast(m, x, v)
    gx, gv, gm, fns, bns = s.adjoint(steps, ckpt, gxT, gvT)
    rx, rv, rm = O.adjoint(m, x, v, G_SI, 10.0, 1e6, steps, gxT, gvT)
    print(f"rel err x {rel(gx, rx):.2e} v {rel(gv, rv):.2e} m {rel(gm, rm):.2e}; fwd {fns/1e6:.2f} ms bwd {bns/1e6:.2f} ms")
    assert rel(gx, rx) < 1e-9 and rel(gv, rv) < 1e-9 and rel(gm, rm) < 1e-9
    # the state is left at time T (the forward's end state)
    xT, vT = O.forward(m, x, v, G_SI, 10.0, 1e6, steps)
    assert rel(s.state()["pos"], xT) < 1e-13


@pytest.mark.parametrize("ics", ["si_ball", "unit_plummer"])
def test_adjoint_checkpoint_interval_invariant(ics):
    """Recomputed segments must land on the forward trajectory bit for bit,
    whatever the checkpoint interval.  The unit-scale Plummer case (G = 1,
    dt = 1e-2, |v| ~ 0.5, |x| ~ 1) makes dt*|v| and h*|a| large against
    ulp(x) and ulp(v), so a kick or drift rounded differently from the forward
    step (a fused multiply-add on one side only) would change the gradients."""
    print(f"case checkpoint invariance {ics}", flush=True)
    if ics == "si_ball":
        m, x, v = capi.ic_uniform_ball(22, 300, 1e20, 1e22, 1e9, 1e3)
        kw = {}
    else:
        m, x, v = capi.ic_plummer(22, 300)
        kw = dict(dt=1e-2, eps=0.05, G=1.0)
    rng = np.random.default_rng(2)
    gxT, gvT = rng.normal(size=(300, 3)), rng.normal(size=(300, 3))
    outs = [sys_fast(m, x, v, **kw).adjoint(50, ck, gxT, gvT)[:3] for ck in (1, 7, 50, 64)]
    for o in outs[1:]:
        for a, b in zip(outs[0], o):
            assert np.array_equal(a, b)


def test_adjoint_graphs_match_eager():
    """Full forward segments and interior reverse segments replay captured HIP
    graphs; with graphs off every launch is issued directly.  Same kernels in
    the same order, so the gradients must be bit-identical."""
    m, x, v = capi.ic_uniform_ball(23, 257, 1e20, 1e22, 1e9, 1e3)
    rng = np.random.default_rng(3)
    gxT, gvT = rng.normal(size=(257, 3)), rng.normal(size=(257, 3))
    outs = []
    for graphs in (1, 2):
        s = sys_fast(m, x, v)
        s.set_tuning("step_graphs", graphs)
        outs.append(s.adjoint(203, 16, gxT, gvT)[:3] + (s.state()["pos"],))
    for a, b in zip(*outs):
        assert np.array_equal(a, b)


def test_config4_finite_differences():
    """N = 4096 uniform ball (seed 4), eps = 1e6 m, dt = 10 s, 2000 steps,
    checkpoints every 64; loss = sum |x_T - x*_T|^2 with targets from ICs
    perturbed by 1e-6 relative Gaussian noise."""
    n, steps, dt, eps = 4096, 2000, 10.0, 1e6
    m, x, v = capi.ic_uniform_ball(4, n, 1e20, 1e22, 1e9, 1e3)
    rng = np.random.default_rng(4)
    xp = x * (1 + 1e-6 * rng.normal(size=x.shape))
    vp = v * (1 + 1e-6 * rng.normal(size=v.shape))

    def final(m_, x_, v_):
        s = sys_fast(m_, x_, v_, dt, eps, steps)
        s.evolve()
        st = s.state()
        s.close()
        return st["pos"]

    target = final(m, xp, vp)

    def loss(m_, x_, v_):
        return float(((final(m_, x_, v_) - target) ** 2).sum())

    xT = final(m, x, v)
    s = sys_fast(m, x, v, dt, eps, steps)
    gx, gv, gm, fns, bns = s.adjoint(steps, 64, 2 * (xT - target), np.zeros((n, 3)))
    print(f"forward {fns/1e6:.1f} ms, backward {bns/1e6:.1f} ms, ratio {bns/fns:.2f}", flush=True)
    # Normwise agreement: |fd - adjoint| over the gradient's max-norm.  (The
    # difference quotient of a small component carries FP64 rounding noise of
    # the 2000-step trajectories, ~1e-4 of that component.)
    picks = [(0, 0), (17, 1), (1000, 2), (2048, 0), (4095, 1)]
    worst = 0.0
    for (i, k) in picks:
        for which in ("x", "v"):
            base = x if which == "x" else v
            grad = gx if which == "x" else gv
            h = 1e-7 * abs(base[i, k])
            p, q = base.copy(), base.copy()
            p[i, k] += h
            q[i, k] -= h
            lp = loss(m, p, v) if which == "x" else loss(m, x, p)
            lq = loss(m, q, v) if which == "x" else loss(m, x, q)
            fd = (lp - lq) / (2 * h)
            g = grad[i, k]
            err = abs(fd - g) / np.abs(grad).max()
            worst = max(worst, err)
            print(f"d/d{which}[{i},{k}]: adjoint {g:.10e} fd {fd:.10e} normwise {err:.1e}", flush=True)
    assert worst < 1e-5


def test_config4_full_n_vs_oracle_adjoint():
    """Config 4 at its full size (N = 4096 uniform ball, seed 4, eps = 1e6 m,
    dt = 10 s, checkpoints every 64) on a shortened horizon of 128 steps (two
    checkpoint segments): the config's loss, sum |x_T - x*_T|^2 with targets
    from ICs perturbed by 1e-6 relative Gaussian noise, and all three
    gradients -- dL/dx0, dL/dv0 and dL/dm -- within 1e-9 (max-norm relative)
    of the oracle's FP64 reverse-mode adjoint on the same cotangents."""
    n, steps, dt, eps = 4096, 128, 10.0, 1e6
    print(f"case config4 full N n={n} steps={steps}", flush=True)
    m, x, v = capi.ic_uniform_ball(4, n, 1e20, 1e22, 1e9, 1e3)
    rng = np.random.default_rng(4)
    xp = x * (1 + 1e-6 * rng.normal(size=x.shape))
    vp = v * (1 + 1e-6 * rng.normal(size=v.shape))
    t = sys_fast(m, xp, vp, dt, eps, steps)
    t.evolve()
    target = t.state()["pos"]
    t.close()
    s = sys_fast(m, x, v, dt, eps, steps)
    s.evolve()
    gxT = 2 * (s.state()["pos"] - target)
    s.close()
    gvT = np.zeros((n, 3))
    s = sys_fast(m, x, v, dt, eps, steps)
    gx, gv, gm, fns, bns = s.adjoint(steps, 64, gxT, gvT)
    print("case config4 oracle adjoint", flush=True)
    rx, rv, rm = O.adjoint(m, x, v, G_SI, dt, eps, steps, gxT, gvT)
    print(f"rel err x {rel(gx, rx):.2e} v {rel(gv, rv):.2e} m {rel(gm, rm):.2e}", flush=True)
    assert rel(gx, rx) < 1e-9 and rel(gv, rv) < 1e-9 and rel(gm, rm) < 1e-9


def test_mass_gradient_vs_oracle_midsize():
    n, steps = 512, 200
    m, x, v = capi.ic_uniform_ball(4, n, 1e20, 1e22, 1e9, 1e3)
    rng = np.random.default_rng(5)
    gxT = rng.normal(size=(n, 3))
    gx, gv, gm, _, _ = sys_fast(m, x, v).adjoint(steps, 64, gxT, np.zeros((n, 3)))
    rx, rv, rm = O.adjoint(m, x, v, G_SI, 10.0, 1e6, steps, gxT, np.zeros((n, 3)))
    assert rel(gm, rm) < 1e-9 and rel(gx, rx) < 1e-9 and rel(gv, rv) < 1e-9


def test_autograd_function_harness():
    import torch

    from gregpackagescpp_amd.autograd import KDKEvolve, KDKParams

    n = 40
    m, x, v = capi.ic_uniform_ball(23, n, 1e20, 1e22, 1e9, 1e3)
    p = KDKParams(G=G_SI, dt=10.0, eps=1e6, steps=30, ckpt_every=8)
    tx = torch.tensor(x, requires_grad=True)
    tv = torch.tensor(v, requires_grad=True)
    tm = torch.tensor(m, requires_grad=True)
    tgt = torch.tensor(x * 1.000001)
    xT, vT = KDKEvolve.apply(tx, tv, tm, p)
    loss = ((xT - tgt) ** 2).sum() + 1e-3 * (vT ** 2).sum()
    loss.backward()
    xTn = xT.detach().numpy()
    rx, rv, rm = O.adjoint(m, x, v, G_SI, 10.0, 1e6, 30, 2 * (xTn - x * 1.000001), 2e-3 * vT.detach().numpy())
    assert rel(tx.grad.numpy(), rx) < 1e-9
    assert rel(tv.grad.numpy(), rv) < 1e-9
    assert rel(tm.grad.numpy(), rm) < 1e-9
