"""torch.autograd.Function over the engine's KDK forward and its adjoint.

    xT, vT = KDKEvolve.apply(x0, v0, m, params)

x0, v0: (N, 3) float64 tensors, m: (N,) float64; params: KDKParams.  The
forward runs `steps` direct-summation KDK steps on the GPU through the C ABI
(FP64 split kernels); backward calls nb_adjoint, which checkpoints every
`ckpt_every` steps and returns dL/dx0, dL/dv0, dL/dm.  Tensors travel through
host memory (the C ABI takes host arrays), so they may live on any device.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from . import capi


@dataclass(frozen=True)
class KDKParams:
    G: float
    dt: float
    eps: float
    steps: int
    ckpt_every: int = 64
    device: int = 0


def _system(x0, v0, m, p: KDKParams):
    s = capi.System(p.device)
    s.set_G(p.G)
    s.set_timestep(p.dt)
    s.set_softening(p.eps)
    s.set_precision(capi.PREC_FP64_FAST)
    s.set_iterations(p.steps)
    s.set_bodies(m, x0, v0)
    return s


class KDKEvolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x0, v0, m, params: KDKParams):
        xn = x0.detach().cpu().numpy().astype(np.float64)
        vn = v0.detach().cpu().numpy().astype(np.float64)
        mn = m.detach().cpu().numpy().astype(np.float64)
        s = _system(xn, vn, mn, params)
        s.evolve(capi.LEAPFROG_KDK)
        st = s.state()
        s.close()
        ctx.params = params
        ctx.save_for_backward(x0, v0, m)
        return (torch.from_numpy(st["pos"]).to(x0.device), torch.from_numpy(st["vel"]).to(v0.device))

    @staticmethod
    def backward(ctx, gxT, gvT):
        x0, v0, m = ctx.saved_tensors
        p: KDKParams = ctx.params
        s = _system(x0.detach().cpu().numpy(), v0.detach().cpu().numpy(), m.detach().cpu().numpy(), p)
        gx = np.zeros(x0.shape) if gxT is None else gxT.detach().cpu().numpy().astype(np.float64)
        gv = np.zeros(v0.shape) if gvT is None else gvT.detach().cpu().numpy().astype(np.float64)
        gx0, gv0, gm, _, _ = s.adjoint(p.steps, p.ckpt_every, gx, gv)
        s.close()
        dev = x0.device
        return torch.from_numpy(gx0).to(dev), torch.from_numpy(gv0).to(dev), torch.from_numpy(gm).to(dev), None
