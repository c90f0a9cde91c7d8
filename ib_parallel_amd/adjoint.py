"""The discrete adjoint of the GPU time step as a torch.autograd.Function.

``ForceObjective.apply(re, omega0, params, bodies, nsteps, cx, cy)`` runs
``nsteps`` steps of the bodies and Schur form ``params``/``bodies`` give from the vorticity
``omega0`` (mgridlev blocks of (m-1)(n-1); the fluxes are vort2flux of it)
at Reynolds number ``re`` and returns the time-mean weighted force

    J = mean over the steps of sum_b cx[b] fx_b + cy[b] fy_b

(fx, fy as ib_get_forces reports them).  ``backward`` hands out dJ/dre and
dJ/domega0 from ib_adjoint_run, which runs the transposed step operators
backwards from checkpoints; the forward and the adjoint are one library call.
"""
import numpy as np
import torch

from . import GpuSim, abi


def _with_re(params, re):
    p = abi.IBParams.from_buffer_copy(params)
    p.re = float(re)
    return p


def start_state(sim, omega0, rhs_old=None):
    """Load omega0 (and the fluxes vort2flux gives it) into sim as its state."""
    w, s, qx, qy = sim.vort2flux(np.ascontiguousarray(omega0, dtype=np.float64))
    P, Fx, Fy = sim.P, sim.Fx, sim.Fy
    for k in range(sim.L):
        ro = None if rhs_old is None else rhs_old[k * P:(k + 1) * P]
        sim.set_level(k + 1, omega=w[k * P:(k + 1) * P], s=s[k * P:(k + 1) * P], qx=qx[k * Fx:(k + 1) * Fx],
                      qy=qy[k * Fy:(k + 1) * Fy], rhs_old=ro)


class ForceObjective(torch.autograd.Function):
    @staticmethod
    def forward(ctx, re, omega0, params, bodies, nsteps, cx, cy, ckpt_every=None, rhs_old=None):
        sim = GpuSim(_with_re(params, re.item()), bodies)
        try:
            start_state(sim, omega0.detach().cpu().numpy().astype(np.float64), rhs_old)
            J, dre, dw = sim.adjoint(nsteps, cx, cy, ckpt_every=ckpt_every, omega_grad=True)
        finally:
            sim.close()
        ctx.save_for_backward(torch.tensor(dre, dtype=torch.float64), torch.from_numpy(dw))
        ctx.shapes = (re.shape, omega0.shape, omega0.dtype, omega0.device)
        return torch.tensor(J, dtype=torch.float64)

    @staticmethod
    def backward(ctx, gJ):
        dre, dw = ctx.saved_tensors
        re_shape, w_shape, w_dtype, w_dev = ctx.shapes
        g_re = (gJ * dre).reshape(re_shape)
        g_w = (gJ * dw).reshape(w_shape).to(dtype=w_dtype, device=w_dev)
        return g_re, g_w, None, None, None, None, None, None, None
