"""torch.autograd binding of the differentiable LJ trajectory.

TrajectoryLoss.apply(x0, eps, sigma, system, v0, dt, nsteps, checkpoint_every,
p_target, w_p) returns the scalar loss
    L = <E_pot>_t + w_p (<P>_t - p_target)^2
of `nsteps` velocity-Verlet (EVN) steps of `system` (an ljamd.LJSystem) from
positions x0 (float32 [N,4] on the device) and velocities v0.  Its backward
calls the library's adjoint sweep (lj_trajectory_grad: one hand-written
Hessian-vector-product pass per state, checkpoints every `checkpoint_every`
steps) and returns dL/dx0, dL/deps, dL/dsigma.  Device tensors travel by
pointer over the C-ABI on torch's current stream; nothing is computed in torch.
"""
import ctypes as C

import numpy as np
import torch

from ._lib import ARRAY_POS, ARRAY_VEL, check, lib


def _set_state(system, x0, v0):
    st = C.c_void_p(torch.cuda.current_stream(x0.device).cuda_stream)
    check(lib().lj_set_array_device(system.h, ARRAY_POS, C.c_void_p(x0.data_ptr()), st))
    check(lib().lj_set_array_device(system.h, ARRAY_VEL, C.c_void_p(v0.data_ptr()), st))


class TrajectoryLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x0, eps, sigma, system, v0, dt, nsteps, checkpoint_every=20, p_target=1.0,
                w_p=1.0):
        if x0.dtype != torch.float32 or not x0.is_cuda or x0.shape != (system.N, 4):
            raise ValueError("x0 must be a float32 [N,4] device tensor")
        x0c = x0.contiguous()
        v0c = v0.contiguous().to(torch.float32)
        system.set_lj_params(float(eps), float(sigma))
        _set_state(system, x0c, v0c)
        out = system.trajectory_loss(dt, nsteps, p_target, w_p)
        ctx.system = system
        ctx.args = (dt, nsteps, checkpoint_every, p_target, w_p, float(eps), float(sigma))
        ctx.save_for_backward(x0c, v0c)
        return torch.tensor(out[0], dtype=torch.float64, device=x0.device)

    @staticmethod
    def backward(ctx, g):
        x0, v0 = ctx.saved_tensors
        dt, nsteps, k, pt, wp, eps, sigma = ctx.args
        system = ctx.system
        system.set_lj_params(eps, sigma)
        _set_state(system, x0, v0)
        dx0 = torch.zeros_like(x0)
        torch.cuda.current_stream(x0.device).synchronize()
        _, g2, _, _ = system.trajectory_grad(dt, nsteps, k, pt, wp, d_dx0=dx0.data_ptr())
        gs = g.to(torch.float64)
        dx = (dx0.to(torch.float64) * gs).to(torch.float32)
        dx[:, 3] = 0
        geps = gs * float(g2[0])
        gsig = gs * float(g2[1])
        return dx, geps, gsig, None, None, None, None, None, None, None


def trajectory_loss(x0, eps, sigma, system, v0, dt=0.005, nsteps=200, checkpoint_every=20,
                    p_target=1.0, w_p=1.0):
    return TrajectoryLoss.apply(x0, eps, sigma, system, v0, dt, nsteps, checkpoint_every,
                                p_target, w_p)
