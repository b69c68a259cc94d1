"""torch.autograd.Function over the reverse mode of libpbd.

Rollout.apply(v0, rho0, eps, x0, params, frames, k_ckpt) runs `frames` frames
of the PBF step (Fluid::compute x substeps per frame, src/app/simulation/
Fluid.cu:14-97) from positions x0 and velocities v0 through pbd_adj_forward and
returns the final positions and velocities in the particles' initial order.
Its backward calls pbd_adj_backward: gradients with respect to v0, x0, the rest
density rho0 and the relaxation eps.  Torch only carries the tensors (CPU,
float32); the forward and reverse passes run in the HIP kernels of libpbd.so.
"""
from __future__ import annotations

import numpy as np
import torch

from .capi import Context, PbdParams


def _params_for(params: PbdParams, rho0: float, eps: float) -> PbdParams:
    p = params.copy()
    p.rest_density = float(rho0)
    p.rest_density_from_lattice = 0
    p.relaxation = float(eps)
    p.box_spawn_in_the_beginning = 0
    return p


class Rollout(torch.autograd.Function):
    @staticmethod
    def forward(ctx, v0, rho0, eps, x0, params, frames, k_ckpt):
        x0n = x0.detach().cpu().numpy().astype(np.float32).reshape(-1, 3)
        v0n = v0.detach().cpu().numpy().astype(np.float32).reshape(-1, 3)
        n = x0n.shape[0]
        fluid = Context(_params_for(params, float(rho0), float(eps)), capacity=n)
        pos = np.zeros((n, 4), np.float32)
        pos[:, :3] = x0n
        pos[:, 3] = 1.0
        vel = np.zeros((n, 4), np.float32)
        vel[:, :3] = v0n
        fluid.add_particles(pos, vel)
        fluid.adj_forward(int(frames), int(k_ckpt))
        ids = fluid.adj_ids().astype(np.int64)
        xT = np.zeros((n, 3), np.float32)
        vT = np.zeros((n, 3), np.float32)
        xT[ids] = fluid.download("positions")[:, :3]
        vT[ids] = fluid.download("velocities")[:, :3]
        ctx.fluid = fluid
        ctx.ids = ids
        return torch.from_numpy(xT), torch.from_numpy(vT)

    @staticmethod
    def backward(ctx, gx, gv):
        ids = ctx.ids
        gxf = None if gx is None else gx.detach().cpu().numpy()[ids]
        gvf = None if gv is None else gv.detach().cpu().numpy()[ids]
        g = ctx.fluid.adj_backward(gxf, gvf)
        ctx.fluid.close()
        return (torch.from_numpy(g["v0"]), torch.tensor(g["rho0"], dtype=torch.float32),
                torch.tensor(g["eps"], dtype=torch.float32), torch.from_numpy(g["x0"]), None, None, None)


def rollout(v0, rho0, eps, x0, params: PbdParams, frames: int, k_ckpt: int = 1):
    """(x_T, v_T) in the initial particle order; differentiable in v0, rho0, eps and x0."""
    return Rollout.apply(v0, rho0, eps, x0, params, frames, k_ckpt)
