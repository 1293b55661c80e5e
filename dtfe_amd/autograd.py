"""Differentiable DTFE: a torch.autograd.Function around the gfx950 kernels.

Forward: vertex densities (rho_v = 4 w_v / (rho_bar W_v)) and the sampled
fields of `api.sample`. Backward: the hand-written adjoint kernels
(`api.adjoint`), which re-locate the sample points (starting from the cells
the forward located) and recompute their barycentric coordinates. The
tessellation is held fixed (gradients at fixed topology), as are the
Monte-Carlo sample points.

`dtfe_fields` is the particle-level entry: periodic images are gathered from
the particles with an index map, so torch folds the image gradients back to
the particles; rho_bar = sum(w) / 35000^3 (the reference's normalisation) is
part of the graph.
"""
from __future__ import annotations

from typing import Dict, Optional, Sequence

import torch

from . import api
from .mesh import REFERENCE_BOX_VOLUME


class _DTFESample(torch.autograd.Function):
    @staticmethod
    def forward(ctx, pos, weight, vel, rho_bar_t, mesh_cells, mesh_nbrs, spec, fields, method, samples, seed,
                value_bug):
        mesh = api.DeviceMesh(pos=pos.detach().contiguous(), cells=mesh_cells, nbrs=mesh_nbrs)
        if mesh.pos.shape[0] >= 4 and mesh.cells.shape[0] > 0:
            mesh.build_locator()  # these positions, for the forward and its backward
        rho_bar = float(rho_bar_t.item())
        rho = api.vertex_density(mesh, weight.detach(), rho_bar)
        out = api.sample(mesh, rho, None if vel is None else vel.detach(), None, fields, spec, method=method,
                         samples=samples, seed=seed, value_bug=value_bug, want_cells=True)
        ctx.save_for_backward(pos, weight, rho, *(() if vel is None else (vel,)))
        ctx.locator = mesh.locator
        ctx.located = out["located_cell"]
        ctx.has_vel = vel is not None
        ctx.meta = (mesh_cells, mesh_nbrs, spec, list(fields), method, samples, seed, value_bug, rho_bar)
        return tuple(out[f] for f in fields)

    @staticmethod
    def backward(ctx, *grads):
        saved = ctx.saved_tensors
        pos, weight, rho = saved[:3]
        vel = saved[3] if ctx.has_vel else None
        cells, nbrs, spec, fields, method, samples, seed, value_bug, rho_bar = ctx.meta
        mesh = api.DeviceMesh(pos=pos.detach().contiguous(), cells=cells, nbrs=nbrs, locator=ctx.locator)
        upstream = {f: g for f, g in zip(fields, grads) if g is not None}
        g, g_rho_bar = api.adjoint(mesh, rho, weight.detach(), None if vel is None else vel.detach(), rho_bar, spec,
                                   upstream, method=method, samples=samples, seed=seed, value_bug=value_bug,
                                   start_cells=ctx.located)
        g_pos = g.get("position")
        g_w = g.get("weight")
        g_v = g.get("velocity") if ctx.has_vel else None
        g_rb = torch.tensor(g_rho_bar, dtype=torch.float64, device=pos.device)
        return (None if g_pos is None else g_pos.to(pos.dtype), g_w, g_v, g_rb, None, None, None, None, None, None,
                None, None)


def dtfe_sample(pos: torch.Tensor, weight: torch.Tensor, vel: Optional[torch.Tensor], rho_bar: torch.Tensor,
                cells: torch.Tensor, nbrs: torch.Tensor, spec, fields: Sequence[str], method: int = 0,
                samples: int = 1, seed: int = 0, value_bug: bool = True):
    """Differentiable sampling on a fixed mesh (vertex-level inputs).

    pos [nv, 3] float64, weight [nv], vel [nv, 3] (field dtype), rho_bar a
    0-d float64 tensor. Returns a tuple of the requested output fields."""
    return _DTFESample.apply(pos, weight, vel, rho_bar, cells, nbrs, spec, tuple(fields), method, samples, seed,
                             value_bug)


def dtfe_fields(particle_pos: torch.Tensor, particle_weight: torch.Tensor, particle_vel: Optional[torch.Tensor],
                image_of: torch.Tensor, image_shift: torch.Tensor, cells: torch.Tensor, nbrs: torch.Tensor, spec,
                fields: Sequence[str], method: int = 0, samples: int = 1, seed: int = 0,
                value_bug: bool = True) -> Dict[str, torch.Tensor]:
    """Particle-level differentiable DTFE on a fixed tessellation of the
    padded (periodic-image) vertex set.

    image_of [nv] maps each mesh vertex to its particle; image_shift [nv, 3]
    is the vertex position minus the particle position (a multiple of the box
    length). rho_bar = sum(particle weights) / 35000^3."""
    pos = particle_pos[image_of] + image_shift
    w = particle_weight[image_of]
    v = None if particle_vel is None else particle_vel[image_of]
    rho_bar = particle_weight.to(torch.float64).sum() / REFERENCE_BOX_VOLUME
    outs = dtfe_sample(pos, w, v, rho_bar, cells, nbrs, spec, fields, method, samples, seed, value_bug)
    return dict(zip(fields, outs))
