"""Gradients of a disk simulation through the hand-written adjoint.

``DiskEvolve`` is a ``torch.autograd.Function``: forward runs ``nsteps`` of
the device loop from an initial surface density with a given planet mass and
viscosity alpha (checkpointing every ``ckpt_every`` steps; 0 keeps every
step's intermediates in device memory and recomputes nothing) and returns the
final surface density; backward runs the adjoint of every step
(pg_adjoint_backward) and returns dL/dSigma0, dL/dq, dL/dalpha.  The step
sizes of the forward pass are frozen in the backward pass.

``gap_loss`` is the loss of the gap-opening workload: the L2 distance of the
azimuthally averaged final Sigma(r) to a target profile with a Gaussian gap.
"""
from __future__ import annotations

import numpy as np
import torch

from .disk import Disk
from .params import Params


def radial_centres(p: Params) -> np.ndarray:
    """Cell centres 0.5 (x_a[i] + x_a[i+1]) of the log radial grid (gridx 1)
    or uniform grid (gridx 0), as the library builds them."""
    n = int(p.xres)
    i = np.arange(n + 1, dtype=np.float64)
    if int(p.gridx) == 1:
        xa = p.xmin * np.exp(np.log(p.xmax / p.xmin) * i / n)
    else:
        xa = p.xmin + (p.xmax - p.xmin) * i / n
    return 0.5 * (xa[1:] + xa[:-1])


def gap_profile(p: Params, depth: float = 0.5, width_h: float = 2.0) -> np.ndarray:
    r = radial_centres(p)
    w = width_h * p.sc_h
    return (r / p.planet_radius) ** (-p.p_alpha) * (1.0 - depth * np.exp(-(r - 1.0) ** 2 / (2.0 * w * w)))


def gap_loss(sigma: torch.Tensor, target: torch.Tensor) -> torch.Tensor:
    """sum_i (<Sigma>_phi(r_i) - Sigma_t(r_i))^2 for sigma of shape (nphi, nr)."""
    return torch.sum((sigma.mean(dim=0) - target) ** 2)


class DiskEvolve(torch.autograd.Function):
    """Sigma_final = evolve(Sigma0, q, alpha); backward by the adjoint."""

    @staticmethod
    def forward(ctx, sigma0: torch.Tensor, q: torch.Tensor, alpha: torch.Tensor, params: Params,
                nsteps: int, ckpt_every: int = 20):
        p = params.replace(planet_mass=float(q), ss_alpha=float(alpha))
        d = Disk(p)
        gas, _ = d.get_state()
        gas[0] = sigma0.detach().cpu().double().numpy()
        d.set_state(gas)
        d.adjoint_forward(nsteps, ckpt_every)
        out, _ = d.get_state()
        if any(ctx.needs_input_grad[:3]):
            # the handle holds the checkpoints (or, with ckpt_every 0, every
            # step's intermediates) until backward runs or the graph that
            # owns ctx is freed, whichever comes first
            ctx.disk = d
        else:
            d.close()
            ctx.disk = None
        return torch.from_numpy(out[0].copy()).to(sigma0.dtype)

    @staticmethod
    def backward(ctx, g_sigma: torch.Tensor):
        d = ctx.disk
        if d is None:
            raise RuntimeError("DiskEvolve.backward: the forward pass kept no adjoint workspace "
                               "(no input required a gradient), or backward already ran")
        dL = np.zeros((5, d.ny, d.nx))
        dL[0] = g_sigma.detach().cpu().double().numpy()
        dq, dal, d0 = d.adjoint_backward(dL)
        d.close()
        ctx.disk = None
        return (torch.from_numpy(d0[0].copy()).to(g_sigma.dtype), torch.tensor(dq, dtype=g_sigma.dtype),
                torch.tensor(dal, dtype=g_sigma.dtype), None, None, None)


def evolve(sigma0: torch.Tensor, q: torch.Tensor, alpha: torch.Tensor, params: Params, nsteps: int,
           ckpt_every: int = 20) -> torch.Tensor:
    return DiskEvolve.apply(sigma0, q, alpha, params, nsteps, ckpt_every)


__all__ = ["DiskEvolve", "evolve", "gap_loss", "gap_profile", "radial_centres"]
