"""Differentiable waveform misfit: a ``torch.autograd.Function`` over the GPU time loop.

``waveform_misfit(rho, vp, vs, sim, obs)`` maps the MODEL=gll parameters of a simulation (per GLL point,
[nspec, 25] tensors) to chi = 1/2 dt sum_n sum_r |s_n - o_n|^2, with s the displacement seismograms of a run
from rest.  The forward pass runs the elastic time loop on the GPU and keeps a checkpoint of the wavefield every
``chunk`` steps in device memory; the backward pass recomputes each segment from its checkpoint and runs the
hand-written reverse-mode kernels (the exact discrete adjoint of the Newmark loop, not a backward
reconstruction), accumulating the gradient in FP64.  The reference's SIMULATION_TYPE=3 kernels are the same
gradient in FP64 (K_kappa = kappa g_kappa / (w J), K_rho = rho g_rho / (w J), and K_mu up to the reference's
3D deviator convention).

One simulation holds the checkpoints of one forward pass: if another forward ran on the same simulation
before this backward, the backward repeats its own forward first.
"""
from __future__ import annotations

import itertools

import numpy as np
import torch

from .driver import Simulation

_tokens = itertools.count(1)


def _host(t: torch.Tensor) -> np.ndarray:
    return t.detach().to("cpu", torch.float64).contiguous().numpy().reshape(-1)


class WaveformMisfit(torch.autograd.Function):
    @staticmethod
    def forward(ctx, rho, vp, vs, sim: Simulation, obs, chunk: int = 100):
        model = tuple(_host(t) for t in (rho, vp, vs))
        obs_np = np.asarray(obs.detach().cpu().numpy() if torch.is_tensor(obs) else obs, np.float64)
        sim.set_model(*model)
        chi, _ = sim.gradient_forward(obs_np, chunk)
        token = next(_tokens)
        sim._grad_token = token
        ctx.sim, ctx.model, ctx.obs, ctx.chunk, ctx.token = sim, model, obs_np, chunk, token
        ctx.meta = [(t.shape, t.dtype, t.device) for t in (rho, vp, vs)]
        return torch.tensor(chi, dtype=torch.float64, device=rho.device)

    @staticmethod
    def backward(ctx, gchi):
        sim = ctx.sim
        if getattr(sim, "_grad_token", None) != ctx.token:
            sim.set_model(*ctx.model)
            sim.gradient_forward(ctx.obs, ctx.chunk)
            sim._grad_token = ctx.token
        g = sim.gradient_backward()
        scale = float(gchi)
        outs = []
        for name, (shape, dtype, device) in zip(("rho", "vp", "vs"), ctx.meta):
            outs.append(torch.as_tensor(g[name] * scale, dtype=torch.float64).reshape(shape).to(device, dtype))
        return outs[0], outs[1], outs[2], None, None, None


def waveform_misfit(rho: torch.Tensor, vp: torch.Tensor, vs: torch.Tensor, sim: Simulation, obs,
                    chunk: int = 100) -> torch.Tensor:
    """chi(rho, vp, vs) as a differentiable FP64 scalar; see the module docstring."""
    return WaveformMisfit.apply(rho, vp, vs, sim, obs, chunk)
