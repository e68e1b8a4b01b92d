"""torch.autograd binding of the backward kernel (jmm_pair_observables).

    U = PairEnergy.apply(x, L, theta, chain, precision)

x: (n_configs, N) float64 CUDA tensor of sorted positions, L: (n_configs,)
float64 CUDA tensor, theta: (2,) float64 tensor of potential parameters
((eps, sigma) for LJ, (A, kappa) for Yukawa).  Returns U per configuration;
its backward gives dU/dx and dU/dtheta from the same kernel launch (FP32 or
FP64 pair arithmetic, FP64 sums).  The kernel runs on torch's current stream.

virial_pressure() and fluctuation_gradient() build the observables the
differentiation path reports: P = (N T + W)/L with W = -sum d phi'(d), and
d<A>/dtheta = <dA/dtheta> - beta (<A dU/dtheta> - <A><dU/dtheta>).
"""
from __future__ import annotations

import dataclasses

import numpy as np
import torch

from .engine import Chain, pair_observables


def _observables(x: torch.Tensor, L: torch.Tensor, chain: Chain, theta, precision: int,
                 want_grad: bool):
    if not (x.is_cuda and L.is_cuda):
        raise ValueError("PairEnergy needs CUDA (HIP) tensors: the backward kernel runs on the GPU")
    if x.dtype != torch.float64 or L.dtype != torch.float64:
        raise ValueError("positions and lengths are float64")
    x = x.contiguous()
    L = L.contiguous()
    n, N = x.shape
    th = [float(t) for t in theta]
    p = list(chain.pot_p) + [0.0] * (4 - len(chain.pot_p))
    p[0], p[1] = th[0], th[1]
    ch = dataclasses.replace(chain, N=N, pot_p=tuple(p), L0=float(L[0].item()) if n else 1.0)
    obs = torch.empty((n, 6), dtype=torch.float64, device=x.device)
    grad = torch.empty_like(x) if want_grad else None
    stream = torch.cuda.current_stream(x.device).cuda_stream
    pair_observables(ch, x.data_ptr(), L.data_ptr(), obs.data_ptr(),
                     grad.data_ptr() if want_grad else None, n, precision, stream)
    return obs, grad


class PairEnergy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, L, theta, chain: Chain, precision: int = 64):
        obs, grad = _observables(x, L, chain, theta.detach().cpu().numpy(), precision,
                                 want_grad=True)
        ctx.save_for_backward(grad, obs[:, 2:4].clone())
        ctx.theta_device = theta.device
        return obs[:, 0].clone()

    @staticmethod
    def backward(ctx, gU):
        grad, dUdth = ctx.saved_tensors
        gx = gU[:, None] * grad
        gth = (gU[:, None] * dUdth).sum(0)
        return gx, None, gth.to(device=ctx.theta_device, dtype=torch.float64), None, None


def observables(x, L, chain: Chain, theta, precision: int = 64):
    """(n, 6) tensor: U, W, dU/dtheta0, dU/dtheta1, dW/dtheta0, dW/dtheta1."""
    obs, _ = _observables(x, L, chain, theta, precision, want_grad=False)
    return obs


def virial_pressure(obs: torch.Tensor, L: torch.Tensor, N: int, T: float):
    """P = (N T + W)/L and dP/dtheta = (dW/dtheta)/L per configuration (the
    1D virial route; hard-core contact terms are not included)."""
    P = (N * T + obs[:, 1]) / L
    dP = obs[:, 4:6] / L[:, None]
    return P, dP


def fluctuation_gradient(A, dA_dtheta, dU_dtheta, beta):
    """d<A>/dtheta = <dA/dtheta> - beta (<A dU/dtheta> - <A><dU/dtheta>),
    over samples (first axis).  Works on numpy arrays or torch tensors."""
    if isinstance(A, torch.Tensor):
        Am = A.mean()
        return dA_dtheta.mean(0) - beta * ((A[:, None] * dU_dtheta).mean(0) - Am * dU_dtheta.mean(0))
    A = np.asarray(A)
    return (np.asarray(dA_dtheta).mean(0)
            - beta * ((A[:, None] * dU_dtheta).mean(0) - A.mean() * np.asarray(dU_dtheta).mean(0)))
