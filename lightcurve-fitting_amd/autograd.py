"""``torch.autograd.Function`` wrappers over the adjoint kernels.

``SviElbo.apply(mu, log_sigma, gpu_ctx, data, model_ids, prior_c, prior_w,
n_samples, seed, step)`` returns one Monte-Carlo ELBO estimate per job (the
objective of one step of the batched SVI, with that step's Philox noise) and
back-propagates through the hand-written adjoint kernel: the kernel computes
the ELBO and its exact gradient w.r.t. (mu, log sigma) in one fused pass, the
forward stores the gradient, and backward scales it by the incoming gradient.
Tensors are [n_jobs, 7] float64 (rows padded like the C ABI); they cross the
C ABI through host memory.

``Gp2dNlml.apply(theta, gpu_ctx, csr)`` returns the negative log marginal
likelihood of each source's 2D GP (separable squared-exponential kernel,
theta = ln(amp, l_t, l_lambda, sigma_n) as [n_sources, 4] float64) and
back-propagates through the Cholesky adjoint kernel (``lcf_gp2d_nlml_grad``:
1/2 tr((K^-1 - alpha alpha^T) dK/dtheta) from one blocked FP64 factorisation);
sources whose K is not positive definite give NaN and a NaN gradient.

Imported on demand (it imports torch); nothing else in the package needs it.
"""
from __future__ import annotations

import numpy as np
import torch


def _np(t: torch.Tensor) -> np.ndarray:
    return np.ascontiguousarray(t.detach().to("cpu", torch.float64).numpy())


class SviElbo(torch.autograd.Function):
    @staticmethod
    def forward(ctx, mu, log_sigma, gpu_ctx, data, model_ids, prior_c, prior_w, n_samples,
                seed=42, step=0, flags=0):
        elbo, gmu, gls = gpu_ctx.svi_elbo_grad(data, model_ids, _np(mu), _np(log_sigma), _np(prior_c),
                                               _np(prior_w), int(n_samples), seed=int(seed),
                                               step=int(step), flags=int(flags))
        ctx.save_for_backward(torch.from_numpy(gmu).to(mu.device, mu.dtype),
                              torch.from_numpy(gls).to(log_sigma.device, log_sigma.dtype))
        return torch.from_numpy(elbo).to(mu.device, mu.dtype)

    @staticmethod
    def backward(ctx, grad_out):
        gmu, gls = ctx.saved_tensors
        g = grad_out.reshape(-1, 1)
        return g * gmu, g * gls, None, None, None, None, None, None, None, None, None


class Gp2dNlml(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta, gpu_ctx, csr):
        r = gpu_ctx.gp2d_nlml_grad(csr, _np(theta))
        ctx.save_for_backward(torch.from_numpy(r["grad"]).to(theta.device, theta.dtype))
        return torch.from_numpy(r["nlml"]).to(theta.device, theta.dtype)

    @staticmethod
    def backward(ctx, grad_out):
        (g,) = ctx.saved_tensors
        return grad_out.reshape(-1, 1) * g, None, None
