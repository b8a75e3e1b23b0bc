"""torch.autograd front end of the batched spectra and their parameter derivatives.

``lensed_cl(model, theta)`` returns the observed spectra (lensed TT, TE, EE on
l = 2..l_max, or the unlensed ones without lensing) of every row of
``theta`` ([n][7]: omega_b, omega_cdm, h, ln 1e10 A_s, n_s, tau_reio, sum m_nu)
as a [n][3][n_l] tensor that takes part in torch's autograd: any loss built
from it with torch operations back-propagates to ``theta``.  The reference gets
the same from ``jax.grad`` through ``Model.__call__`` (abcmb/main.py:165-184).

The forward pass runs ``tangent_gradients``' sequence (tangent-linear ODE and
LOS launches, the lensing differenced along each parameter on the moved
spectra) and keeps the Jacobian d C / d theta ([n][7][3][n_l], on the GPU);
the backward pass is its transpose applied to the incoming cotangent, one
contraction.  Cotangents are returned for ``theta`` only (the host tables'
derivatives are not formed).
"""
from __future__ import annotations

import numpy as np

from . import likelihood as LK


class _NoLoss:
    """Stands in for a likelihood: the spectra and their Jacobian are the outputs."""

    def loglike_cl(self, cl, mock_index):
        import torch
        return torch.zeros(cl.shape[0], dtype=cl.dtype, device=cl.device)


def _function():
    import torch

    class LensedCl(torch.autograd.Function):
        @staticmethod
        def forward(ctx, theta, model, k_grids, chunk):
            th = np.ascontiguousarray(theta.detach().cpu().numpy(), dtype=np.float64)
            cl_out, jac_out = [], []
            LK.tangent_gradients(model, _NoLoss(), th, k_grids=k_grids, chunk=chunk, cl_out=cl_out,
                                 jac_out=jac_out)
            jac = torch.cat(jac_out)
            ctx.save_for_backward(jac)
            ctx.theta_device, ctx.theta_dtype = theta.device, theta.dtype
            return torch.cat(cl_out).to(theta.device)

        @staticmethod
        def backward(ctx, grad_cl):
            (jac,) = ctx.saved_tensors
            g = torch.einsum("nxl,npxl->np", grad_cl.to(jac.device, jac.dtype), jac)
            return g.to(ctx.theta_device, ctx.theta_dtype), None, None, None

    return LensedCl


_LensedCl = None


def lensed_cl(model, theta, k_grids=None, chunk: int = 64):
    """Observed spectra of each row of ``theta`` ([n][7] tensor, float64), differentiable
    with respect to ``theta``; ``k_grids``: one perturbation k grid per row (default:
    the model's)."""
    global _LensedCl
    if _LensedCl is None:
        _LensedCl = _function()
    if theta.dim() != 2 or theta.shape[1] != len(LK.NAMES):
        raise ValueError(f"theta must be [n][{len(LK.NAMES)}] ({', '.join(LK.NAMES)})")
    return _LensedCl.apply(theta, model, k_grids, chunk)
