"""torch.autograd front end of the batched spectra and their parameter derivatives.

``lensed_cl(model, theta)`` returns the observed spectra (lensed TT, TE, EE on
l = 2..l_max, or the unlensed ones without lensing) of every row of
``theta`` ([n][7]: omega_b, omega_cdm, h, ln 1e10 A_s, n_s, tau_reio, sum m_nu)
as a [n][3][n_l] tensor that takes part in torch's autograd: any loss built
from it with torch operations back-propagates to ``theta``.  The reference gets
the same from ``jax.grad`` through ``Model.__call__`` (abcmb/main.py:165-184).

``mode="adjoint"`` (default): the forward pass is one primal ODE launch and
the downstream (``likelihood.adjoint_forward``); the backward pass is reverse
mode on the GPU (``likelihood.adjoint_backward``): the lensing, l-spline,
LOS and Limber adjoints, the discrete-adjoint ODE kernel (checkpoints every 64
accepted steps, transposed stage solves), and the contraction of the host
inputs' cotangents with their derivatives along the parameters.  The
backward also leaves those cotangents (cosmology fields, tau / x_e / T_m
tables, LOS background vectors, C_l^phiphi) in ``host_cotangents`` of the
returned tensor's function node, for callers that chain them through their
own background (the reference's jax.vjp).

``mode="jacobian"``: the forward pass runs ``tangent_gradients``' sequence
(tangent-linear ODE and LOS, the lensing differenced along each parameter)
and keeps d C / d theta ([n][7][3][n_l]); the backward is its transpose
applied to the cotangent, one contraction.

``spectra(model, values, names, params)``: one cosmology's spectra as a
function of any named input parameters (user fluids' own included), through
``Model.cl_jacobian``.
"""
from __future__ import annotations

import numpy as np

from . import likelihood as LK


class _NoLoss:
    """Stands in for a likelihood: the spectra and their Jacobian are the outputs."""

    def loglike_cl(self, cl, mock_index):
        import torch
        return torch.zeros(cl.shape[0], dtype=cl.dtype, device=cl.device)


def _functions():
    import torch

    class LensedClJacobian(torch.autograd.Function):
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

    class LensedClAdjoint(torch.autograd.Function):
        @staticmethod
        def forward(ctx, theta, model, k_grids, chunk):
            th = np.ascontiguousarray(theta.detach().cpu().numpy(), dtype=np.float64)
            ctx.model, ctx.chunk = model, max(1, chunk)
            ctx.theta_device, ctx.theta_dtype = theta.device, theta.dtype
            k_grids = list(k_grids) if k_grids is not None else [None] * th.shape[0]
            ctx.states, obs = [], []
            for a in range(0, th.shape[0], ctx.chunk):
                st, o = LK.adjoint_forward(model, th[a:a + ctx.chunk], k_grids[a:a + ctx.chunk])
                ctx.states.append(st)
                obs.append(o.detach().clone())
            ctx.host_cotangents = []
            return torch.cat(obs).to(theta.device)

        @staticmethod
        def backward(ctx, grad_cl):
            grads, a = [], 0
            for st in ctx.states:
                n = st["thetas"].shape[0]
                gb = grad_cl[a:a + n].to("cuda", torch.float64).contiguous()
                grads.append(LK.adjoint_backward(ctx.model, st, gb, cotangents_out=ctx.host_cotangents))
                a += n
            ctx.states = None
            g = torch.as_tensor(np.concatenate(grads))
            return g.to(ctx.theta_device, ctx.theta_dtype), None, None, None

    class SpectraJacobian(torch.autograd.Function):
        @staticmethod
        def forward(ctx, values, model, names, params):
            v = values.detach().cpu().numpy().astype(np.float64)
            p = dict(params)
            p.update({n: float(x) for n, x in zip(names, v)})
            J = model.cl_jacobian(p, names)
            pre = "lensed_" if model.specs["lensing"] else ""
            obs = J["lensed"][:3] if pre else J["cl"]
            ctx.save_for_backward(torch.as_tensor(np.stack([J[pre + n][:3] for n in names])))
            ctx.values_device, ctx.values_dtype = values.device, values.dtype
            return torch.as_tensor(obs).to(values.device, values.dtype)

        @staticmethod
        def backward(ctx, grad_cl):
            (jac,) = ctx.saved_tensors                                    # [n_names][3][n_l]
            g = torch.tensordot(jac, grad_cl.to(jac.device, jac.dtype), dims=([1, 2], [0, 1]))
            return g.to(ctx.values_device, ctx.values_dtype), None, None, None

    return {"jacobian": LensedClJacobian, "adjoint": LensedClAdjoint, "spectra": SpectraJacobian}


_FUNCS = None


def lensed_cl(model, theta, k_grids=None, chunk: int = 64, mode: str = "adjoint"):
    """Observed spectra of each row of ``theta`` ([n][7] tensor, float64), differentiable
    with respect to ``theta``; ``k_grids``: one perturbation k grid per row (default:
    the model's); ``mode``: "adjoint" (reverse mode on the GPU in the backward
    pass) or "jacobian" (the forward-mode Jacobian kept from the forward pass)."""
    global _FUNCS
    if _FUNCS is None:
        _FUNCS = _functions()
    if mode not in ("jacobian", "adjoint"):
        raise ValueError("mode must be one of ('jacobian', 'adjoint')")
    if theta.dim() != 2 or theta.shape[1] != len(LK.NAMES):
        raise ValueError(f"theta must be [n][{len(LK.NAMES)}] ({', '.join(LK.NAMES)})")
    return _FUNCS[mode].apply(theta, model, k_grids, chunk)


def spectra(model, values, names, params: dict):
    """Observed spectra of one cosmology ([3][n_l]: lensed TT, TE, EE on
    l = 2..l_max, or the unlensed ones without lensing), differentiable with
    respect to ``values`` (a 1-D tensor): the values of the parameters ``names``
    -- any inputs that reach the perturbations, at most five, the user fluids'
    own included (``Model.cl_jacobian``) -- the other parameters taken from
    ``params``.  The backward pass contracts the cotangent with the forward-mode
    Jacobian kept from the forward pass."""
    global _FUNCS
    if _FUNCS is None:
        _FUNCS = _functions()
    names = list(names)
    if values.dim() != 1 or values.shape[0] != len(names):
        raise ValueError(f"values must be a 1-D tensor of {len(names)} entries, one per name")
    return _FUNCS["spectra"].apply(values, model, names, params)
