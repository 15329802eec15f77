"""torch.autograd functions over the sweep and the scattering solve.  Forward passes and
vector-Jacobian products both run in libtsw.so (tsw_sweep_apply / tsw_sweep_vjp and
tsw_gmres_solve / tsw_solve_vjp); tensors are float64 CUDA tensors in the reference layouts:
  src, psi  [nx+2][ny+2][nq][p+1][p+1]  (halo cells hold the inflow)
  sigma     [nx+2][ny+2]                 (cell means)
The reference has no adjoint (its SweepOperator::Apply, SweepOperator.cpp:118-135, is the forward map
differentiated here).
"""
from __future__ import annotations

import math

import torch

from . import FP64


class Sweep(torch.autograd.Function):
    """psi = L^{-1} src with inflow from src's halo cells; differentiable in src and sigma_t."""

    @staticmethod
    def forward(fctx, src, sigma_t, ctx, dt=math.inf, precision=FP64):
        sigma_t = sigma_t.detach().contiguous()
        psi = src.detach().clone(memory_format=torch.contiguous_format)
        ctx.sweep(psi, sigma_t, dt=dt, precision=precision)
        fctx.save_for_backward(psi, sigma_t)
        fctx.args = (ctx, dt, precision)
        return psi

    @staticmethod
    def backward(fctx, grad_psi):
        psi, sigma_t = fctx.saved_tensors
        ctx, dt, precision = fctx.args
        seed = grad_psi.contiguous()
        gsrc = torch.empty_like(psi)
        gst = torch.empty_like(sigma_t) if fctx.needs_input_grad[1] else None
        ctx.sweep_vjp(seed, gsrc, sigma_t, psi=psi if gst is not None else None, grad_sigma_t=gst, dt=dt,
                      precision=precision)
        return gsrc, gst, None, None, None


class Solve(torch.autograd.Function):
    """psi = (L - sigma_s P / 4pi)^{-1} q by Peierls GMRES; differentiable in q (source and inflow halos),
    sigma_t and sigma_s.  The backward pass solves the adjoint problem with the same GMRES."""

    @staticmethod
    def forward(fctx, q, sigma_t, sigma_s, ctx, dt=math.inf, abs_tol=1e-12, adj_tol=1e-12, restart=30,
                precision=FP64):
        sigma_t = sigma_t.detach().contiguous()
        sigma_s = sigma_s.detach().contiguous()
        psi = q.detach().clone(memory_format=torch.contiguous_format)
        ctx.gmres_solve(psi, sigma_t, sigma_s, dt=dt, abs_tol=abs_tol, restart=restart, precision=precision)
        fctx.save_for_backward(psi, sigma_t, sigma_s)
        fctx.args = (ctx, dt, adj_tol, restart, precision)
        return psi

    @staticmethod
    def backward(fctx, grad_psi):
        psi, sigma_t, sigma_s = fctx.saved_tensors
        ctx, dt, adj_tol, restart, precision = fctx.args
        seed = grad_psi.contiguous()
        gsrc = torch.empty_like(psi)
        gst = torch.empty_like(sigma_t) if fctx.needs_input_grad[1] else None
        gss = torch.empty_like(sigma_s) if fctx.needs_input_grad[2] else None
        need_psi = gst is not None or gss is not None
        ctx.solve_vjp(seed, gsrc, sigma_t, sigma_s, psi=psi if need_psi else None, grad_sigma_t=gst,
                      grad_sigma_s=gss, dt=dt, abs_tol=adj_tol, restart=restart, precision=precision)
        return gsrc, gst, gss, None, None, None, None, None, None


def sweep(src, sigma_t, ctx, dt=math.inf, precision=FP64):
    return Sweep.apply(src, sigma_t, ctx, dt, precision)


def solve(q, sigma_t, sigma_s, ctx, dt=math.inf, abs_tol=1e-12, adj_tol=1e-12, restart=30, precision=FP64):
    return Solve.apply(q, sigma_t, sigma_s, ctx, dt, abs_tol, adj_tol, restart, precision)
