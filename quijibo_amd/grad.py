"""Differentiation of the Julia field with respect to the root sets.

JuliaField: a torch.autograd.Function whose forward evaluates F = log|f^k(q0)|
at points (qj_points_eval) and whose backward is the hand-written adjoint
kernel (qj_points_adjoint), giving dJ/d(root positions) and dJ/d(powers).

centered_root_gradient: the reference's central-difference root-position
gradient (OPTIMIZE_3D::computeLogRationalCenteredGradient,
OPTIMIZE_3D.cpp:2872-2951: eps per component, entries [top 4 x nt, bottom 4 x nb])
for any scalar objective evaluated on the GPU.
"""
import numpy as np
import torch

from . import api
from ._lib import FP64


class JuliaField(torch.autograd.Function):
    """F(points) for roots/powers given as float64 tensors; differentiable in
    top_roots, top_powers, bottom_roots, bottom_powers."""

    @staticmethod
    def forward(ctx, points, top_roots, top_powers, bottom_roots, bottom_powers, params, precision=FP64,
                band=api.DEFAULT_BAND, top_scalar=1.0, bottom_scalar=1.0):
        top = api.POLYNOMIAL_4D(top_roots.detach().cpu().numpy(), top_powers.detach().cpu().numpy(), top_scalar)
        bot = api.POLYNOMIAL_4D(bottom_roots.detach().cpu().numpy(), bottom_powers.detach().cpu().numpy(),
                                bottom_scalar)
        r = api.Rational(top, bot)
        F = api.points_eval(points.detach(), r, params, precision, band)
        ctx.save_for_backward(points)
        ctx.rational = r
        ctx.params = params
        ctx.precision = precision
        ctx.nt = top.totalRoots()
        ctx.devices = (top_roots.device, top_powers.device, bottom_roots.device, bottom_powers.device)
        ctx.dtypes = (top_roots.dtype, top_powers.dtype, bottom_roots.dtype, bottom_powers.dtype)
        return F

    @staticmethod
    def backward(ctx, gF):
        (points,) = ctx.saved_tensors
        g = torch.where(torch.isfinite(gF), gF, torch.zeros_like(gF))
        gr, gp = api.points_adjoint(points, ctx.rational, ctx.params, g, ctx.precision)
        nt = ctx.nt
        outs = [gr[:nt], gp[:nt], gr[nt:], gp[nt:]]
        outs = [o.to(device=d, dtype=t) for o, d, t in zip(outs, ctx.devices, ctx.dtypes)]
        return (None, outs[0], outs[1], outs[2], outs[3], None, None, None, None, None)


def julia_field(points, top_roots, top_powers, bottom_roots, bottom_powers, params, precision=FP64,
                band=api.DEFAULT_BAND, top_scalar=1.0, bottom_scalar=1.0):
    return JuliaField.apply(points, top_roots, top_powers, bottom_roots, bottom_powers, params, precision, band,
                            top_scalar, bottom_scalar)


def centered_root_gradient(objective, top, bottom, eps):
    """Central differences of objective(rational) over every root component.

    objective: callable taking an api.Rational and returning a float.
    Returns a numpy vector [top 4 x nt, bottom 4 x nb] (OPTIMIZE_3D.cpp:2872-2951)."""
    nt, nb = top.totalRoots(), bottom.totalRoots()
    grad = np.zeros(4 * (nt + nb))
    r = api.Rational(top, bottom)
    for which, poly, off in ((0, top, 0), (1, bottom, 4 * nt)):
        base = poly.roots.copy()
        for x in range(poly.totalRoots()):
            for y in range(4):
                vals = []
                for sgn in (1.0, -1.0):
                    poly.roots = base.copy()
                    poly.roots[x, y] += sgn * eps
                    r.update(top, bottom)
                    vals.append(objective(r))
                grad[off + 4 * x + y] = (vals[0] - vals[1]) / (2.0 * eps)
        poly.roots = base
    r.update(top, bottom)
    return grad


def superquadric_objective_seed(F, n_surface):
    """dJ/dF for J = sum_{surface} F^2 + 0.1 mean_{exterior} exp(-F) (non-finite F masked)."""
    g = torch.empty_like(F)
    fs = F[:n_surface]
    fe = F[n_surface:]
    g[:n_surface] = 2.0 * fs
    g[n_surface:] = -0.1 / max(1, fe.numel()) * torch.exp(-fe)
    return torch.where(torch.isfinite(g), g, torch.zeros_like(g))


def superquadric_objective(F, n_surface):
    fs = F[:n_surface]
    fe = F[n_surface:]
    ms = torch.isfinite(fs)
    me = torch.isfinite(fe)
    return float((fs[ms] ** 2).sum() + 0.1 * torch.exp(-fe[me]).sum() / max(1, fe.numel()))
