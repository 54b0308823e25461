"""torch.autograd binding of the static solve and its adjoint.

The element modulus scales s_e (E_e = E * s_e) are the differentiable
parameters.  K(s) = sum_e s_e K~_e^0, so for u = K(s)^-1 f and a loss L(u)
with upstream gradient g = dL/du:

    w        = K(s)^-1 g              (adjoint solve: the same PCG, K symmetric)
    dL/ds_e  = -w_e' K~_e^0 u_e       (femgpu_material_grad, K~_e^0 recomputed)
    dL/df    = w

`compliance(...)` is the special case L = f'u with the problem's own loads:
its adjoint solution is u itself, so forward and gradient cost one solve
(femgpu_compliance_grad).

The reference (femfunc.cpp) has no differentiation; this is the differentiable
face of the same EbE solve.  Tensors may live on any device: the C-ABI takes
host arrays, so values pass through host memory as float64.
"""
from __future__ import annotations

import numpy as np
import torch

from ._lib import FP64, TOL_REL_RES, Context


def _np(t: torch.Tensor) -> np.ndarray:
    return np.ascontiguousarray(t.detach().to("cpu", torch.float64).numpy())


class _Material:
    """Remembers which scales are loaded so backward can restore them."""

    def __init__(self, context: Context, E: float, nu: float):
        self.context, self.E, self.nu = context, E, nu
        self.loaded: np.ndarray | None = None

    def load(self, s: np.ndarray):
        if self.loaded is None or self.loaded.shape != s.shape or not np.array_equal(self.loaded, s):
            self.context.set_material(self.E, self.nu, s)
            self.loaded = s.copy()


class ElasticSolve(torch.autograd.Function):
    """u = K(scale)^-1 f; differentiable in scale (n_elems) and f (n_nodes x dim)."""

    @staticmethod
    def forward(ctx, scale: torch.Tensor, f: torch.Tensor, material: _Material, opts: dict):
        s = _np(scale)
        material.load(s)
        u, stats = material.context.solve_rhs(_np(f), **opts)
        ctx.material, ctx.opts, ctx.s = material, opts, s
        ctx.stats = stats
        out = torch.from_numpy(u).to(device=f.device, dtype=f.dtype)
        ctx.save_for_backward(out)
        return out

    @staticmethod
    def backward(ctx, grad_u: torch.Tensor):
        (u,) = ctx.saved_tensors
        ctx.material.load(ctx.s)
        w, _ = ctx.material.context.solve_rhs(_np(grad_u), **ctx.opts)
        g = ctx.material.context.material_grad(_np(u), w)
        grad_scale = torch.from_numpy(g).to(device=grad_u.device, dtype=grad_u.dtype)
        grad_f = torch.from_numpy(w).to(device=grad_u.device, dtype=grad_u.dtype)
        return grad_scale, grad_f, None, None


class Compliance(torch.autograd.Function):
    """C(scale) = f'u with K(scale) u = f for the problem's own loads."""

    @staticmethod
    def forward(ctx, scale: torch.Tensor, material: _Material, opts: dict):
        s = _np(scale)
        material.load(s)
        c, g, _, stats = material.context.compliance_grad(**opts)
        ctx.stats = stats
        ctx.save_for_backward(torch.from_numpy(g).to(device=scale.device, dtype=scale.dtype))
        return torch.tensor(c, device=scale.device, dtype=scale.dtype)

    @staticmethod
    def backward(ctx, grad_c: torch.Tensor):
        (g,) = ctx.saved_tensors
        return grad_c * g, None, None


class ElasticProblem:
    """A loaded problem (mesh, constraints, loads) whose element modulus
    scales are the differentiable input.

    >>> ep = ElasticProblem(context, E=2e7, nu=0.3)
    >>> rho = torch.rand(n_elems, dtype=torch.float64, requires_grad=True)
    >>> C = ep.compliance(rho ** 3); C.backward()        # rho.grad = dC/drho
    >>> u = ep.solve(rho ** 3, f); (u ** 2).sum().backward()
    """

    def __init__(self, context: Context, E: float, nu: float, precision: int = FP64,
                 tol: float = 1e-10, tol_kind: int = TOL_REL_RES, max_iter: int = 0):
        self.material = _Material(context, E, nu)
        self.opts = dict(precision=precision, tol=tol, tol_kind=tol_kind, max_iter=max_iter)

    def solve(self, scale: torch.Tensor, f: torch.Tensor) -> torch.Tensor:
        return ElasticSolve.apply(scale, f, self.material, self.opts)

    def compliance(self, scale: torch.Tensor) -> torch.Tensor:
        return Compliance.apply(scale, self.material, self.opts)
