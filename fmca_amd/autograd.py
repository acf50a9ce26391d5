"""Differentiable H2 kernel matvec: y = K(targets, points; l) alpha.

The backward pass returns
  d alpha   = K^T G                                (transposed H2 action)
  d l       = <G, (dK/dl) alpha>                    (H2 action with dk/dl)
  d points  = sum_r alpha_r (s (Phi^T g_r) - Phi^T (g_r * t))   per coordinate
  d targets = sum_r g_r   (t (Phi a_r)   - Phi (a_r * s))
with Phi the H2 action of phi(r) = k'(r)/r on the same tree, bases and blocks
(for targets = points both terms add up).  The H2 operator is linear in the
kernel function, so each of these is one action with a derivative kernel; no
moments are stored between forward and backward -- every action recomputes
them from the points and the rhs.

FMCA has no differentiation code; these forms are checked against centred
finite differences of the H2 forward and against dense FP64 autograd.
"""
from __future__ import annotations

import numpy as np
import torch

from . import FP64, CovarianceKernel, H2kernelActor
from . import _C

KERNEL_VALUE = _C.KERNEL_VALUE
KERNEL_DL = _C.KERNEL_DL
KERNEL_PHI = _C.KERNEL_PHI


class H2Spec:
    """Kernel type, interpolation degree, admissibility and precision of an H2
    matvec, plus a cache of the plan for the current point set(s)."""

    def __init__(self, ktype: str = "MATERN32", deg: int = 4, eta: float = 0.8,
                 precision: int = FP64, c: float = 1.0, nu: float = 1.0):
        self.ktype = ktype.upper()
        self.deg = int(deg)
        self.eta = float(eta)
        self.precision = int(precision)
        self.c = float(c)
        self.nu = float(nu)
        self._key = None
        self._actor = None

    def actor(self, points: torch.Tensor, targets, l: float) -> H2kernelActor:
        key = (points.data_ptr(), points._version, tuple(points.shape),
               None if targets is None else (targets.data_ptr(), targets._version,
                                             tuple(targets.shape)))
        if self._actor is None or key != self._key:
            P = np.asfortranarray(points.detach().cpu().numpy().astype(np.float64))
            T = P if targets is None else np.asfortranarray(
                targets.detach().cpu().numpy().astype(np.float64))
            self._actor = H2kernelActor(CovarianceKernel(self.ktype, l, self.c, self.nu), T, P,
                                        self.deg, self.eta)
            self._key = key
        self._actor.set_kernel(CovarianceKernel(self.ktype, l, self.c, self.nu))
        return self._actor


def _variant(actor: H2kernelActor, variant: int, x: torch.Tensor, transpose: bool,
             precision: int) -> torch.Tensor:
    x = x if x.dim() == 2 else x.reshape(-1, 1)
    if x.shape[1] == 1:
        x = x.contiguous()
    elif x.stride(0) != 1 or x.stride(1) < x.shape[0]:
        x = x.t().contiguous().t()  # column-major copy
    rows_out = actor.n if transpose else actor.m
    rows_in = actor.m if transpose else actor.n
    nrhs = x.shape[1]
    out = torch.empty((nrhs, rows_out), dtype=torch.float64, device=x.device).t()
    actor._a.action_variant_device(variant, transpose, x.data_ptr(),
                                   x.stride(1) if nrhs > 1 else rows_in, nrhs, out.data_ptr(),
                                   out.stride(1) if nrhs > 1 else rows_out, precision,
                                   torch.cuda.current_stream(x.device).cuda_stream)
    return out


class _H2Matvec(torch.autograd.Function):
    @staticmethod
    def forward(ctx, alpha, lengthscale, points, targets, spec):
        l = float(lengthscale.detach().item())
        actor = spec.actor(points, targets, l)
        a2 = alpha if alpha.dim() == 2 else alpha.reshape(-1, 1)
        y = _variant(actor, KERNEL_VALUE, a2, False, spec.precision)
        ctx.actor = actor
        ctx.spec = spec
        ctx.vector = alpha.dim() == 1
        ctx.symmetric = targets is None
        ctx.save_for_backward(alpha, lengthscale, points,
                              targets if targets is not None else points)
        return y[:, 0] if alpha.dim() == 1 else y

    @staticmethod
    def backward(ctx, grad):
        alpha, lengthscale, points, targets = ctx.saved_tensors
        actor, prec = ctx.actor, ctx.spec.precision
        G = grad.reshape(-1, 1) if ctx.vector else grad
        G = G.to(torch.float64)
        A = alpha.reshape(-1, 1) if ctx.vector else alpha
        d_alpha = d_l = d_points = d_targets = None
        if ctx.needs_input_grad[0]:
            d_alpha = _variant(actor, KERNEL_VALUE, G, True, prec)
            d_alpha = d_alpha[:, 0] if ctx.vector else d_alpha
        if ctx.needs_input_grad[1]:
            Yl = _variant(actor, KERNEL_DL, A, False, prec)
            d_l = (G * Yl).sum().reshape(lengthscale.shape).to(lengthscale.dtype)
        need_s = ctx.needs_input_grad[2]
        need_t = ctx.needs_input_grad[3] or (ctx.symmetric and need_s)
        if need_s or need_t:
            dev = A.device
            S = points.detach().to(dev, torch.float64)     # (d, n) sources
            T = targets.detach().to(dev, torch.float64)    # (d, m) targets
            d, r = S.shape[0], A.shape[1]
            if need_t:
                # targets: sum_r g_r (t (Phi a_r) - Phi (a_r * s))
                U = torch.cat([A] + [A * S[k][:, None] for k in range(d)], dim=1)
                V = _variant(actor, KERNEL_PHI, U, False, prec)
                Va = V[:, :r]
                gt = torch.stack([(G * (T[k][:, None] * Va - V[:, (k + 1) * r:(k + 2) * r])).sum(1)
                                  for k in range(d)])
            if need_s:
                # sources: sum_r a_r (s (Phi^T g_r) - Phi^T (g_r * t))
                U = torch.cat([G] + [G * T[k][:, None] for k in range(d)], dim=1)
                W = _variant(actor, KERNEL_PHI, U, True, prec)
                Wg = W[:, :r]
                gs = torch.stack([(A * (S[k][:, None] * Wg - W[:, (k + 1) * r:(k + 2) * r])).sum(1)
                                  for k in range(d)])
            if ctx.symmetric:
                d_points = (gs + gt).to(points.device, points.dtype)
            else:
                d_points = gs.to(points.device, points.dtype) if need_s else None
                d_targets = gt.to(targets.device, targets.dtype) if ctx.needs_input_grad[3] else None
        return d_alpha, d_l, d_points, d_targets, None


def h2_matvec(alpha: torch.Tensor, lengthscale: torch.Tensor, points: torch.Tensor,
              spec: H2Spec, targets: torch.Tensor | None = None) -> torch.Tensor:
    """y = K_H2(targets, points; l) alpha, differentiable in alpha, l, points, targets.

    alpha: (n,) or (n, nrhs) float64 CUDA tensor; lengthscale: 0-dim tensor;
    points / targets: (dim, n) / (dim, m) float64 tensors (targets None: the
    points themselves, one shared tree)."""
    if not torch.is_tensor(lengthscale):
        lengthscale = torch.tensor(float(lengthscale), dtype=torch.float64)
    return _H2Matvec.apply(alpha, lengthscale, points, targets, spec)
