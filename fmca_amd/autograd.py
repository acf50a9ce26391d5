"""Differentiable H2 kernel matvec: y = K(targets, points; l) alpha.

One point set (targets = points, one tree, tensor-product interpolation, at
most 16 columns) with point gradients requested: the backward is
fmca_h2_gradient_device -- one action on
[alpha | G | alpha] (the last group with dk/dl) that stops after the downward
pass, then gradient kernels that give
  d alpha   = K~ G
  d l       = <G, dK~/dl alpha>
  d points  = the derivative of the H2 forward itself: near blocks through
              phi(r) = k'(r)/r and the pair weights G_p.alpha_j + alpha_p.G_j,
              far blocks through the gradient of the interpolation basis at
              the point dotted with the far-field sums of alpha and G
with the tree and boxes held fixed.

Otherwise (no point gradients, two point sets, total degree, more columns)
the backward applies
H2 actions with derivative kernels, which the H2 operator's linearity in the
kernel function allows:
  d alpha   = K^T G                                (transposed H2 action)
  d l       = <G, (dK/dl) alpha>                    (H2 action with dk/dl)
  d points  = sum_r alpha_r (s (Phi^T g_r) - Phi^T (g_r * t))   per coordinate
  d targets = sum_r g_r   (t (Phi a_r)   - Phi (a_r * s))
with Phi the H2 action of phi on the same tree, bases and blocks, fused into
one action whose 16-column groups carry different kernel variants
(fmca_h2_action_variants_device).  In both forms no moments are stored
between forward and backward: the backward recomputes them from the points
and its right-hand sides.

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
    """Kernel type, interpolation degree, admissibility, precision,
    interpolator ("tensor_product" or "total_degree") and planner ("auto",
    "host" or "device") of an H2 matvec, plus a cache of the plan for the
    current point set(s)."""

    def __init__(self, ktype: str = "MATERN32", deg: int = 4, eta: float = 0.8,
                 precision: int = FP64, c: float = 1.0, nu: float = 1.0,
                 interpolator: str = "tensor_product", planner: str = "auto"):
        self.ktype = ktype.upper()
        self.interpolator = interpolator
        self.planner = planner
        self.deg = int(deg)
        self.eta = float(eta)
        self.precision = int(precision)
        self.c = float(c)
        self.nu = float(nu)
        self._key = None
        self._actor = None
        self._held = None  # (points, targets) the cached plan was built from

    def actor(self, points: torch.Tensor, targets, l: float) -> H2kernelActor:
        """The actor planned on `points` / `targets`, re-planned when they
        change.  The spec holds the tensors the plan was built from and a copy
        of their values: a fresh tensor that reuses the memory of a freed one
        (same address, version 0) is a different object, and values changed
        behind the version counter (through .data) differ from the copy."""
        key = (id(points), points._version, tuple(points.shape),
               None if targets is None else (id(targets), targets._version,
                                             tuple(targets.shape)))
        same = (self._actor is not None and key == self._key
                and self._held[0] is points and self._held[1] is targets
                and torch.equal(self._copy[0], points.detach())
                and (targets is None or torch.equal(self._copy[1], targets.detach())))
        if not same:
            P = np.asfortranarray(points.detach().cpu().numpy().astype(np.float64))
            T = P if targets is None else np.asfortranarray(
                targets.detach().cpu().numpy().astype(np.float64))
            self._actor = H2kernelActor(CovarianceKernel(self.ktype, l, self.c, self.nu), T, P,
                                        self.deg, self.eta, self.interpolator,
                                        self.planner)
            self._key = key
            self._held = (points, targets)
            self._copy = (points.detach().clone(),
                          None if targets is None else targets.detach().clone())
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


def _variants(actor: H2kernelActor, blocks, transpose: bool, precision: int):
    """Applies the H2 matrix of kernel variant v to every (v, X) of `blocks`
    (X: rows_in x c) in fused actions: each block fills ceil(c / 16) column
    groups of one action of up to 8 groups.  Returns the products in order,
    as column-major views."""
    if not blocks:
        return []
    dev = blocks[0][1].device
    rows_in = actor.m if transpose else actor.n
    rows_out = actor.n if transpose else actor.m
    stream = torch.cuda.current_stream(dev).cuda_stream
    pieces = []  # (variant, block, first column, end column), 16 columns at most
    for b, (v, X) in enumerate(blocks):
        for c0 in range(0, X.shape[1], 16):
            pieces.append((v, b, c0, min(c0 + 16, X.shape[1])))
    parts = [[] for _ in blocks]
    for p0 in range(0, len(pieces), 8):
        call = pieces[p0:p0 + 8]
        # the rhs of the action as (ncols, rows_in) row-major = column-major (rows_in, ncols)
        tens = []
        for g, (v, b, c0, c1) in enumerate(call):
            tens.append(blocks[b][1][:, c0:c1].t())
            if c1 - c0 < 16 and g < len(call) - 1:
                tens.append(torch.zeros((16 - (c1 - c0), rows_in), dtype=torch.float64,
                                        device=dev))
        Ut = torch.cat(tens, dim=0).to(torch.float64)
        ncols = Ut.shape[0]
        Yt = torch.empty((ncols, rows_out), dtype=torch.float64, device=dev)
        actor._a.action_variants_device([v for v, _, _, _ in call], transpose, Ut.data_ptr(),
                                        rows_in, ncols, Yt.data_ptr(), rows_out, precision,
                                        stream)
        # the groups of one block are adjacent (padding only ends a block)
        first = {}
        for g, (v, b, c0, c1) in enumerate(call):
            g0, w = first.get(b, (g, 0))
            first[b] = (g0, 16 * (g - g0) + (c1 - c0))
        for b, (g0, w) in first.items():
            parts[b].append(Yt[16 * g0:16 * g0 + w])
    return [(p[0] if len(p) == 1 else torch.cat(p, dim=0)).t() for p in parts]


def _colmajor(X: torch.Tensor) -> torch.Tensor:
    X = X.to(torch.float64)
    if X.shape[1] == 1:
        return X.contiguous()
    return X if (X.stride(0) == 1 and X.stride(1) == X.shape[0]) else X.t().contiguous().t()


def _gradient(actor: H2kernelActor, A: torch.Tensor, G: torch.Tensor, precision: int):
    """(K~ G, dL/dP, dL/dl) of L = <G, K~ A> for one shared point set, from the
    gradient kernels (fmca_h2_gradient_device); None where that form does not
    apply (two point sets, total degree, > 16 columns, no matrix-core fit)."""
    n, r = A.shape
    A, G = _colmajor(A), _colmajor(G)
    dev = A.device
    dalpha = torch.empty((r, n), dtype=torch.float64, device=dev).t()
    dpoints = torch.empty((actor.dim, n), dtype=torch.float64, device=dev)
    dl = torch.empty(1, dtype=torch.float64, device=dev)
    try:
        actor._a.gradient_device(A.data_ptr(), n, G.data_ptr(), n, r, precision,
                                 dalpha.data_ptr(), n, dpoints.data_ptr(), n, dl.data_ptr(),
                                 torch.cuda.current_stream(dev).cuda_stream)
    except ValueError as e:
        if str(e).startswith("not supported"):
            return None
        raise
    return dalpha, dpoints, dl


class _H2Matvec(torch.autograd.Function):
    @staticmethod
    def forward(ctx, alpha, lengthscale, points, targets, spec):
        l = float(lengthscale.detach().item())
        actor = spec.actor(points, targets, l)
        a2 = alpha if alpha.dim() == 2 else alpha.reshape(-1, 1)
        y = _variant(actor, KERNEL_VALUE, a2, False, spec.precision)
        ctx.actor = actor
        ctx.l = l
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
        # the actor is shared with later forwards of the same spec, each of
        # which sets its own lengthscale: restore this forward's
        spec = ctx.spec
        actor.set_kernel(CovarianceKernel(spec.ktype, ctx.l, spec.c, spec.nu))
        G = (grad.reshape(-1, 1) if ctx.vector else grad).to(torch.float64)
        A = (alpha.reshape(-1, 1) if ctx.vector else alpha).to(torch.float64)
        need_a, need_l = ctx.needs_input_grad[0], ctx.needs_input_grad[1]
        need_s = ctx.needs_input_grad[2]
        need_t = ctx.needs_input_grad[3] or (ctx.symmetric and need_s)
        dev = A.device
        r = A.shape[1]
        if ctx.symmetric and need_s and r <= 16:
            # one point set with point gradients: the gradient kernels give all
            # three at once (without them the fused derivative actions below
            # need fewer columns)
            out = _gradient(actor, A, G, prec)
            if out is not None:
                dal, dpo, dl = out
                d_alpha = (dal[:, 0] if ctx.vector else dal) if need_a else None
                d_l = dl.reshape(lengthscale.shape).to(lengthscale.device,
                                                       lengthscale.dtype) if need_l else None
                d_points = dpo.to(points.device, points.dtype) if need_s else None
                return d_alpha, d_l, d_points, None, None
        S = points.detach().to(dev, torch.float64) if (need_s or need_t) else None  # (d, n)
        T = targets.detach().to(dev, torch.float64) if (need_s or need_t) else None  # (d, m)
        d = S.shape[0] if S is not None else 0
        # blocks of the action on the forward pattern and of the transposed one
        fwd, trn = [], []
        if need_l:
            fwd.append(("dl", KERNEL_DL, A))
        if ctx.symmetric:
            if need_a:
                fwd.append(("alpha", KERNEL_VALUE, G))
            if need_s:
                fwd.append(("phi", KERNEL_PHI, torch.cat(
                    [A, G] + [A * S[k][:, None] for k in range(d)]
                    + [G * S[k][:, None] for k in range(d)], dim=1)))
        else:
            if need_a:
                trn.append(("alpha", KERNEL_VALUE, G))
            if need_t:
                fwd.append(("phi_t", KERNEL_PHI, torch.cat(
                    [A] + [A * S[k][:, None] for k in range(d)], dim=1)))
            if need_s:
                trn.append(("phi_s", KERNEL_PHI, torch.cat(
                    [G] + [G * T[k][:, None] for k in range(d)], dim=1)))
        res = {}
        for blocks, transpose in ((fwd, False), (trn, True)):
            outs = _variants(actor, [(v, X) for _, v, X in blocks], transpose, prec)
            res.update({name: Y for (name, _, _), Y in zip(blocks, outs)})
        d_alpha = d_l = d_points = d_targets = None
        if need_a:
            d_alpha = res["alpha"][:, 0] if ctx.vector else res["alpha"]
        if need_l:
            d_l = (G * res["dl"]).sum().reshape(lengthscale.shape).to(lengthscale.dtype)
        if ctx.symmetric and need_s:
            V = res["phi"]
            PA, PG = V[:, :r], V[:, r:2 * r]
            g = torch.stack([
                (G * (S[k][:, None] * PA - V[:, (2 + k) * r:(3 + k) * r])).sum(1)
                + (A * (S[k][:, None] * PG - V[:, (2 + d + k) * r:(3 + d + k) * r])).sum(1)
                for k in range(d)])
            d_points = g.to(points.device, points.dtype)
        elif not ctx.symmetric:
            if need_t and ctx.needs_input_grad[3]:
                V = res["phi_t"]
                gt = torch.stack([(G * (T[k][:, None] * V[:, :r]
                                        - V[:, (1 + k) * r:(2 + k) * r])).sum(1)
                                  for k in range(d)])
                d_targets = gt.to(targets.device, targets.dtype)
            if need_s:
                W = res["phi_s"]
                gs = torch.stack([(A * (S[k][:, None] * W[:, :r]
                                        - W[:, (1 + k) * r:(2 + k) * r])).sum(1)
                                  for k in range(d)])
                d_points = gs.to(points.device, points.dtype)
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
