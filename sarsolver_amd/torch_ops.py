"""Device-resident SAR operators for torch: autograd functions and a batched operator.

`SarForwardFn` computes y = A x with the gfx950 forward kernels; its backward
is the hand-written backprojection kernel, grad_x = A^H grad_y, which is the
vector-Jacobian product of a C-linear map under PyTorch's complex convention
(for a real loss L, x.grad = dL/dRe(x) + i dL/dIm(x)).  `SarAdjointFn` is the
mirror image.  The gradient is exact for the adjoint pairs `nearest` and
`linear`; with `nearest_ref` the backward uses the reference's backprojection,
which is not the exact adjoint of its forward (sarsolver_cxx.cpp:289 vs :347).

Shapes: images are [M] (or [B, M] for a batch), phase histories [P, N]
(or [B, P, N]); all complex64 on the plan's device.
"""
from __future__ import annotations

from typing import Sequence

import torch

from ._native import SarPlan


class SarForwardFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, plan: SarPlan, order: str = "fft"):
        ctx.plan, ctx.order = plan, order
        return plan.forward(x.contiguous(), order=order)

    @staticmethod
    def backward(ctx, gy):
        return ctx.plan.adjoint(gy.contiguous(), order=ctx.order), None, None


class SarAdjointFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, y, plan: SarPlan, order: str = "fft"):
        ctx.plan, ctx.order = plan, order
        return plan.adjoint(y.contiguous(), order=order)

    @staticmethod
    def backward(ctx, gx):
        return ctx.plan.forward(gx.contiguous(), order=ctx.order), None, None


def sar_forward(x: torch.Tensor, plan: SarPlan, order: str = "fft") -> torch.Tensor:
    """Differentiable A x."""
    return SarForwardFn.apply(x, plan, order)


def sar_adjoint(y: torch.Tensor, plan: SarPlan, order: str = "fft") -> torch.Tensor:
    """Differentiable A^H y."""
    return SarAdjointFn.apply(y, plan, order)


class SarOperator:
    """A (and A^H, A^H A) on device vectors for one plan, with an optional
    Tikhonov shift: normal(x) = A^H A x + lam x."""

    def __init__(self, plan: SarPlan, lam: float = 0.0, order: str = "fft"):
        self.plan, self.lam, self.order = plan, float(lam), order
        self.shape_in = (plan.num_scatterers,)
        self.shape_out = (plan.num_slow_times, plan.num_fast_times)
        self.device = plan.device

    def forward(self, x):
        return sar_forward(x, self.plan, self.order)

    def adjoint(self, y):
        return sar_adjoint(y, self.plan, self.order)

    def normal(self, x):
        out = self.adjoint(self.forward(x))
        if self.lam:
            out = out + self.lam * x
        return out

    __call__ = forward


class BatchedSarOperator:
    """A batch of scenes, each with its own aperture (one plan per scene):
    forward maps [B, M] -> [B, P, N], differentiably."""

    def __init__(self, plans: Sequence[SarPlan], order: str = "fft"):
        if not plans:
            raise ValueError("at least one plan is needed")
        self.plans, self.order = list(plans), order

    def forward(self, x):
        if x.shape[0] != len(self.plans):
            raise ValueError(f"batch {x.shape[0]} != {len(self.plans)} plans")
        return torch.stack([sar_forward(x[b], p, self.order) for b, p in enumerate(self.plans)])

    def adjoint(self, y):
        if y.shape[0] != len(self.plans):
            raise ValueError(f"batch {y.shape[0]} != {len(self.plans)} plans")
        return torch.stack([sar_adjoint(y[b], p, self.order) for b, p in enumerate(self.plans)])

    __call__ = forward


def data_fidelity(A, x, y) -> torch.Tensor:
    """0.5 ||A x - y||^2 (real, summed over the batch); its gradient in x is
    A^H (A x - y)."""
    r = A(x) - y
    return 0.5 * torch.sum(r.real.double() ** 2 + r.imag.double() ** 2)
