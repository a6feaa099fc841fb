"""Device solvers for (A^H A + lam I) x = A^H b: CGLS and flexible preconditioned CG.

Operators are any objects with `forward(x)` and `adjoint(y)` on torch tensors
(for instance `torch_ops.SarOperator`).  Vectors stay on the operator's device;
each iteration synchronises once to read the residual norm.

* `cgls`: conjugate gradients on the normal equations written with A and A^H
  (one A and one A^H per iteration).  It is the fallback and the judge: every
  solve is scored by `true_residual`, recomputed from scratch.
* `fpcg`: flexible PCG with a preconditioner callable z = M(r) that may be
  nonlinear (a trained network): Polak-Ribiere beta.  Non-finite values or
  non-positive curvature p^H K p <= 0, or a residual grown to 10x the
  smallest seen so far, switch the solve to CGLS from the best iterate
  (reported in the result).
Stopping test: ||A^H b - (A^H A + lam I) x|| / ||A^H b|| <= tol.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field
from typing import Callable, List, Optional

import torch


def _dot(a: torch.Tensor, b: torch.Tensor) -> torch.Tensor:
    """a^H b accumulated in float64 / complex128."""
    if a.is_complex():
        return torch.vdot(a.reshape(-1).to(torch.complex128), b.reshape(-1).to(torch.complex128))
    return torch.dot(a.reshape(-1).double(), b.reshape(-1).double())


def _norm(a: torch.Tensor) -> float:
    return float(torch.linalg.vector_norm(a.reshape(-1).to(
        torch.complex128 if a.is_complex() else torch.float64)))


@dataclass
class SolveResult:
    x: torch.Tensor
    iterations: int
    residuals: List[float] = field(default_factory=list)  # relative, per iteration
    converged: bool = False
    seconds: float = 0.0
    fallback: Optional[str] = None  # why fpcg handed over to CGLS, if it did
    true_residual: Optional[float] = None


def normal_apply(op, x: torch.Tensor, lam: float) -> torch.Tensor:
    out = op.adjoint(op.forward(x))
    return out + lam * x if lam else out


def true_residual(op, x: torch.Tensor, b: torch.Tensor, lam: float = 0.0,
                  rhs: Optional[torch.Tensor] = None) -> float:
    """||A^H b - (A^H A + lam I) x|| / ||A^H b||, recomputed with fresh operator calls."""
    c = op.adjoint(b) if rhs is None else rhs
    r = c - normal_apply(op, x, lam)
    return _norm(r) / max(_norm(c), 1e-300)


def cgls(op, b: torch.Tensor, lam: float = 0.0, tol: float = 1e-3, max_iter: int = 200,
         x0: Optional[torch.Tensor] = None, sync_time: bool = True) -> SolveResult:
    """CGLS for min ||A x - b||^2 + lam ||x||^2."""
    t0 = time.perf_counter()
    c = op.adjoint(b)
    c_norm = max(_norm(c), 1e-300)
    if x0 is None:
        x = torch.zeros_like(c)
        r = b.clone()
        s = c.clone()
    else:
        x = x0.clone()
        r = b - op.forward(x)
        s = op.adjoint(r) - lam * x
    p = s.clone()
    gamma = _dot(s, s).real
    res = [math.sqrt(float(gamma)) / c_norm]
    it = 0
    converged = res[-1] <= tol
    while not converged and it < max_iter:
        q = op.forward(p)
        delta = _dot(q, q).real + lam * _dot(p, p).real
        if not (float(delta) > 0.0):
            break
        alpha = (gamma / delta).to(x.dtype)
        x = x + alpha * p
        r = r - alpha * q
        s = op.adjoint(r) - lam * x
        gamma_new = _dot(s, s).real
        it += 1
        res.append(math.sqrt(float(gamma_new)) / c_norm)
        if not math.isfinite(res[-1]):
            break
        converged = res[-1] <= tol
        p = s + (gamma_new / gamma).to(x.dtype) * p
        gamma = gamma_new
    if sync_time and x.is_cuda:
        torch.cuda.synchronize(x.device)
    return SolveResult(x=x, iterations=it, residuals=res, converged=converged,
                       seconds=time.perf_counter() - t0)


def fpcg(op, b: torch.Tensor, precond: Callable[[torch.Tensor], torch.Tensor], lam: float = 0.0,
         tol: float = 1e-3, max_iter: int = 200, sync_time: bool = True) -> SolveResult:
    """Flexible preconditioned CG on (A^H A + lam I) x = A^H b."""
    t0 = time.perf_counter()
    c = op.adjoint(b)
    c_norm = max(_norm(c), 1e-300)
    x = torch.zeros_like(c)
    r = c.clone()
    z = precond(r)
    p = z.clone()
    rz = _dot(r, z)
    res = [1.0]
    it = 0
    fallback = None
    converged = False
    best_x, best_res = x, 1.0
    while it < max_iter:
        if not bool(torch.isfinite(torch.view_as_real(z) if z.is_complex() else z).all()):
            fallback = "preconditioner returned non-finite values"
            x = best_x
            break
        Kp = normal_apply(op, p, lam)
        pKp = _dot(p, Kp)
        if not (float(pKp.real) > 0.0) or not math.isfinite(float(pKp.real)):
            fallback = "non-positive curvature p^H K p <= 0"
            x = best_x
            break
        alpha = (rz / pKp).to(x.dtype)
        x = x + alpha * p
        r_new = r - alpha * Kp
        it += 1
        res.append(_norm(r_new) / c_norm)
        if res[-1] <= tol:
            converged = True
            r = r_new
            break
        if res[-1] < best_res:
            best_x, best_res = x, res[-1]
        elif not math.isfinite(res[-1]) or res[-1] > 10.0 * best_res:
            fallback = "residual grew to 10x its minimum"
            x = best_x
            break
        z_new = precond(r_new)
        rz_new = _dot(r_new, z_new)
        beta = (_dot(r_new - r, z_new) / rz).to(x.dtype)  # Polak-Ribiere
        p = z_new + beta * p
        r, z, rz = r_new, z_new, rz_new
    if fallback is not None:
        rest = cgls(op, b, lam=lam, tol=tol, max_iter=max(1, max_iter - it), x0=x,
                    sync_time=False)
        x = rest.x
        it += rest.iterations
        res.extend(rest.residuals[1:])
        converged = rest.converged
    if sync_time and x.is_cuda:
        torch.cuda.synchronize(x.device)
    return SolveResult(x=x, iterations=it, residuals=res, converged=converged,
                       seconds=time.perf_counter() - t0, fallback=fallback)


class DenseOperator:
    """A dense matrix as an operator (tests and small problems)."""

    def __init__(self, A: torch.Tensor):
        self.A = A

    def forward(self, x):
        return self.A @ x

    def adjoint(self, y):
        return self.A.conj().T @ y
