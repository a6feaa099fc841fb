"""Device solvers for (A^H A + lam I) x = A^H b: CGLS and flexible preconditioned CG.

Operators are any objects with `forward(x)` and `adjoint(y)` on torch tensors
(for instance `torch_ops.SarOperator`).  Vectors stay on the operator's device;
each iteration synchronises with the host once, to read the scalars that
decide whether to stop (both solvers are written the same way, so their
wall times compare the methods and not the bookkeeping).

* `cgls`: conjugate gradients on the normal equations written with A and A^H
  (one A and one A^H per iteration).  It is the fallback and the yardstick:
  every solve is scored by `true_residual`, recomputed from scratch.
* `fpcg`: flexible PCG with a preconditioner callable z = M(r) that may be
  nonlinear (a trained network): Polak-Ribiere beta.  Non-finite values,
  non-positive curvature p^H K p <= 0, or a residual grown to 10x the
  smallest seen so far, switch the solve to CGLS from the best iterate
  (reported in the result).
Stopping test: ||A^H b - (A^H A + lam I) x|| / ||A^H b|| <= tol.
Inner products accumulate in the vectors' precision (complex64 on the GPU);
`true_residual` recomputes the residual of a result in FP64.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field
from typing import Callable, List, Optional

import torch


def _dot(a: torch.Tensor, b: torch.Tensor) -> torch.Tensor:
    """a^H b as a 0-d tensor (no host synchronisation)."""
    if a.is_complex():
        return torch.vdot(a.reshape(-1), b.reshape(-1))
    return torch.dot(a.reshape(-1), b.reshape(-1))


def _norm(a: torch.Tensor) -> float:
    """||a|| accumulated in FP64 (synchronises)."""
    return float(torch.linalg.vector_norm(a.reshape(-1).to(
        torch.complex128 if a.is_complex() else torch.float64)))


def _host(*ts: torch.Tensor) -> List[float]:
    """Real parts of several 0-d tensors in one device-to-host transfer."""
    return torch.stack([t.real if t.is_complex() else t for t in ts]).double().tolist()


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
        s = c
    else:
        x = x0.clone()
        r = b - op.forward(x)
        s = op.adjoint(r) - lam * x
    p = s
    gamma = _dot(s, s).real
    res = [math.sqrt(max(_host(gamma)[0], 0.0)) / c_norm]
    it = 0
    converged = res[-1] <= tol
    while not converged and it < max_iter:
        q = op.forward(p)
        delta = _dot(q, q).real + lam * _dot(p, p).real
        alpha = (gamma / delta).to(x.dtype)
        x_prev = x
        x = x + alpha * p
        r = r - alpha * q
        s = op.adjoint(r) - lam * x
        gamma_new = _dot(s, s).real
        d, g = _host(delta, gamma_new)
        if not (d > 0.0) or not math.isfinite(d) or not math.isfinite(g):
            x = x_prev  # the step was not defined: keep the last good iterate
            break
        it += 1
        res.append(math.sqrt(max(g, 0.0)) / c_norm)
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
    r = c
    z = precond(r)
    p = z
    rz = _dot(r, z)
    res = [1.0]
    it = 0
    fallback = None
    converged = False
    best_x, best_res = x, 1.0
    while it < max_iter:
        Kp = normal_apply(op, p, lam)
        pKp = _dot(p, Kp)
        alpha = (rz / pKp).to(x.dtype)
        x_new = x + alpha * p
        r_new = r - alpha * Kp
        rr = _dot(r_new, r_new).real
        curv, rr_h = _host(pKp, rr)
        # a non-finite z (preconditioner) or step shows up here as well
        if not (curv > 0.0) or not math.isfinite(curv) or not math.isfinite(rr_h):
            fallback = "non-positive or non-finite curvature p^H K p"
            x = best_x
            break
        x, it = x_new, it + 1
        res.append(math.sqrt(max(rr_h, 0.0)) / c_norm)
        if res[-1] <= tol:
            converged = True
            break
        if res[-1] < best_res:
            best_x, best_res = x, res[-1]
        elif res[-1] > 10.0 * best_res:
            fallback = "residual grew to 10x its minimum"
            x = best_x
            break
        z_new = precond(r_new)
        rz_new = _dot(r_new, z_new)
        beta = ((rz_new - _dot(r, z_new)) / rz).to(x.dtype)  # Polak-Ribiere
        p = z_new + beta * p
        r, rz = r_new, rz_new
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
