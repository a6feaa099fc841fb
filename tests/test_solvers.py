"""CGLS / flexible PCG on dense complex systems (CPU).

CGLS has no reference counterpart with assertions (the reference's test_cgls
asserts nothing), so it is pinned by closed forms: after n iterations on an
n-column system it equals the least-squares solution (numpy.linalg.lstsq /
the regularised normal equations), and it agrees with scipy's LSQR.
"""
import numpy as np
import pytest
import torch

from sarsolver_amd.solvers import DenseOperator, cgls, fpcg, true_residual


def system(m=60, n=12, seed=0, cond=30.0):
    rng = np.random.default_rng(seed)
    U, _ = np.linalg.qr(rng.standard_normal((m, n)) + 1j * rng.standard_normal((m, n)))
    V, _ = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    s = np.geomspace(1.0, 1.0 / cond, n)
    A = (U * s) @ V.conj().T
    b = rng.standard_normal(m) + 1j * rng.standard_normal(m)
    return torch.tensor(A, dtype=torch.complex128), torch.tensor(b, dtype=torch.complex128)


def test_cgls_equals_lstsq_after_n_iterations():
    """Krylov termination: with 4 distinct singular values CGLS is exact after 4
    iterations (up to rounding)."""
    rng = np.random.default_rng(0)
    m, n = 60, 12
    U, _ = np.linalg.qr(rng.standard_normal((m, n)) + 1j * rng.standard_normal((m, n)))
    V, _ = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    s = np.repeat([1.0, 0.5, 0.2, 0.1], 3)
    A = torch.tensor((U * s) @ V.conj().T, dtype=torch.complex128)
    b = torch.tensor(rng.standard_normal(m) + 1j * rng.standard_normal(m), dtype=torch.complex128)
    res = cgls(DenseOperator(A), b, tol=0.0, max_iter=4)
    ref = np.linalg.lstsq(A.numpy(), b.numpy(), rcond=None)[0]
    assert np.allclose(res.x.numpy(), ref, rtol=1e-8, atol=1e-10)


@pytest.mark.parametrize("lam", [1e-3, 0.1])
def test_cgls_regularised_normal_equations(lam):
    A, b = system(seed=1)
    res = cgls(DenseOperator(A), b, lam=lam, tol=1e-12, max_iter=100)
    An = A.numpy()
    ref = np.linalg.solve(An.conj().T @ An + lam * np.eye(An.shape[1]), An.conj().T @ b.numpy())
    assert np.allclose(res.x.numpy(), ref, rtol=1e-8, atol=1e-10)
    assert res.converged and true_residual(DenseOperator(A), res.x, b, lam) <= 1e-11


def test_cgls_agrees_with_scipy_lsqr():
    from scipy.sparse.linalg import lsqr
    A, b = system(seed=2, cond=5.0)
    res = cgls(DenseOperator(A), b, tol=1e-12, max_iter=50)
    ref = lsqr(A.numpy(), b.numpy(), atol=1e-14, btol=1e-14)[0]
    assert np.allclose(res.x.numpy(), ref, rtol=1e-7, atol=1e-9)


def test_fpcg_identity_and_exact_preconditioner():
    A, b = system(seed=3)
    op = DenseOperator(A)
    lam = 1e-2
    r1 = fpcg(op, b, lambda r: r, lam=lam, tol=1e-10, max_iter=100)
    r2 = cgls(op, b, lam=lam, tol=1e-10, max_iter=100)
    assert np.allclose(r1.x.numpy(), r2.x.numpy(), rtol=1e-6, atol=1e-8)
    K = A.conj().T @ A + lam * torch.eye(A.shape[1], dtype=A.dtype)
    Kinv = torch.linalg.inv(K)
    r3 = fpcg(op, b, lambda r: Kinv @ r, lam=lam, tol=1e-10, max_iter=100)
    assert r3.iterations == 1 and r3.converged and r3.fallback is None


@pytest.mark.parametrize("bad,reason", [(lambda r: r * float("nan"), "non-finite"),
                                        (lambda r: torch.zeros_like(r), "curvature")])
def test_fpcg_falls_back_to_cgls(bad, reason):
    """A preconditioner that returns NaN, or a zero direction (p^H K p = 0), hands
    the solve to CGLS, which still converges."""
    A, b = system(seed=4)
    res = fpcg(DenseOperator(A), b, bad, lam=1e-3, tol=1e-8, max_iter=200)
    assert res.fallback is not None and reason in res.fallback
    assert res.converged
    assert true_residual(DenseOperator(A), res.x, b, 1e-3) <= 1e-7


def test_fpcg_hands_over_to_cgls_when_the_residual_grows():
    """A preconditioner that rotates residuals by i points every step uphill; the
    solve must detect the growth, restart from the best iterate with CGLS and
    still converge."""
    torch = pytest.importorskip("torch")
    from sarsolver_amd.solvers import DenseOperator, fpcg, true_residual
    g = torch.Generator().manual_seed(9)
    A = torch.complex(torch.randn(50, 30, generator=g, dtype=torch.float64),
                      torch.randn(50, 30, generator=g, dtype=torch.float64))
    b = torch.complex(torch.randn(50, generator=g, dtype=torch.float64),
                      torch.randn(50, generator=g, dtype=torch.float64))
    op = DenseOperator(A)
    res = fpcg(op, b, lambda r: 1j * r, lam=0.1, tol=1e-8, max_iter=200)
    assert res.fallback is not None
    assert res.converged and true_residual(op, res.x, b, 0.1) <= 1e-7
