"""The FP64 oracle, pinned (CPU only).

Each check is chosen so that a plausible slip (a dropped term, a wrong sign
or index, a transposed operand, a shifted bin rule) fails at least one:
  * FFT vs numpy.fft (pocketfft), powers of two and 2^7 * 251 (Bluestein);
  * operators vs a brute-force direct sum without FFTs (oracle/brute.py),
    all interpolation modes, odd and even N, bistatic, complex weights;
  * the reference's own fixtures: little_backproj.npy (A^H) and
    little_soln.npy (A and A^H through two PDHG updates);
  * closed forms: a scatterer at the SRP; (A^H A e_m)_m = P*N for unit weights;
  * exact bin ties (u = k +- 1/2) and negative u;
  * adjointness of the consistent modes, non-adjointness of the reference pair;
  * blocking invariance over thread counts.
"""
import numpy as np
import pytest

from conftest import LITTLE_NORM, rel_l2

import oracle.brute as B


def rand_c(rng, shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


@pytest.mark.parametrize("n", [1, 2, 8, 64, 251, 1024, 4096, 32128, 7, 12])
@pytest.mark.parametrize("sign", [-1, 1])
def test_fft_matches_numpy(oracle_lib, n, sign):
    rng = np.random.default_rng(n)
    x = rand_c(rng, n)
    ref = np.fft.fft(x) if sign == -1 else np.fft.ifft(x) * n
    got = oracle_lib.fft(x, sign)
    assert np.abs(got - ref).max() <= 1e-13 * max(1.0, np.abs(ref).max())


def _tiny_problem(rng, P, N, M, bistatic=True):
    tx = np.stack([2000 + 10 * rng.standard_normal(P), 30 * rng.standard_normal(P),
                   500 + rng.standard_normal(P)], 1)
    rx = tx + (50 * rng.standard_normal((P, 3)) if bistatic else 0.0)
    srp = 0.1 * rng.standard_normal((P, 3))
    pos = 3 * rng.standard_normal((M, 3))
    return tx, rx, srp, pos, rand_c(rng, M), rand_c(rng, N), rand_c(rng, P)


@pytest.mark.parametrize("N", [7, 8, 9, 16])
@pytest.mark.parametrize("interp", ["nearest_ref", "nearest", "linear"])
@pytest.mark.parametrize("bistatic", [True, False])
def test_operators_match_brute_force(oracle_lib, N, interp, bistatic):
    rng = np.random.default_rng(100 + N)
    P, M = 5, 6
    tx, rx, srp, pos, a, W, w = _tiny_problem(rng, P, N, M, bistatic)
    fc, fs, U, sgn = 1.0e9, 1.0e8, 4.5, -1.0 if N % 2 else 1.0
    Y = oracle_lib.forward(tx, rx, srp, pos, a, W, w, fc, fs, N, upsample_ratio=U,
                           sign_multiplier=sgn, interp=interp)
    Yb = B.forward(tx, rx, srp, pos, a, W, w, fc, fs, N, U=U, sgn=sgn, interp=interp)
    assert rel_l2(Y, Yb) <= 1e-12
    y = rand_c(rng, (P, N))
    x = oracle_lib.adjoint(tx, rx, srp, pos, y, W, w, fc, fs, upsample_ratio=U,
                           sign_multiplier=sgn, interp=interp)
    xb = B.adjoint(tx, rx, srp, pos, y, W, w, fc, fs, U=U, sgn=sgn, interp=interp)
    assert rel_l2(x, xb) <= 1e-12


@pytest.mark.parametrize("interp", ["nearest", "linear"])
def test_consistent_modes_are_adjoint(oracle_lib, interp):
    rng = np.random.default_rng(3)
    P, N, M = 9, 16, 40
    tx, rx, srp, pos, a, W, w = _tiny_problem(rng, P, N, M)
    y = rand_c(rng, (P, N))
    Y = oracle_lib.forward(tx, rx, srp, pos, a, W, w, 1e9, 1e8, N, upsample_ratio=8.0,
                           interp=interp)
    x = oracle_lib.adjoint(tx, rx, srp, pos, y, W, w, 1e9, 1e8, upsample_ratio=8.0, interp=interp)
    err = abs(np.vdot(y, Y) - np.vdot(x, a)) / (np.linalg.norm(Y) * np.linalg.norm(y))
    assert err <= 1e-13


def test_reference_pair_is_not_adjoint(oracle_lib):
    """The reference's adjoint reads bin round(0.5+u) while its forward writes round(u)
    (sarsolver_cxx.cpp:289 vs :347): the dot-product test must fail for nearest_ref."""
    rng = np.random.default_rng(4)
    P, N, M = 9, 16, 40
    tx, rx, srp, pos, a, W, w = _tiny_problem(rng, P, N, M)
    y = rand_c(rng, (P, N))
    Y = oracle_lib.forward(tx, rx, srp, pos, a, W, w, 1e9, 1e8, N, upsample_ratio=2.0)
    x = oracle_lib.adjoint(tx, rx, srp, pos, y, W, w, 1e9, 1e8, upsample_ratio=2.0)
    err = abs(np.vdot(y, Y) - np.vdot(x, a)) / (np.linalg.norm(Y) * np.linalg.norm(y))
    assert err > 1e-3


def test_scatterer_at_srp_gives_flat_spectrum(oracle_lib):
    """dR = 0 -> bin 0, phase 0: Y[p, :] = a * w_p * W (CPP:284-307 closed form)."""
    rng = np.random.default_rng(5)
    P, N = 6, 10
    tx = 1000 * rng.standard_normal((P, 3))
    rx = 1000 * rng.standard_normal((P, 3))
    srp = np.tile(np.array([[1.5, -2.0, 0.25]]), (P, 1))
    W, w = rand_c(rng, N), rand_c(rng, P)
    a = np.array([0.3 - 0.7j])
    Y = oracle_lib.forward(tx, rx, srp, srp[:1], a, W, w, 9e9, 5e8, N, upsample_ratio=3.0)
    assert np.allclose(Y, a[0] * w[:, None] * W[None, :], rtol=1e-14, atol=0)


def test_linear_phase_ramp_at_integer_bin(oracle_lib):
    """A scatterer whose dR is exactly k bins gives Y[p, f] = W_f w_p a e^{i k_c dR}
    e^{-2 pi i nu_f k / N_w}: a pure linear phase ramp across frequency."""
    N, U = 12, 4.0
    nw = int(U * N)
    c, fs, fc = 1.0, 1.0, 0.3
    by_dx = fs * nw / (N * c)  # = 4 bins per metre
    k = 3
    # monostatic at +x, SRP at origin, scatterer on the x-axis: dR = -2 d
    tx = np.array([[100.0, 0.0, 0.0]])
    d = k / by_dx / 2.0
    pos = np.array([[-d, 0.0, 0.0]])
    Y = oracle_lib.forward(tx, tx, np.zeros((1, 3)), pos, np.array([1.0 + 0j]), np.ones(N),
                           np.ones(1), fc, fs, N, c_eff=c, upsample_ratio=U)
    nu = np.rint(np.fft.fftfreq(N) * N)
    dr = 2 * d
    expect = np.exp(1j * 2 * np.pi * fc / c * dr) * np.exp(-2j * np.pi * nu * k / nw)
    assert np.allclose(Y[0], expect, rtol=0, atol=1e-12)


@pytest.mark.parametrize("u,fwd,adj", [(2.5, 3, 3), (-2.5, -3, -2), (3.49999, 3, 4),
                                        (-0.5, -1, 0), (0.5, 1, 1), (4.0, 4, 5), (-4.0, -4, -4)])
def test_bin_rule_ties(oracle_lib, u, fwd, adj):
    """round() is half away from zero (C/C++), not numpy's half-to-even; the adjoint
    adds 0.5 before rounding (CPP:289, CPP:347); negative bins wrap mod N_w."""
    nw = 16
    assert oracle_lib.bin_index(u, 1.0, nw, "nearest_ref", adjoint=False) == fwd % nw
    assert oracle_lib.bin_index(u, 1.0, nw, "nearest_ref", adjoint=True) == adj % nw
    assert oracle_lib.bin_index(u, 1.0, nw, "nearest", adjoint=True) == fwd % nw
    n0, f = oracle_lib.bin_index(u, 1.0, nw, "linear")
    assert n0 == int(np.floor(u)) % nw and f == pytest.approx(u - np.floor(u), abs=0)


def test_working_length_truncates(oracle_lib):
    """N_w = (size_t)(U * N) in double (CPP:223)."""
    assert oracle_lib.working_len(251, 128.0) == 32128
    assert oracle_lib.working_len(10, 2.55) == 25
    assert oracle_lib.working_len(3, 1.0 / 3.0 * 3.0) == 3


def test_normal_operator_diagonal(oracle_lib):
    """(A^H A e_m)_m = sum_p |w_p|^2 sum_f |W_f|^2 for the consistent nearest pair."""
    rng = np.random.default_rng(8)
    P, N, M = 7, 11, 5
    tx, rx, srp, pos, _, W, w = _tiny_problem(rng, P, N, M)
    e = np.zeros(M, complex)
    e[2] = 1.0
    Y = oracle_lib.forward(tx, rx, srp, pos, e, W, w, 1e9, 1e8, N, upsample_ratio=6.0,
                           interp="nearest")
    x = oracle_lib.adjoint(tx, rx, srp, pos, Y, W, w, 1e9, 1e8, upsample_ratio=6.0,
                           interp="nearest")
    expect = np.sum(np.abs(w) ** 2) * np.sum(np.abs(W) ** 2)
    assert x[2] == pytest.approx(expect, rel=1e-12)


@pytest.mark.parametrize("threads", [3, 7])
def test_blocking_invariance(oracle_lib, threads):
    rng = np.random.default_rng(9)
    P, N, M = 6, 16, 101
    tx, rx, srp, pos, a, W, w = _tiny_problem(rng, P, N, M)
    y = rand_c(rng, (P, N))
    Y1 = oracle_lib.forward(tx, rx, srp, pos, a, W, w, 1e9, 1e8, N, upsample_ratio=4.0)
    Yt = oracle_lib.forward(tx, rx, srp, pos, a, W, w, 1e9, 1e8, N, upsample_ratio=4.0,
                            num_threads=threads)
    assert rel_l2(Yt, Y1) <= 1e-14
    x1 = oracle_lib.adjoint(tx, rx, srp, pos, y, W, w, 1e9, 1e8, upsample_ratio=4.0)
    xt = oracle_lib.adjoint(tx, rx, srp, pos, y, W, w, 1e9, 1e8, upsample_ratio=4.0,
                            num_threads=threads)
    assert np.array_equal(x1, xt)


def _little_ops(O, ap, sc):
    pos = sc.positions
    W = np.fft.ifftshift(ap.waveform_fft)
    tx, srp, w = ap.transmitter_positions, ap.stab_ref_posns, ap.slow_time_weighting
    fc, fs, N = ap.centre_frequency, ap.sample_bandwidth, ap.num_freqs

    def A(x):
        Y = O.forward(tx, tx, srp, pos, x.reshape(-1), W, w, fc, fs, N, upsample_ratio=128.0,
                      num_threads=8)
        return np.fft.fftshift(Y, axes=1)

    def AH(y):
        return O.adjoint(tx, tx, srp, pos, np.fft.ifftshift(y, axes=1), W, w, fc, fs,
                         upsample_ratio=128.0, num_threads=8)
    return A, AH


def test_little_scene_rebuild(little, little_arrays):
    """The rebuilt grid has the fixture's shape and the backprojection peaks sit on the
    ground-truth scatterers with the closed-form height P*N (= 62 499) x amplitude."""
    ap, sc = little
    assert tuple(sc.grid_shape) == little_arrays["gt"].shape == (1, 244, 241)
    gt, bp = little_arrays["gt"], little_arrays["backproj"]
    idx = tuple(np.argwhere(np.abs(gt) > 0).T)
    assert np.allclose(bp[idx] / gt[idx], 249 * 251, rtol=0.01)


def test_little_backproj_fixture(oracle_lib, little, little_arrays):
    """A^H(little_noisy_meas) vs the reference's little_backproj.npy: elementwise
    rel 1e-3 (reference tests/test_cil.py:15) and relative L2 1e-12."""
    ap, sc = little
    _, AH = _little_ops(oracle_lib, ap, sc)
    x = AH(little_arrays["noisy_meas"]).reshape(sc.grid_shape)
    ref = little_arrays["backproj"]
    assert rel_l2(x, ref) <= 1e-12
    assert np.all(np.abs(x - ref) <= 1e-3 * np.abs(ref) + 1e-12)


def test_little_pdhg_fixture(oracle_lib, little, little_arrays):
    """Two PDHG updates (sigma = tau = 1/||A||, f = ||Ax - b||^2, g = 0) reproduce the
    reference's little_soln.npy, which pins the forward operator too."""
    ap, sc = little
    A, AH = _little_ops(oracle_lib, ap, sc)
    b = little_arrays["noisy_meas"]
    s = t = 1.0 / LITTLE_NORM
    x = np.zeros(sc.grid_shape, complex)
    xbar = x.copy()
    y = np.zeros_like(b)
    for _ in range(2):
        x_old = x
        y = (y + s * A(xbar) - s * b) / (1 + s / 2)
        x = x_old - t * AH(y).reshape(sc.grid_shape)
        xbar = 2 * x - x_old
    assert rel_l2(x, little_arrays["soln"]) <= 1e-12
