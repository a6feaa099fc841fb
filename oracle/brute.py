"""Brute-force direct sums of the SAR operators, no FFT (numpy, tiny inputs only).

TEST INFRASTRUCTURE ONLY.  An independent formulation used to pin the C
oracle: the reference's per-pulse scatter + unnormalised FFT + spectral
truncation (src/sarsolver/cxx_binding/sarsolver_cxx.cpp:271-310) equals

    Y[p, k] = W_k w_p sum_m a_m e^{i k_c dR_pm} e^{-2 pi i nu_k n_pm / N_w}

and its backprojection (sarsolver_cxx.cpp:319-353) equals

    x_m = sum_p w_p' e^{-i k_c dR_pm} sum_k conj(W_k) Y[p, k] e^{+2 pi i nu_k n'_pm / N_w}

where nu_k is the signed FFT-order frequency of index k (numpy.fft.fftfreq
times N) and n, n' are the bin indices of the interpolation rule.  Ranges are
taken from numpy (np.sqrt of an explicit three-term sum), bins from a
round-half-away-from-zero written with floor.
"""
from __future__ import annotations

import numpy as np

C_LIGHT = 299792458.0


def round_half_away(u):
    u = np.asarray(u, dtype=np.float64)
    return np.sign(u) * np.floor(np.abs(u) + 0.5)


def ranges(tx, rx, srp, pos):
    tx = np.asarray(tx, np.float64).reshape(-1, 3)
    rx = np.asarray(rx, np.float64).reshape(-1, 3)
    srp = np.asarray(srp, np.float64).reshape(-1, 3)
    pos = np.asarray(pos, np.float64).reshape(-1, 3)

    def dist(a, b):  # a [P,3], b [M,3] -> [P,M]
        d = a[:, None, :] - b[None, :, :]
        return np.sqrt(d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1] + d[..., 2] * d[..., 2])

    r = dist(tx, pos) + dist(rx, pos)
    d_t = tx - srp
    d_r = rx - srp
    r_srp = (np.sqrt(d_t[:, 0] * d_t[:, 0] + d_t[:, 1] * d_t[:, 1] + d_t[:, 2] * d_t[:, 2])
             + np.sqrt(d_r[:, 0] * d_r[:, 0] + d_r[:, 1] * d_r[:, 1] + d_r[:, 2] * d_r[:, 2]))
    return r - r_srp[:, None]


def constants(N, fs, fc, c_eff, U, sgn):
    nw = int(U * float(N))
    by_dx = (fs * float(nw)) / (float(N) * c_eff)
    kc = 2.0 * sgn * np.pi * fc / c_eff
    return nw, by_dx, kc


def signed_freq_index(N):
    return np.rint(np.fft.fftfreq(N) * N).astype(np.int64)


def _weights(dr, by_dx, nw, interp, adjoint):
    """Per (p, m): list of (bin, weight) pairs."""
    u = dr * by_dx
    if interp == "linear":
        n0 = np.floor(u)
        f = u - n0
        return [(np.mod(n0, nw), 1.0 - f), (np.mod(n0 + 1, nw), f)]
    if interp == "nearest_ref" and adjoint:
        n = round_half_away(0.5 + u)
    else:
        n = round_half_away(u)
    return [(np.mod(n, nw), np.ones_like(u))]


def forward(tx, rx, srp, pos, amps, W_fft, w, fc, fs, N, *, c_eff=C_LIGHT, U=128.0, sgn=1.0,
            interp="nearest_ref"):
    nw, by_dx, kc = constants(N, fs, fc, c_eff, U, sgn)
    dr = ranges(tx, rx, srp, pos)                          # [P, M]
    nu = signed_freq_index(N)                               # [N]
    c = np.exp(1j * dr * kc) * np.asarray(amps)[None, :] * np.asarray(w)[:, None]
    Y = np.zeros((dr.shape[0], N), dtype=np.complex128)
    for n, wt in _weights(dr, by_dx, nw, interp, False):
        ph = np.exp(-2j * np.pi * (n[:, :, None] * nu[None, None, :] % nw) / nw)  # [P, M, N]
        Y += np.einsum("pm,pmk->pk", c * wt, ph)
    return Y * np.asarray(W_fft)[None, :]


def adjoint(tx, rx, srp, pos, Y_fft, W_fft, w, fc, fs, *, c_eff=C_LIGHT, U=128.0, sgn=1.0,
            interp="nearest_ref"):
    Y_fft = np.asarray(Y_fft)
    P, N = Y_fft.shape
    nw, by_dx, kc = constants(N, fs, fc, c_eff, U, sgn)
    dr = ranges(tx, rx, srp, pos)
    nu = signed_freq_index(N)
    K = np.conj(np.asarray(W_fft))[None, :] * Y_fft        # [P, N]
    wp = np.asarray(w, dtype=np.complex128)
    if interp != "nearest_ref":
        wp = np.conj(wp)
    x = np.zeros(dr.shape[1], dtype=np.complex128)
    for n, wt in _weights(dr, by_dx, nw, interp, True):
        ph = np.exp(2j * np.pi * (n[:, :, None] * nu[None, None, :] % nw) / nw)    # [P, M, N]
        h = np.einsum("pk,pmk->pm", K, ph)
        x += np.sum(wt * h * np.exp(-1j * dr * kc) * wp[:, None], axis=0)
    return x
