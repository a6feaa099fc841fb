"""FP64 numpy restatement of the reference's dataset resampling (test
infrastructure: only the tests and the benchmark's CPU baseline call it).

Follows, step by step:
  * get_downsample_kernel           utils/signal.py:7-25
  * fast_freq_downsample            utils/signal.py:45-73
  * move_srps                       utils/geodesy.py:33-73
  * NumpySimpleSarDataset.from_resampling, the per-window loop
                                    numpy/measurement.py:342-457
The window loop is restated once per output pulse exactly as the reference
runs it (move_srps and fast_freq_downsample on each window's pulses), so the
restatement also pins that a GPU path may do them once per input pulse.
"""
from __future__ import annotations

import numpy as np
from scipy.constants import c as C_LIGHT
from scipy.constants import pi
from scipy.signal.windows import hann, tukey


def downsample_kernel(downsample_ratio: float) -> np.ndarray:
    """Five-point-stencil spectral kernel with a Hann weighting (signal.py:7-25)."""
    num_samples = int(np.round(5 * downsample_ratio))
    freq_filter = np.zeros(num_samples, dtype="complex64")
    freq_filter[:2] = 1.0
    freq_filter[-1:] = 1.0
    return np.fft.fftshift(np.fft.ifft(freq_filter)) * hann(num_samples)


def fast_freq_downsample(phase_history: np.ndarray, new_num_freqs: int,
                         safety_factor: float = 1.2) -> np.ndarray:
    """Crop the range profiles to new_num_freqs samples with a Tukey taper
    (signal.py:45-73); input and output in centred frequency order."""
    safety_factor = max(safety_factor, 1.0)
    window = np.fft.fftshift(tukey(new_num_freqs, alpha=(safety_factor - 1.0)))[None, :]
    out = np.zeros((phase_history.shape[0], new_num_freqs), dtype=np.complex128)
    prof = np.fft.ifft(np.fft.ifftshift(phase_history, axes=[1]), axis=1)
    lo = new_num_freqs // 2
    hi = new_num_freqs - lo
    out[:, :lo] = prof[:, :lo]
    out[:, -hi:] = prof[:, -hi:]
    return np.fft.fftshift(np.fft.fft(window * out, axis=1), axes=[1])


def srp_range_change(old_srps, new_srp, tx, rx) -> np.ndarray:
    """new minus old bistatic range to the SRP, per pulse (geodesy.py:64-66)."""
    new_srp = np.asarray(new_srp, dtype=np.float64)[None, :]
    new = np.linalg.norm(tx - new_srp, axis=-1) + np.linalg.norm(rx - new_srp, axis=-1)
    old = np.linalg.norm(tx - old_srps, axis=-1) + np.linalg.norm(rx - old_srps, axis=-1)
    return new - old


def move_srps(phase_history, old_srps, new_srp, tx, rx, centre_freq, sample_bandwidth,
              sign_factor=1.0, c_eff=C_LIGHT) -> np.ndarray:
    """Rephase to a new SRP (geodesy.py:33-73)."""
    n = phase_history.shape[1]
    dr = srp_range_change(old_srps, new_srp, tx, rx)[:, None]
    bb_k = (2.0 * pi / c_eff) * np.fft.fftshift(np.fft.fftfreq(n, 1.0 / sample_bandwidth))[None, :]
    ramp = np.exp(1.j * (bb_k * dr))
    centre = np.exp((-2.0j * pi * sign_factor * centre_freq / c_eff) * dr)
    return phase_history * ramp * centre


def resample(phase_history, tx, rx, srps, slow_weights, waveform_fft, centre_freq,
             sample_bandwidth, sign_factor, c_eff, bad, new_srp, new_num_freqs,
             slow_factor, safety_factor=1.2):
    """The window loop of from_resampling (measurement.py:397-457) for given
    factors.  Returns (tx, rx, srps, weights, phase history, waveform) of the
    resampled dataset."""
    kern = downsample_kernel(slow_factor)[:, None]
    K = kern.shape[0]
    P = phase_history.shape[0]
    f = int(slow_factor)
    new_p = ((P - K) // f) - 1
    new_wave = fast_freq_downsample(np.asarray(waveform_fft)[None, :], new_num_freqs,
                                    safety_factor)[0]
    bad = set(int(b) for b in bad)
    T, R, S, Wt, Ph = [], [], [], [], []
    for i in range(new_p):
        o = f * i
        idx = list(range(o, o + K))
        if bad & set(idx):
            continue
        T.append(tx[o + K // 2])
        R.append(rx[o + K // 2])
        S.append(np.asarray(new_srp, dtype=np.float64))
        sh = move_srps(phase_history[o:o + K], srps[o:o + K], new_srp, tx[o:o + K], rx[o:o + K],
                       centre_freq, sample_bandwidth, sign_factor, c_eff)
        ds = fast_freq_downsample(sh, new_num_freqs, safety_factor)
        Ph.append(np.sum(ds * kern, axis=0))
        Wt.append(np.sum(kern[:, 0] * slow_weights[o:o + K]))
    return (np.array(T), np.array(R), np.array(S), np.array(Wt), np.array(Ph), new_wave)
