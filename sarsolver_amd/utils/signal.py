"""Spectral helpers of dataset resampling (reference: utils/signal.py).

Host (numpy) versions of the reference's public utilities, operating on
centred-order phase histories.  The resampling of a whole dataset runs on the
GPU (`NumpySimpleSarDataset.from_resampling`, `sarhip_resample`); these
functions serve single rows (the waveform) and callers of the reference API.
"""
from __future__ import annotations

import numpy as np
from numpy.typing import NDArray
from scipy.signal.windows import hann, tukey


def get_downsample_kernel(downsample_ratio: float) -> NDArray:
    """Slow-time kernel: a five-point spectral stencil (two lowest positive and
    the lowest negative frequency of a round(5 r)-point grid) brought to the
    time domain, centred and Hann-weighted (utils/signal.py:7-25)."""
    num_samples = int(np.round(5 * downsample_ratio))
    spectrum = np.zeros(num_samples, dtype="complex64")
    spectrum[:2] = 1.0
    spectrum[-1:] = 1.0
    return np.fft.fftshift(np.fft.ifft(spectrum)) * hann(num_samples)


def downsample_window(new_num_freqs: int, safety_factor: float = 1.2) -> NDArray:
    """The Tukey taper (alpha = safety_factor - 1, at least 0) of the cropped
    range profiles, in FFT order (utils/signal.py:62)."""
    return np.fft.fftshift(tukey(new_num_freqs, alpha=(max(safety_factor, 1.0) - 1.0)))


def fast_freq_downsample(phase_history: NDArray, new_num_freqs: int,
                         safety_factor: float = 1.2) -> NDArray:
    """Keep the first floor(n/2) and last n - floor(n/2) samples of each range
    profile (n = new_num_freqs), taper them and return to frequency
    (utils/signal.py:45-73).  [P, N] -> [P, n], centred order."""
    prof = np.fft.ifft(np.fft.ifftshift(phase_history, axes=[1]), axis=1)
    lo = new_num_freqs // 2
    hi = new_num_freqs - lo
    cropped = np.concatenate([prof[:, :lo], prof[:, prof.shape[1] - hi:]], axis=1)
    out = np.fft.fft(downsample_window(new_num_freqs, safety_factor)[None, :] * cropped, axis=1)
    return np.fft.fftshift(out, axes=[1])


def to_db_abs(x: NDArray) -> NDArray:
    """10 log10(|x|), floored at -300 dB (utils/signal.py:28-42)."""
    return np.real(10.0 * np.log10(np.abs(x) + 1.0E-30))
