"""Apertures and datasets held as numpy arrays (reference: numpy/measurement.py)."""
from __future__ import annotations

from typing import List, Optional

import numpy as np
from numpy.typing import NDArray
from scipy.constants import c

from ..base import BaseSimpleSarAperture


class NumpySimpleSarAperture(BaseSimpleSarAperture):
    """Bistatic scalar SAR collection manifold (numpy/measurement.py:13-241).

    ``waveform_fft`` (centred order) defaults to ones of length ``num_freqs``;
    ``slow_time_weighting`` defaults to ones."""

    def __init__(self, trans_posns: NDArray, rec_posns: NDArray, srps: NDArray,
                 centre_freq: float, sample_freq: float, waveform_fft: Optional[NDArray] = None,
                 slow_time_weighting: Optional[NDArray] = None, num_freqs: Optional[int] = None,
                 sign_multiplier: float = 1.0, c_eff: float = c):
        if waveform_fft is None:
            if num_freqs is None:
                raise ValueError("Number of frequencies needs to be set somehow")
            waveform_fft = np.ones(num_freqs, dtype=np.complex128)
        self._transmitter_positions = trans_posns
        self._receiver_positions = rec_posns
        self._stab_ref_posns = srps
        self._centre_freq = centre_freq
        self._sample_bandwidth = sample_freq
        self._waveform_fft = waveform_fft
        self._num_freqs = waveform_fft.shape[0]
        self._sign_multiplier = sign_multiplier
        self._c_eff = c_eff
        self._num_slow_times = trans_posns.shape[0]
        if slow_time_weighting is None:
            slow_time_weighting = np.ones(self._num_slow_times, dtype=np.complex128)
        self._slow_time_weighting = slow_time_weighting

    def _sub(self, s: slice) -> "NumpySimpleSarAperture":
        # The reference rebuilds blocks without passing c_eff (numpy/measurement.py:77-103),
        # so blocked evaluation always uses the vacuum speed of light. Kept for parity.
        return NumpySimpleSarAperture(trans_posns=self.transmitter_positions[s, :],
                                      rec_posns=self.receiver_positions[s, :],
                                      srps=self.stab_ref_posns[s, :],
                                      centre_freq=self.centre_frequency,
                                      sample_freq=self.sample_bandwidth,
                                      waveform_fft=self.waveform_fft, num_freqs=self.num_freqs,
                                      sign_multiplier=self.sign_multiplier,
                                      slow_time_weighting=self.slow_time_weighting[s])

    def to_blocks(self, num_blocks: int = 1) -> List["NumpySimpleSarAperture"]:
        """Split into contiguous slow-time blocks (numpy/measurement.py:64-103)."""
        if num_blocks == 1:
            return [self._sub(slice(None))]
        step = self.num_slow_times // num_blocks
        bounds = [k * step for k in range(num_blocks)] + [self.num_slow_times]
        return [self._sub(slice(bounds[k], bounds[k + 1])) for k in range(num_blocks)]

    @property
    def sample_bandwidth(self) -> float:
        return self._sample_bandwidth

    @property
    def waveform_fft(self) -> NDArray:
        return self._waveform_fft

    @property
    def num_slow_times(self) -> int:
        return self._num_slow_times

    @property
    def stab_ref_posns(self) -> NDArray:
        return self._stab_ref_posns

    @property
    def num_freqs(self) -> int:
        return self._num_freqs

    @property
    def c_eff(self) -> float:
        return self._c_eff

    @property
    def centre_frequency(self) -> float:
        return self._centre_freq

    @property
    def transmitter_positions(self) -> NDArray:
        return self._transmitter_positions

    @property
    def sign_multiplier(self) -> float:
        return self._sign_multiplier

    @property
    def receiver_positions(self) -> NDArray:
        return self._receiver_positions

    @property
    def slow_time_weighting(self) -> NDArray:
        return self._slow_time_weighting

    def get_subset(self, item: slice = None) -> "NumpySimpleSarAperture":
        """Slow-time slice (numpy/measurement.py:224-241)."""
        return NumpySimpleSarAperture(self.transmitter_positions[item, :],
                                      self.receiver_positions[item, :],
                                      self.stab_ref_posns[item, :], self.centre_frequency,
                                      self.sample_bandwidth, self.waveform_fft,
                                      self.slow_time_weighting[item],
                                      sign_multiplier=self.sign_multiplier, c_eff=self.c_eff)


class NumpySimpleSarDataset:
    """An aperture plus its measured phase history [P, N] (centred order)
    (numpy/measurement.py:244-340; resampling constructors are not provided)."""

    def __init__(self, aperture: BaseSimpleSarAperture, phase_history: NDArray):
        if phase_history.shape != (aperture.num_slow_times, aperture.num_freqs):
            raise ValueError("phase history shape does not match the aperture")
        self._aperture = aperture
        self._phase_history = phase_history

    @property
    def aperture(self) -> BaseSimpleSarAperture:
        return self._aperture

    def phase_history(self, item: slice = slice(None)) -> NDArray:
        return self._phase_history[item, :]

    def get_subset(self, item: slice = None) -> "NumpySimpleSarDataset":
        return NumpySimpleSarDataset(self._aperture.get_subset(item), self._phase_history[item, :])
