"""Aperture / dataset interfaces, as in the reference's base/measurement.py."""
from __future__ import annotations

from abc import ABC, abstractmethod
from typing import List, Optional

import numpy as np
from numpy.typing import NDArray
from scipy.constants import pi


def normalise(x: NDArray) -> NDArray:
    """Unit vectors along the last axis (base/measurement.py:9-23)."""
    return x / np.linalg.norm(x, axis=-1, keepdims=True)


def bad_indices(x: NDArray) -> NDArray:
    """Row indices of a 2-D array containing a non-finite value (base/measurement.py:26-41)."""
    return np.argwhere(np.any(~np.isfinite(x), axis=-1))[:, 0]


def block_phase_history_array(m: NDArray, num_blocks: int = 1) -> List[NDArray]:
    """Split a [P, N] array into ``num_blocks`` slow-time blocks of P//num_blocks rows,
    the last taking the remainder (base/measurement.py:44-69)."""
    if num_blocks == 1:
        return [m]
    step = m.shape[0] // num_blocks
    bounds = [k * step for k in range(num_blocks)] + [m.shape[0]]
    return [m[bounds[k]:bounds[k + 1], :] for k in range(num_blocks)]


class BaseSimpleSarAperture(ABC):
    """A bistatic SAR collection: per-pulse Tx/Rx/SRP positions, fast-frequency sampling,
    waveform spectrum and slow-time weights (base/measurement.py:72-355)."""

    @property
    @abstractmethod
    def num_slow_times(self) -> int:
        raise NotImplementedError

    @property
    @abstractmethod
    def num_freqs(self) -> int:
        raise NotImplementedError

    @property
    @abstractmethod
    def centre_frequency(self) -> float:
        raise NotImplementedError

    @property
    @abstractmethod
    def sample_bandwidth(self) -> float:
        raise NotImplementedError

    @property
    def freqs(self) -> NDArray:
        """Fast frequencies in centred order: fc + fftshift(fftfreq(N, 1/fs))
        (base/measurement.py:129-139)."""
        return self.centre_frequency + np.fft.fftshift(np.fft.fftfreq(self.num_freqs,
                                                                      1.0 / self.sample_bandwidth))

    @property
    @abstractmethod
    def transmitter_positions(self) -> NDArray:
        raise NotImplementedError

    @property
    @abstractmethod
    def receiver_positions(self) -> NDArray:
        raise NotImplementedError

    @property
    @abstractmethod
    def stab_ref_posns(self) -> NDArray:
        raise NotImplementedError

    @property
    @abstractmethod
    def waveform_fft(self) -> NDArray:
        """[num_freqs] complex waveform spectrum, centred order."""
        raise NotImplementedError

    @property
    @abstractmethod
    def slow_time_weighting(self) -> NDArray:
        raise NotImplementedError

    @property
    @abstractmethod
    def sign_multiplier(self) -> float:
        raise NotImplementedError

    @property
    @abstractmethod
    def c_eff(self) -> float:
        raise NotImplementedError

    @abstractmethod
    def to_blocks(self, num_blocks: int = 1) -> List["BaseSimpleSarAperture"]:
        raise NotImplementedError

    def _unit_k_vecs(self, real_space_origin: Optional[NDArray] = None) -> NDArray:
        """Bistatic momentum-transfer directions at 1 Hz: unit(Rx - o) - unit(o - Tx)
        (base/measurement.py:247-270)."""
        o = (self.stab_ref_posns if real_space_origin is None
             else np.expand_dims(real_space_origin, axis=0))
        return normalise(self.receiver_positions - o) - normalise(o - self.transmitter_positions)

    def _band_edge_k_vecs(self, freq: float, real_space_origin=None) -> NDArray:
        return self._unit_k_vecs(real_space_origin) * (2.0 * pi / self.c_eff) * freq

    def edge_k_vecs(self, real_space_origin: Optional[NDArray] = None) -> NDArray:
        """k-vectors at the lower then upper band edge, stacked on axis 0
        (base/measurement.py:272-323)."""
        lo = self.centre_frequency - 0.5 * self.sample_bandwidth
        hi = self.centre_frequency + 0.5 * self.sample_bandwidth
        return np.concatenate([self._band_edge_k_vecs(lo, real_space_origin),
                               self._band_edge_k_vecs(hi, real_space_origin)], axis=0)

    @property
    def bad_slow_times(self) -> List[int]:
        """Sorted pulse indices with a non-finite Tx, Rx or SRP (base/measurement.py:325-339)."""
        bad = set(bad_indices(self.transmitter_positions).tolist())
        bad |= set(bad_indices(self.receiver_positions).tolist())
        bad |= set(bad_indices(self.stab_ref_posns).tolist())
        return sorted(bad)

    @abstractmethod
    def get_subset(self, item: slice = None) -> "BaseSimpleSarAperture":
        raise NotImplementedError


class BaseSimpleSarDataset(ABC):
    """An aperture with its measured phase history (base/measurement.py:358-402)."""

    @property
    @abstractmethod
    def aperture(self) -> BaseSimpleSarAperture:
        raise NotImplementedError

    @abstractmethod
    def phase_history(self, item: slice = slice(None)) -> Optional[NDArray]:
        """[num_requested, num_freqs] complex phase history (centred order)."""
        raise NotImplementedError

    @abstractmethod
    def get_subset(self, item: slice = None) -> "BaseSimpleSarDataset":
        raise NotImplementedError
