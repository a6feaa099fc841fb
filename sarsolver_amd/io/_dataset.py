"""One channel of a CPHD file as a dataset (reference: io/cphd101.py:13-285
and io/cphd03.py:13-289), read with sarsolver_amd.io.cphd instead of sarpy.

The datasets expose the reference's properties: positions from the per-vector
parameters (TxPos, RcvPos, SRPPos), centre frequency f0 + N df / 2 and
sample bandwidth N df with (f0, df) = (SC0, SCSS) for 1.0.x and
(Fx0, Fx_SS) for 0.3 read from the first vector, the sign convention from
Global/SGN (Global/PhaseSGN), unit waveform and slow-time weighting unless
given.  ``phase_history(item)`` returns host complex64 samples;
``phase_history_device(item, device)`` moves the raw big-endian bytes to the
GPU and decodes them there (`sarhip_decode_signal`), so 8- and 16-bit files
cross the bus at a quarter or half of the complex64 size.
"""
from __future__ import annotations

from typing import List, Optional

import numpy as np
from numpy.typing import NDArray
from scipy.constants import c

from ..base import BaseSimpleSarDataset
from ..numpy import NumpySimpleSarAperture, NumpySimpleSarDataset
from .cphd import CphdFile, open_cphd


class _CphdDataset(BaseSimpleSarDataset):
    _F0, _DF, _VERSION = "SC0", "SCSS", "1."

    def __init__(self, file_address: str, channel_index: int = 0, c_eff: float = c,
                 waveform_fft: Optional[NDArray] = None,
                 slow_time_weighting: Optional[NDArray] = None):
        self._file_address = str(file_address)
        self._file: Optional[CphdFile] = open_cphd(self._file_address)
        if not self._file.version.startswith(self._VERSION):
            raise ValueError(f"{file_address} is CPHD {self._file.version}, "
                             f"not {self._VERSION}x")
        if not 0 <= channel_index < len(self._file.channels):
            raise IndexError(f"channel {channel_index} not in file "
                             f"({len(self._file.channels)} channels)")
        self._channel_index = channel_index
        self._waveform_fft = waveform_fft
        self._slow_time_weighting = slow_time_weighting
        self._c_eff = c_eff
        ch = self._file.channels[channel_index]
        # the reference takes the smaller of NumVectors and the number of
        # transmitter positions read back (io/cphd101.py:39-42)
        self._num_slow_times = min(ch.num_vectors, self.transmitter_positions.shape[0])
        self._num_freqs = ch.num_samples
        self._sign_multiplier = 1.0 if self._file.sign > 0 else -1.0

    @property
    def reader(self) -> CphdFile:
        if self._file is None:
            self._file = open_cphd(self._file_address)
        return self._file

    def _pvp(self, name: str, index=slice(None)) -> NDArray:
        return self.reader.pvp_field(name, self._channel_index, index)

    @property
    def aperture(self) -> NumpySimpleSarAperture:
        return NumpySimpleSarAperture(self.transmitter_positions, self.receiver_positions,
                                      self.stab_ref_posns, self.centre_frequency,
                                      self.sample_bandwidth, self.waveform_fft,
                                      self.slow_time_weighting, num_freqs=None,
                                      sign_multiplier=self.sign_multiplier, c_eff=self.c_eff)

    @property
    def num_slow_times(self) -> int:
        return self._num_slow_times

    @property
    def num_freqs(self) -> int:
        return self._num_freqs

    @property
    def centre_frequency(self) -> float:
        f0 = float(self._pvp(self._F0, 0))
        return f0 + 0.5 * self.sample_bandwidth

    @property
    def sample_bandwidth(self) -> float:
        return self.num_freqs * float(self._pvp(self._DF, 0))

    @property
    def transmitter_positions(self) -> NDArray:
        return self._pvp("TxPos")

    @property
    def receiver_positions(self) -> NDArray:
        return self._pvp("RcvPos")

    @property
    def stab_ref_posns(self) -> NDArray:
        return self._pvp("SRPPos")

    @property
    def waveform_fft(self) -> NDArray:
        if self._waveform_fft is None:
            self._waveform_fft = np.ones(self.num_freqs, dtype=np.complex128)
        return self._waveform_fft

    @property
    def slow_time_weighting(self) -> NDArray:
        if self._slow_time_weighting is None:
            self._slow_time_weighting = np.ones(self.num_slow_times, dtype=np.complex128)
        return self._slow_time_weighting

    @property
    def sign_multiplier(self) -> float:
        return self._sign_multiplier

    @property
    def c_eff(self) -> float:
        return self._c_eff

    def to_blocks(self, num_blocks: int = 1) -> List[NumpySimpleSarDataset]:
        """Datasets without phase history over num_blocks slow-time blocks
        (io/cphd101.py:203-244)."""
        n = self.num_slow_times
        bounds = [0, n] if num_blocks == 1 else \
            [b * (n // num_blocks) for b in range(num_blocks)] + [n]
        tx, rx, srp = self.transmitter_positions, self.receiver_positions, self.stab_ref_posns
        w = self.slow_time_weighting
        return [NumpySimpleSarDataset.from_scratch(
            trans_posns=tx[a:b], rec_posns=rx[a:b], srps=srp[a:b],
            centre_freq=self.centre_frequency, sample_freq=self.sample_bandwidth,
            c_eff=self.c_eff, waveform_fft=self.waveform_fft, num_freqs=self.num_freqs,
            sign_multiplier=self.sign_multiplier, slow_time_weighting=w[a:b],
            phase_history=None) for a, b in zip(bounds[:-1], bounds[1:])]

    def phase_history(self, item: slice = slice(None)) -> NDArray:
        """Decoded complex64 samples [n, num_freqs] (host)."""
        return self.reader.signal(self._channel_index, item)

    def phase_history_device(self, item: slice = slice(None), device=None):
        """The same samples decoded on the GPU: a complex64 torch tensor."""
        from .._native import decode_signal
        f = self.reader
        ch = f.channels[self._channel_index]
        start, stop, step = item.indices(ch.num_vectors)
        if step != 1:
            raise ValueError("phase_history_device takes contiguous slices")
        raw = f.raw_signal(self._channel_index, start, stop)
        sf = f.amp_sf(self._channel_index, start, stop)
        return decode_signal(raw, f.format_code, ch.num_samples, sf, device)

    def __getstate__(self):
        state = dict(self.__dict__)
        state["_file"] = None
        return state

    def __setstate__(self, state):
        self.__dict__ = state

    def get_subset(self, item: slice = None) -> NumpySimpleSarDataset:
        item = slice(None) if item is None else item
        return NumpySimpleSarDataset(self.aperture.get_subset(item), self.phase_history(item))
