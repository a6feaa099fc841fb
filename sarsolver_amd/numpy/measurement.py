"""Apertures and datasets held as numpy arrays (reference: numpy/measurement.py)."""
from __future__ import annotations

from typing import List, Optional

import numpy as np
from numpy.typing import NDArray
from scipy.constants import c

from ..base import BaseSimpleSarAperture, BaseSimpleSarDataset


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


class NumpySimpleSarDataset(BaseSimpleSarDataset):
    """An aperture plus its measured phase history [P, N] (centred order)
    (numpy/measurement.py:244-501)."""

    def __init__(self, aperture: BaseSimpleSarAperture, phase_history: Optional[NDArray]):
        if phase_history is not None and \
                phase_history.shape != (aperture.num_slow_times, aperture.num_freqs):
            raise ValueError("phase history shape does not match the aperture")
        self._aperture = aperture
        self._phase_history = phase_history

    @staticmethod
    def from_scratch(trans_posns: NDArray, rec_posns: NDArray, srps: NDArray, centre_freq: float,
                     sample_freq: float, phase_history: Optional[NDArray] = None,
                     waveform_fft: Optional[NDArray] = None,
                     slow_time_weighting: Optional[NDArray] = None, num_freqs: Optional[int] = None,
                     sign_multiplier: float = 1.0, c_eff: float = c) -> "NumpySimpleSarDataset":
        """Aperture and dataset from the aperture's arguments (numpy/measurement.py:258-293)."""
        aperture = NumpySimpleSarAperture(trans_posns=trans_posns, rec_posns=rec_posns, srps=srps,
                                          centre_freq=centre_freq, sample_freq=sample_freq,
                                          waveform_fft=waveform_fft,
                                          slow_time_weighting=slow_time_weighting,
                                          num_freqs=num_freqs, sign_multiplier=sign_multiplier,
                                          c_eff=c_eff)
        return NumpySimpleSarDataset(aperture, phase_history)

    @property
    def aperture(self) -> BaseSimpleSarAperture:
        return self._aperture

    def phase_history(self, item: slice = slice(None)) -> Optional[NDArray]:
        if self._phase_history is None:
            return None
        return self._phase_history[item]

    def remove_phase_history(self):
        self._phase_history = None

    def get_subset(self, item: slice = None) -> "NumpySimpleSarDataset":
        ph = None if self._phase_history is None else self._phase_history[item, :]
        return NumpySimpleSarDataset(self._aperture.get_subset(item), ph)

    @staticmethod
    def from_resampling(original_dataset: "NumpySimpleSarDataset", range_amb: Optional[float] = None,
                        cross_range_amb: Optional[float] = None,
                        fast_downsample_factor: float = 1.0, slow_downsample_factor: int = 1,
                        new_srp: Optional[NDArray] = None, safety_factor: float = 1.2,
                        num_threads: Optional[int] = None,
                        device=None) -> "NumpySimpleSarDataset":
        """A smaller dataset resampled about `new_srp` (numpy/measurement.py:341-457).

        The factors, the new aperture and the windows (every `slow` pulses, K =
        round(5 slow) input pulses each, windows touching a bad pulse dropped)
        follow the reference; the phase history is computed on the GPU
        (`sarhip_resample`): each input pulse is moved to the new SRP and
        cropped in range once, then each window is summed with the slow-time
        kernel.  `num_threads` is accepted for signature compatibility.  A
        cross-range ambiguity that implies a factor below 1 gives factor 1 (the
        reference sets the float 1.0 there, which its integer indexing then
        rejects)."""
        import torch

        from .. import _native
        from ..utils import OrientedBoundingBox, get_downsample_kernel
        from ..utils.geodesy import srp_range_change
        from ..utils.signal import downsample_window, fast_freq_downsample
        ap = original_dataset.aperture
        if new_srp is None:
            new_srp = ap.stab_ref_posns[0]
        new_srp = np.asarray(new_srp, dtype=np.float64)
        obb = OrientedBoundingBox(ap.edge_k_vecs(new_srp))
        current_range_amb = 2.0 * np.pi * ap.num_freqs / obb.range_span
        current_cross_range_amb = 2.0 * np.pi * ap.num_slow_times / obb.cross_range_span
        if cross_range_amb is not None:
            slow_downsample_factor = int(current_cross_range_amb
                                         / (2.0 * safety_factor * cross_range_amb))
        if slow_downsample_factor < 1:
            slow_downsample_factor = 1
        slow = int(slow_downsample_factor)
        if range_amb is not None:
            fast_downsample_factor = current_range_amb / (2.0 * safety_factor * range_amb)
        old_n = ap.num_freqs
        new_n = min(int(1.0 + old_n / fast_downsample_factor), old_n)
        kernel = get_downsample_kernel(slow)
        K = kernel.shape[0]
        new_p = ((ap.num_slow_times - K) // slow) - 1
        new_wave = fast_freq_downsample(np.expand_dims(ap.waveform_fft, axis=0), new_n,
                                        safety_factor)[0, :]
        bad = set(ap.bad_slow_times)
        starts = np.array([slow * i for i in range(max(new_p, 0))
                           if not bad.intersection(range(slow * i, slow * i + K))], dtype=np.int32)
        tx, rx, srps = ap.transmitter_positions, ap.receiver_positions, ap.stab_ref_posns
        dr = srp_range_change(srps, new_srp, tx, rx)
        if device is None:
            device = torch.device("cuda", torch.cuda.current_device())
        if hasattr(original_dataset, "phase_history_device"):
            # file-backed datasets decode their raw samples on the GPU
            ph = original_dataset.phase_history_device(slice(None), device)
        else:
            ph = original_dataset.phase_history()
        if isinstance(ph, torch.Tensor):
            ph_dev = ph.to(device=device, dtype=torch.complex64)
        else:
            ph_dev = torch.from_numpy(np.ascontiguousarray(ph, dtype=np.complex64)).to(device)
        window = downsample_window(new_n, safety_factor) / old_n
        out = _native.resample(ph_dev, dr, ap.centre_frequency, ap.sample_bandwidth,
                               ap.sign_multiplier, ap.c_eff, new_n, window, kernel, starts)
        centre = starts + K // 2
        w = np.asarray(ap.slow_time_weighting)
        weights = np.array([np.sum(kernel * w[s0:s0 + K]) for s0 in starts], dtype=np.complex128)
        return NumpySimpleSarDataset.from_scratch(
            trans_posns=tx[centre], rec_posns=rx[centre],
            srps=np.repeat(new_srp[None, :], starts.shape[0], axis=0),
            centre_freq=ap.centre_frequency, sample_freq=ap.sample_bandwidth,
            waveform_fft=new_wave, slow_time_weighting=weights,
            phase_history=out.cpu().numpy().astype(np.complex128),
            sign_multiplier=ap.sign_multiplier, c_eff=ap.c_eff)
