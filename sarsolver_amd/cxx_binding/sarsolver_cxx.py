"""Host-record entry points, bound to libsarsolver_hip.so.

Drop-in for the reference's ctypes module (src/sarsolver/cxx_binding/
sarsolver_cxx.py): the same ``SarCalculationInfo`` constructor and ``*_py``
properties, and ``forward_evaluate`` / ``adjoint_evaluate`` that overwrite
the record's phase history / scattering amplitudes in place.  The native side
runs the reference's nearest-neighbour rules with FP64 ranges on the GPU.
"""
from __future__ import annotations

import ctypes

import numpy as np
from numpy.typing import NDArray
from scipy.constants import c

from .._native import SarCalculationInfo as _Record
from .._native import evaluate_host


def _as_ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class SarCalculationInfo(_Record):
    """Marshals numpy inputs into the 128-byte C record (sarsolver_cxx.py:49-167).

    Real arrays become contiguous float64; complex arrays become [..., 2] float64
    (re, im).  Length mismatches raise ValueError, as in the reference."""

    def __init__(self, trans_posns: NDArray, rec_posns: NDArray, srps: NDArray,
                 scat_posns: NDArray, phase_history: NDArray, scat_amplitudes: NDArray,
                 waveform_fft: NDArray, slow_time_weighting: NDArray, centre_freq: float,
                 sample_freq: float, c_eff: float = c, upsample_ratio: float = 128.0,
                 sign_multiplier: float = 1.0):
        super().__init__()
        P, N = phase_history.shape[0], phase_history.shape[1]
        M = scat_posns.shape[0]
        self.num_fast_times, self.num_slow_times, self.num_scatterers = N, P, M
        self.centre_frequency = centre_freq
        self.sample_frequency = sample_freq
        self.c_eff = c_eff
        self.upsample_ratio = upsample_ratio
        self.sign_multiplier = sign_multiplier

        def real(a, n, what):
            arr = np.ascontiguousarray(np.asarray(a).astype("float64")).reshape(-1)
            if arr.shape[0] != n:
                raise ValueError(f"Check {what}.")
            return arr

        def cplx(a, shape, what):
            arr = np.zeros(tuple(shape) + (2,), dtype="float64")
            src = np.asarray(a)
            if src.size != int(np.prod(shape)):
                raise ValueError(f"Check {what}.")
            src = src.reshape(shape)
            arr[..., 0] = np.real(src).astype("float64")
            arr[..., 1] = np.imag(src).astype("float64")
            return arr

        self.trans_posns_py = real(trans_posns, 3 * P, "transmitter positions")
        self.rec_posns_py = real(rec_posns, 3 * P, "receiver positions")
        self.stab_ref_posns_py = real(srps, 3 * P, "stab ref positions")
        self.scat_posns_py = real(scat_posns, 3 * M, "scatterer positions")
        self._phase_history_py = cplx(phase_history, (P, N), "phase history")
        self._scattering_amplitudes_py = cplx(scat_amplitudes, (M,), "scattering amplitudes")
        self._waveform_fft_py = cplx(waveform_fft, (N,), "waveform")
        self._slow_time_weighting_py = cplx(slow_time_weighting, (P,), "slow time weighting")
        self.transmit_posns = _as_ptr(self.trans_posns_py)
        self.receive_posns = _as_ptr(self.rec_posns_py)
        self.stab_ref_posns = _as_ptr(self.stab_ref_posns_py)
        self.scat_posns = _as_ptr(self.scat_posns_py)
        self.phase_history = _as_ptr(self._phase_history_py)
        self.scattering_amplitudes = _as_ptr(self._scattering_amplitudes_py)
        self.waveform_fft = _as_ptr(self._waveform_fft_py)
        self.slow_time_weighting = _as_ptr(self._slow_time_weighting_py)

    @property
    def phase_history_py(self) -> NDArray:
        return self._phase_history_py[:, :, 0] + 1.0j * self._phase_history_py[:, :, 1]

    @property
    def scattering_amplitudes_py(self) -> NDArray:
        return self._scattering_amplitudes_py[:, 0] + 1.0j * self._scattering_amplitudes_py[:, 1]

    @property
    def waveform_fft_py(self) -> NDArray:
        return self._waveform_fft_py[:, 0] + 1.0j * self._waveform_fft_py[:, 1]

    @property
    def slow_time_weighting_py(self) -> NDArray:
        return self._slow_time_weighting_py[:, 0] + 1.0j * self._slow_time_weighting_py[:, 1]


def forward_evaluate(sar_info: SarCalculationInfo) -> None:
    """A x into ``sar_info.phase_history`` (FFT order); reference CPP:26-30."""
    evaluate_host(sar_info, forward=True)


def adjoint_evaluate(sar_info: SarCalculationInfo) -> None:
    """A^H y into ``sar_info.scattering_amplitudes``; reference CPP:32-36."""
    evaluate_host(sar_info, forward=False)
