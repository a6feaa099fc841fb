"""ctypes wrapper around the FP64 CPU oracle (liboracle: oracle/sar_oracle.c).

TEST INFRASTRUCTURE ONLY: the checker for the HIP library.  Only the tests,
the smoke check and the CPU-baseline leg of the benchmark may use it.

The functions take and return numpy arrays in the reference's conventions:
positions [.,3] float64, complex128 vectors, phase history [P, N] in FFT order
(the order the reference hands its C++ core, src/sarsolver/cxx_binding/
forward.py:39,44 and adjoint.py:36-38).
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libsar_oracle.so"

NEAREST_REF, NEAREST, LINEAR = 0, 1, 2
INTERP = {"nearest_ref": NEAREST_REF, "nearest": NEAREST, "linear": LINEAR}

C_LIGHT = 299792458.0


class SarCalculationInfo(ctypes.Structure):
    """Same 128-byte layout as the C record in include/sarsolver_hip.h."""

    _fields_ = [("num_fast_times", ctypes.c_size_t), ("num_slow_times", ctypes.c_size_t),
                ("num_scatterers", ctypes.c_size_t),
                ("transmit_posns", ctypes.POINTER(ctypes.c_double)),
                ("receive_posns", ctypes.POINTER(ctypes.c_double)),
                ("stab_ref_posns", ctypes.POINTER(ctypes.c_double)),
                ("scat_posns", ctypes.POINTER(ctypes.c_double)),
                ("phase_history", ctypes.POINTER(ctypes.c_double)),
                ("scattering_amplitudes", ctypes.POINTER(ctypes.c_double)),
                ("waveform_fft", ctypes.POINTER(ctypes.c_double)),
                ("slow_time_weighting", ctypes.POINTER(ctypes.c_double)),
                ("centre_frequency", ctypes.c_double), ("sample_frequency", ctypes.c_double),
                ("c_eff", ctypes.c_double), ("upsample_ratio", ctypes.c_double),
                ("sign_multiplier", ctypes.c_double)]


def build(force: bool = False) -> Path:
    """Compile liboracle with the committed Makefile (gcc, -ffp-contract=off)."""
    if force or not LIB_PATH.exists():
        subprocess.run(["make", "-s", "-B" if force else "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = ctypes.CDLL(str(LIB_PATH))
        P = ctypes.POINTER(SarCalculationInfo)
        L.oracle_forward.argtypes = [P, ctypes.c_int, ctypes.c_int]
        L.oracle_adjoint.argtypes = [P, ctypes.c_int, ctypes.c_int]
        L.oracle_fft.argtypes = [ctypes.POINTER(ctypes.c_double), ctypes.c_size_t, ctypes.c_int]
        L.oracle_differential_ranges.argtypes = [P, ctypes.POINTER(ctypes.c_double)]
        L.oracle_bin.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_size_t, ctypes.c_int,
                                 ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
        L.oracle_bin.restype = ctypes.c_long
        L.oracle_working_len.argtypes = [ctypes.c_size_t, ctypes.c_double]
        L.oracle_working_len.restype = ctypes.c_size_t
        _lib = L
    return _lib


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _c128_view(a) -> np.ndarray:
    a = np.ascontiguousarray(np.asarray(a, dtype=np.complex128))
    return a


class _Info:
    """Keeps the marshalled arrays alive for the lifetime of the record."""

    def __init__(self, tx, rx, srp, pos, ph, amps, W_fft, w, fc, fs, c_eff, U, sgn):
        self.tx = np.ascontiguousarray(tx, dtype=np.float64).reshape(-1, 3)
        self.rx = np.ascontiguousarray(rx, dtype=np.float64).reshape(-1, 3)
        self.srp = np.ascontiguousarray(srp, dtype=np.float64).reshape(-1, 3)
        self.pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
        self.ph = _c128_view(ph)
        self.amps = _c128_view(amps).reshape(-1)
        self.W = _c128_view(W_fft).reshape(-1)
        self.w = _c128_view(w).reshape(-1)
        P, N = self.ph.shape
        M = self.pos.shape[0]
        if self.tx.shape[0] != P or self.rx.shape[0] != P or self.srp.shape[0] != P:
            raise ValueError("pulse count mismatch")
        if self.amps.shape[0] != M or self.W.shape[0] != N or self.w.shape[0] != P:
            raise ValueError("size mismatch")
        self.rec = SarCalculationInfo(N, P, M, _dptr(self.tx), _dptr(self.rx), _dptr(self.srp),
                                      _dptr(self.pos), _dptr(self.ph.view(np.float64)),
                                      _dptr(self.amps.view(np.float64)),
                                      _dptr(self.W.view(np.float64)),
                                      _dptr(self.w.view(np.float64)),
                                      float(fc), float(fs), float(c_eff), float(U), float(sgn))


def forward(tx, rx, srp, pos, amps, W_fft, w, fc, fs, num_fast_times, *, c_eff=C_LIGHT,
            upsample_ratio=128.0, sign_multiplier=1.0, interp="nearest_ref", num_threads=1):
    """A x: phase history [P, N] complex128 in FFT order (CPP:264-310)."""
    P = np.asarray(tx).reshape(-1, 3).shape[0]
    ph = np.zeros((P, num_fast_times), dtype=np.complex128)
    info = _Info(tx, rx, srp, pos, ph, amps, W_fft, w, fc, fs, c_eff, upsample_ratio,
                 sign_multiplier)
    rc = lib().oracle_forward(ctypes.byref(info.rec), INTERP[interp], int(num_threads))
    if rc:
        raise RuntimeError(f"oracle_forward failed ({rc})")
    return info.ph


def adjoint(tx, rx, srp, pos, ph_fft, W_fft, w, fc, fs, *, c_eff=C_LIGHT, upsample_ratio=128.0,
            sign_multiplier=1.0, interp="nearest_ref", num_threads=1):
    """A^H y: image [M] complex128 from a phase history in FFT order (CPP:312-353)."""
    pos = np.asarray(pos).reshape(-1, 3)
    amps = np.zeros(pos.shape[0], dtype=np.complex128)
    info = _Info(tx, rx, srp, pos, ph_fft, amps, W_fft, w, fc, fs, c_eff, upsample_ratio,
                 sign_multiplier)
    rc = lib().oracle_adjoint(ctypes.byref(info.rec), INTERP[interp], int(num_threads))
    if rc:
        raise RuntimeError(f"oracle_adjoint failed ({rc})")
    return info.amps


def fft(x, sign=-1) -> np.ndarray:
    """Unnormalised DFT with kernel e^{sign 2 pi i jk/n} (sign=-1 is numpy.fft.fft)."""
    a = np.array(x, dtype=np.complex128, copy=True).reshape(-1)
    rc = lib().oracle_fft(_dptr(a.view(np.float64)), a.shape[0], int(sign))
    if rc:
        raise RuntimeError(f"oracle_fft failed ({rc})")
    return a


def differential_ranges(tx, rx, srp, pos) -> np.ndarray:
    """dR[p, m] = bistatic range(pixel m) - bistatic range(SRP_p) (CPP:283,287)."""
    tx = np.asarray(tx, dtype=np.float64).reshape(-1, 3)
    pos = np.asarray(pos, dtype=np.float64).reshape(-1, 3)
    P, M = tx.shape[0], pos.shape[0]
    info = _Info(tx, rx, srp, pos, np.zeros((P, 1), np.complex128), np.zeros(M, np.complex128),
                 np.ones(1, np.complex128), np.ones(P, np.complex128), 1.0, 1.0, 1.0, 1.0, 1.0)
    out = np.zeros((P, M), dtype=np.float64)
    lib().oracle_differential_ranges(ctypes.byref(info.rec), _dptr(out))
    return out


def bin_index(dr: float, by_dx: float, nw: int, interp="nearest_ref", adjoint=False):
    f = ctypes.c_double(0.0)
    n = lib().oracle_bin(float(dr), float(by_dx), int(nw), INTERP[interp], int(adjoint),
                         ctypes.byref(f))
    return (n, f.value) if interp == "linear" else n


def working_len(n: int, upsample_ratio: float) -> int:
    return int(lib().oracle_working_len(int(n), float(upsample_ratio)))
