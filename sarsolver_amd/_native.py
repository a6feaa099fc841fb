"""ctypes binding of libsarsolver_hip.so (the gfx950 SAR operator library).

The library is loaded from the package directory (or ``$SARSOLVER_HIP_LIB``).
There is no CPU fallback: if the library or a gfx950 device is missing, every
entry point raises.  PyTorch-ROCm tensors are used only as device memory and
stream carriers; the C ABI takes raw pointers.
"""
from __future__ import annotations

import ctypes
import os
import threading
from pathlib import Path
from typing import Optional

import numpy as np

PKG = Path(__file__).resolve().parent
LIB_NAME = "libsarsolver_hip.so"

INTERP = {"nearest_ref": 0, "nearest": 1, "linear": 2}
PRECISION = {"fp64": 0, "series": 1, "hilo": 2, "fp32": 3}
ORDER = {"fft": 0, "centred": 1}

_lock = threading.Lock()
_lib: Optional[ctypes.CDLL] = None


class SarCalculationInfo(ctypes.Structure):
    """The 128-byte record of the reference's C ABI (sarsolver_cxx.hpp:27-34)."""

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


# Every symbol declared in include/sarsolver_hip.h.
EXPORTED_SYMBOLS = (
    "forward_evaluate", "adjoint_evaluate", "sarhip_plan_create", "sarhip_plan_destroy",
    "sarhip_forward", "sarhip_adjoint", "sarhip_plan_info", "sarhip_plan_layout",
    "sarhip_plan_set_timing",
    "sarhip_plan_timing", "sarhip_last_error", "sarhip_version",
    "sarhip_resample",
    "sarhip_conv3x3_bf16", "sarhip_conv3x3_wgrad_workspace", "sarhip_conv3x3_wgrad",
    "sarhip_conv_last_error",
)


def library_path() -> Path:
    env = os.environ.get("SARSOLVER_HIP_LIB")
    return Path(env) if env else PKG / LIB_NAME


def load_library() -> ctypes.CDLL:
    """Load and bind the library; raises if it has not been built."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        path = library_path()
        if not path.exists():
            raise RuntimeError(f"{path} not found: build it with `python -m sarsolver_amd.build` "
                               "(hipcc, gfx950). There is no CPU fallback.")
        L = ctypes.CDLL(str(path))
        dp = ctypes.POINTER(ctypes.c_double)
        vp = ctypes.c_void_p
        L.forward_evaluate.argtypes = [ctypes.POINTER(SarCalculationInfo)]
        L.forward_evaluate.restype = ctypes.c_int
        L.adjoint_evaluate.argtypes = [ctypes.POINTER(SarCalculationInfo)]
        L.adjoint_evaluate.restype = ctypes.c_int
        L.sarhip_plan_create.argtypes = [ctypes.POINTER(vp), ctypes.c_size_t, ctypes.c_size_t,
                                         ctypes.c_size_t, dp, dp, dp, dp, dp, dp,
                                         ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                         ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                         ctypes.c_int, ctypes.c_int]
        L.sarhip_plan_create.restype = ctypes.c_int
        L.sarhip_plan_destroy.argtypes = [vp]
        L.sarhip_plan_destroy.restype = ctypes.c_int
        L.sarhip_forward.argtypes = [vp, vp, vp, ctypes.c_int, vp]
        L.sarhip_forward.restype = ctypes.c_int
        L.sarhip_adjoint.argtypes = [vp, vp, vp, ctypes.c_int, vp]
        L.sarhip_adjoint.restype = ctypes.c_int
        L.sarhip_plan_info.argtypes = [vp, ctypes.POINTER(ctypes.c_size_t), dp, dp]
        L.sarhip_plan_info.restype = ctypes.c_int
        ip = ctypes.POINTER(ctypes.c_int)
        L.sarhip_plan_layout.argtypes = [vp, ip, ip, ip, ip, ip, dp, ip]
        L.sarhip_plan_layout.restype = ctypes.c_int
        L.sarhip_plan_set_timing.argtypes = [vp, ctypes.c_int]
        L.sarhip_plan_set_timing.restype = ctypes.c_int
        L.sarhip_plan_timing.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_float),
                                         ctypes.POINTER(ctypes.c_int)]
        L.sarhip_plan_timing.restype = ctypes.c_int
        L.sarhip_last_error.argtypes = []
        L.sarhip_last_error.restype = ctypes.c_char_p
        L.sarhip_version.argtypes = []
        L.sarhip_version.restype = ctypes.c_char_p
        sz, dbl = ctypes.c_size_t, ctypes.c_double
        L.sarhip_resample.argtypes = [vp, sz, sz, dp, dbl, dbl, dbl, dbl, sz, dp, dp,
                                      ctypes.c_int, ip, sz, vp, vp]
        L.sarhip_resample.restype = ctypes.c_int
        _lib = L
        return L


def check(rc: int, what: str) -> None:
    if rc != 0:
        msg = load_library().sarhip_last_error().decode(errors="replace")
        exc = ValueError if rc == -1 else RuntimeError
        raise exc(f"{what} failed ({rc}): {msg}")


def _f64(a, shape=None) -> np.ndarray:
    out = np.ascontiguousarray(np.asarray(a, dtype=np.float64))
    if shape is not None:
        out = out.reshape(shape)
    return out


def _c128_interleaved(a) -> np.ndarray:
    z = np.ascontiguousarray(np.asarray(a, dtype=np.complex128)).reshape(-1)
    return z.view(np.float64)


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _torch():
    import torch
    return torch


def _stream_handle(device) -> ctypes.c_void_p:
    torch = _torch()
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


class SarPlan:
    """A device plan: geometry uploaded once, operators run on device tensors.

    Parameters follow the reference's SarCalculationInfo (sarsolver_cxx.py:64-67):
    positions in metres (FP64), ``waveform_fft`` in FFT order, ``slow_time_weighting``
    per pulse.  ``interp`` is one of ``nearest_ref`` (reference bin rules),
    ``nearest`` or ``linear``; ``precision`` selects the range arithmetic.
    """

    def __init__(self, transmit_posns, receive_posns, stab_ref_posns, scat_posns, waveform_fft,
                 slow_time_weighting, centre_frequency: float, sample_frequency: float,
                 c_eff: float, upsample_ratio: float = 128.0, sign_multiplier: float = 1.0,
                 interp: str = "nearest_ref", precision: str = "fp64", device=None):
        torch = _torch()
        L = load_library()
        if interp not in INTERP:
            raise ValueError(f"interp must be one of {sorted(INTERP)}")
        if precision not in PRECISION:
            raise ValueError(f"precision must be one of {sorted(PRECISION)}")
        tx = _f64(transmit_posns).reshape(-1)
        rx = _f64(receive_posns).reshape(-1)
        srp = _f64(stab_ref_posns).reshape(-1)
        pos = _f64(scat_posns).reshape(-1)
        W = _c128_interleaved(waveform_fft)
        w = _c128_interleaved(slow_time_weighting)
        self.geometry = {"tx": tx.reshape(-1, 3), "rx": rx.reshape(-1, 3),
                         "srp": srp.reshape(-1, 3), "pos": pos.reshape(-1, 3)}
        self.num_slow_times = tx.shape[0] // 3
        self.num_fast_times = W.shape[0] // 2
        self.num_scatterers = pos.shape[0] // 3
        P, N, M = self.num_slow_times, self.num_fast_times, self.num_scatterers
        if tx.shape[0] != 3 * P or rx.shape[0] != 3 * P or srp.shape[0] != 3 * P:
            raise ValueError("transmitter, receiver and SRP arrays must all be [P, 3]")
        if pos.shape[0] != 3 * M or M == 0:
            raise ValueError("scatterer positions must be a non-empty [M, 3] array")
        if w.shape[0] != 2 * P:
            raise ValueError("slow_time_weighting must have one entry per pulse")
        if device is None:
            device = torch.device("cuda", torch.cuda.current_device())
        self.device = torch.device(device)
        self.interp, self.precision = interp, precision
        handle = ctypes.c_void_p()
        rc = L.sarhip_plan_create(ctypes.byref(handle), N, P, M, _dptr(tx), _dptr(rx), _dptr(srp),
                                  _dptr(pos), _dptr(W), _dptr(w), float(centre_frequency),
                                  float(sample_frequency), float(c_eff), float(upsample_ratio),
                                  float(sign_multiplier), INTERP[interp], PRECISION[precision],
                                  self.device.index or 0)
        check(rc, "sarhip_plan_create")
        self._handle = handle
        nw = ctypes.c_size_t()
        by_dx = ctypes.c_double()
        kc = ctypes.c_double()
        check(L.sarhip_plan_info(handle, ctypes.byref(nw), ctypes.byref(by_dx), ctypes.byref(kc)),
              "sarhip_plan_info")
        self.working_num_fast_times = nw.value
        self.by_dx = by_dx.value
        self.centre_wavenumber = kc.value
        vals = [ctypes.c_int() for _ in range(5)]
        err = ctypes.c_double()
        reg = ctypes.c_int()
        check(L.sarhip_plan_layout(handle, *[ctypes.byref(v) for v in vals], ctypes.byref(err),
                                   ctypes.byref(reg)), "sarhip_plan_layout")
        inv = {v: k for k, v in PRECISION.items()}
        self.layout = {"precision": inv[vals[0].value], "monostatic": bool(vals[1].value),
                       "valid_pulses": vals[2].value, "fwd_window_bins": vals[3].value,
                       "fwd_pulses_per_group": vals[4].value, "series_err_bins": err.value,
                       "fwd_region_tiles": reg.value}
        self.effective_precision = self.layout["precision"]

    def close(self) -> None:
        if getattr(self, "_handle", None):
            load_library().sarhip_plan_destroy(self._handle)
            self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check_tensor(self, t, shape, name):
        torch = _torch()
        if not isinstance(t, torch.Tensor):
            raise TypeError(f"{name} must be a torch.Tensor")
        if t.dtype != torch.complex64:
            raise TypeError(f"{name} must be complex64, got {t.dtype}")
        if t.device != self.device:
            raise ValueError(f"{name} is on {t.device}, plan is on {self.device}")
        if tuple(t.shape) != tuple(shape):
            raise ValueError(f"{name} has shape {tuple(t.shape)}, expected {tuple(shape)}")
        if not t.is_contiguous():
            raise ValueError(f"{name} must be contiguous")

    def forward(self, amps, out=None, order: str = "fft"):
        """A x on device: amps [M] complex64 -> phase history [P, N] complex64."""
        torch = _torch()
        M, P, N = self.num_scatterers, self.num_slow_times, self.num_fast_times
        amps = amps.reshape(M)
        self._check_tensor(amps, (M,), "amps")
        if out is None:
            out = torch.empty((P, N), dtype=torch.complex64, device=self.device)
        self._check_tensor(out, (P, N), "out")
        rc = load_library().sarhip_forward(self._handle, ctypes.c_void_p(amps.data_ptr()),
                                           ctypes.c_void_p(out.data_ptr()), ORDER[order],
                                           _stream_handle(self.device))
        check(rc, "sarhip_forward")
        return out

    def adjoint(self, phase_history, out=None, order: str = "fft"):
        """A^H y on device: phase history [P, N] complex64 -> image [M] complex64."""
        torch = _torch()
        M, P, N = self.num_scatterers, self.num_slow_times, self.num_fast_times
        self._check_tensor(phase_history, (P, N), "phase_history")
        if out is None:
            out = torch.empty((M,), dtype=torch.complex64, device=self.device)
        self._check_tensor(out, (M,), "out")
        rc = load_library().sarhip_adjoint(self._handle, ctypes.c_void_p(phase_history.data_ptr()),
                                           ctypes.c_void_p(out.data_ptr()), ORDER[order],
                                           _stream_handle(self.device))
        check(rc, "sarhip_adjoint")
        return out

    def set_timing(self, enable: bool) -> None:
        """Start (and clear) or stop the per-call stage event log."""
        check(load_library().sarhip_plan_set_timing(self._handle, int(bool(enable))),
              "sarhip_plan_set_timing")

    def timing(self, which: str):
        """Mean device ms per logged call of ``which`` ('forward' or 'adjoint'):
        dict with window, fft, kernel (gather/scatter), total and calls."""
        buf = (ctypes.c_float * 4)()
        n = ctypes.c_int()
        check(load_library().sarhip_plan_timing(self._handle, {"forward": 0, "adjoint": 1}[which],
                                                buf, ctypes.byref(n)), "sarhip_plan_timing")
        return {"window": buf[0], "fft": buf[1], "kernel": buf[2], "total": buf[3],
                "calls": n.value}


def resample(phase_history, dr, centre_freq: float, sample_freq: float, sign_multiplier: float,
             c_eff: float, new_num_freqs: int, window, kernel, starts):
    """Device resampling (`sarhip_resample`): phase_history [P, N] complex64
    torch tensor (centred order) on a GPU -> [len(starts), new_num_freqs]
    complex64 tensor on the same device.  dr [P], window [new_num_freqs] (FFT
    order, including 1/N), kernel [K] complex and starts [P2] are host arrays."""
    torch = _torch()
    L = load_library()
    if phase_history.dtype != torch.complex64 or not phase_history.is_cuda:
        raise TypeError("phase_history must be a complex64 tensor on a GPU")
    ph = phase_history.contiguous()
    P, N = ph.shape
    drh = _f64(dr).reshape(-1)
    win = _f64(window).reshape(-1)
    kern = _c128_interleaved(kernel)
    st = np.ascontiguousarray(np.asarray(starts, dtype=np.int32).reshape(-1))
    if drh.shape[0] != P or win.shape[0] != new_num_freqs:
        raise ValueError("dr must have one entry per pulse and window one per new frequency")
    out = torch.empty((st.shape[0], new_num_freqs), dtype=torch.complex64, device=ph.device)
    with torch.cuda.device(ph.device):
        rc = L.sarhip_resample(ctypes.c_void_p(ph.data_ptr()), P, N, _dptr(drh),
                               float(centre_freq), float(sample_freq), float(sign_multiplier),
                               float(c_eff), int(new_num_freqs), _dptr(win), _dptr(kern),
                               kern.shape[0] // 2,
                               st.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), st.shape[0],
                               ctypes.c_void_p(out.data_ptr()), _stream_handle(ph.device))
    check(rc, "sarhip_resample")
    return out


def evaluate_host(info: SarCalculationInfo, forward: bool) -> None:
    """Call the reference-compatible entry point on a host record."""
    L = load_library()
    fn = L.forward_evaluate if forward else L.adjoint_evaluate
    check(fn(ctypes.byref(info)), "forward_evaluate" if forward else "adjoint_evaluate")
