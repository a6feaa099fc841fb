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
    "sarhip_plan_set_timing", "sarhip_plan_forward_fp64_rows", "sarhip_plan_gather_pairs",
    "sarhip_plan_timing", "sarhip_plan_kernel_timing", "sarhip_last_error", "sarhip_version",
    "sarhip_resample", "sarhip_plan_create_device_positions", "sarhip_plan_create_grid",
    "sarhip_grid_positions", "sarhip_baseband_values", "sarhip_decode_signal",
    "sarhip_conv3x3_bf16", "sarhip_conv3x3_wgrad_workspace", "sarhip_conv3x3_wgrad",
    "sarhip_conv_last_error", "sarhip_precond_net_workspace", "sarhip_precond_net",
    "sarhip_precond_net_noalloc", "sarhip_spectral_filter_workspace", "sarhip_spectral_filter",
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
        L.sarhip_plan_layout.argtypes = [vp, ip, ip, ip, ip, ip, dp, ip, ip, dp, ip]
        L.sarhip_plan_layout.restype = ctypes.c_int
        L.sarhip_plan_forward_fp64_rows.argtypes = [vp, dp, ip]
        L.sarhip_plan_forward_fp64_rows.restype = ctypes.c_int
        L.sarhip_plan_gather_pairs.argtypes = [vp, ip]
        L.sarhip_plan_gather_pairs.restype = ctypes.c_int
        L.sarhip_plan_set_timing.argtypes = [vp, ctypes.c_int]
        L.sarhip_plan_set_timing.restype = ctypes.c_int
        L.sarhip_plan_timing.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_float),
                                         ctypes.POINTER(ctypes.c_int)]
        L.sarhip_plan_timing.restype = ctypes.c_int
        L.sarhip_plan_kernel_timing.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_float),
                                                ctypes.POINTER(ctypes.c_int)]
        L.sarhip_plan_kernel_timing.restype = ctypes.c_int
        L.sarhip_last_error.argtypes = []
        L.sarhip_last_error.restype = ctypes.c_char_p
        L.sarhip_version.argtypes = []
        L.sarhip_version.restype = ctypes.c_char_p
        sz, dbl = ctypes.c_size_t, ctypes.c_double
        L.sarhip_resample.argtypes = [vp, sz, sz, dp, dbl, dbl, dbl, dbl, sz, dp, dp,
                                      ctypes.c_int, ip, sz, vp, vp]
        L.sarhip_resample.restype = ctypes.c_int
        L.sarhip_plan_create_device_positions.argtypes = [
            ctypes.POINTER(vp), sz, sz, sz, dp, dp, dp, vp, dp, dp, dbl, dbl, dbl, dbl, dbl,
            ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.sarhip_plan_create_device_positions.restype = ctypes.c_int
        szp = ctypes.POINTER(ctypes.c_size_t)
        L.sarhip_plan_create_grid.argtypes = [
            ctypes.POINTER(vp), sz, sz, szp, dp, dp, dp, dp, dp, dp, dp, dp, dbl, dbl, dbl, dbl,
            dbl, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.sarhip_plan_create_grid.restype = ctypes.c_int
        L.sarhip_grid_positions.argtypes = [szp, dp, dp, dp, vp, vp]
        L.sarhip_grid_positions.restype = ctypes.c_int
        L.sarhip_baseband_values.argtypes = [vp, sz, dp, dp, dbl, vp, vp]
        L.sarhip_baseband_values.restype = ctypes.c_int
        L.sarhip_decode_signal.argtypes = [vp, ctypes.c_int, sz, sz, vp, vp, vp]
        L.sarhip_decode_signal.restype = ctypes.c_int
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


class _GridArgs:
    """ctypes arrays describing a GridSarScene for the device generator."""

    def __init__(self, shape, dims, centre, rot):
        self.shape = (ctypes.c_size_t * 3)(*shape)
        self.dims = np.ascontiguousarray(dims, dtype=np.float64)
        self.centre = np.ascontiguousarray(centre, dtype=np.float64)
        self.rot = np.ascontiguousarray(rot, dtype=np.float64).reshape(9)
        self.num_points = int(shape[0]) * int(shape[1]) * int(shape[2])
        self.key = (tuple(int(v) for v in shape), self.dims.tobytes(), self.centre.tobytes(),
                    self.rot.tobytes())


def grid_args(scene) -> Optional[_GridArgs]:
    """The grid description of a GridSarScene-like object (grid_shape,
    grid_dimensions in (z, y, x) order, grid_centre, rotation_matrix) when
    the device generator reproduces its ``positions`` bit for bit, else None.

    numpy evaluates ``-0.5 * L`` in L's own type, so edge lengths must be
    Python numbers or FP64 (a float32 edge length gives a float32 linspace);
    centre and rotation only need to convert to FP64 exactly, which every
    real numpy dtype does."""
    need = ("grid_shape", "grid_dimensions", "grid_centre", "rotation_matrix")
    if not all(hasattr(scene, a) for a in need):
        return None
    shape = [int(v) for v in scene.grid_shape]
    dims = list(scene.grid_dimensions)
    if len(shape) != 3 or len(dims) != 3 or min(shape) < 1:
        return None
    for d in dims:
        if not (isinstance(d, (int, float, np.integer)) or
                (isinstance(d, np.floating) and np.asarray(d).dtype == np.float64)):
            return None
        if isinstance(d, (int, np.integer)) and abs(int(d)) > 2 ** 53:
            return None
    centre = np.asarray(scene.grid_centre)
    rot = np.asarray(scene.rotation_matrix)
    if centre.shape != (3,) or rot.shape != (3, 3):
        return None
    if centre.dtype.kind not in "fiu" or rot.dtype.kind not in "fiu":
        return None
    return _GridArgs(shape, [float(d) for d in dims], centre, rot)


class _GridScene:
    """Minimal GridSarScene stand-in rebuilt from stored grid arguments."""

    def __init__(self, g: _GridArgs):
        self.grid_shape = tuple(g.shape)
        self.grid_dimensions = [float(v) for v in g.dims]
        self.grid_centre = g.centre
        self.rotation_matrix = g.rot.reshape(3, 3)


def src_scene_of(g: _GridArgs) -> _GridScene:
    return _GridScene(g)


def grid_positions(scene, device=None):
    """GridSarScene positions generated on the device: [M, 3] float64 torch
    tensor, bit-identical to ``scene.positions``."""
    torch = _torch()
    g = grid_args(scene)
    if g is None:
        raise ValueError("scene is not a grid the device generator reproduces exactly")
    dev = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
    out = torch.empty((g.num_points, 3), dtype=torch.float64, device=dev)
    with torch.cuda.device(dev):
        rc = load_library().sarhip_grid_positions(g.shape, _dptr(g.dims), _dptr(g.centre),
                                                   _dptr(g.rot), ctypes.c_void_p(out.data_ptr()),
                                                   _stream_handle(dev))
    check(rc, "sarhip_grid_positions")
    return out


def baseband_values(positions, spot, displacement, sign_multiplier: float):
    """exp(-i sign ((positions - spot) . displacement)) on the device:
    positions [M, 3] float64 CUDA tensor -> [M] complex128 tensor."""
    torch = _torch()
    if not (isinstance(positions, torch.Tensor) and positions.is_cuda and
            positions.dtype == torch.float64 and positions.dim() == 2 and positions.shape[1] == 3):
        raise TypeError("positions must be an [M, 3] float64 tensor on a GPU")
    pos = positions.contiguous()
    sp = _f64(spot).reshape(3)
    dp_ = _f64(displacement).reshape(3)
    out = torch.empty((pos.shape[0],), dtype=torch.complex128, device=pos.device)
    with torch.cuda.device(pos.device):
        rc = load_library().sarhip_baseband_values(ctypes.c_void_p(pos.data_ptr()), pos.shape[0],
                                                    _dptr(sp), _dptr(dp_), float(sign_multiplier),
                                                    ctypes.c_void_p(out.data_ptr()),
                                                    _stream_handle(pos.device))
    check(rc, "sarhip_baseband_values")
    return out


def decode_signal(raw: np.ndarray, format_code: int, num_samples: int, amp_sf=None, device=None):
    """Raw big-endian CPHD samples (uint8 [n, num_samples * bytes per sample],
    host) decoded on the GPU into an [n, num_samples] complex64 tensor;
    amp_sf [n] scales each vector when given."""
    torch = _torch()
    dev = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
    raw = np.ascontiguousarray(raw, dtype=np.uint8)
    n = raw.shape[0] if raw.ndim == 2 else 0
    out = torch.empty((n, int(num_samples)), dtype=torch.complex64, device=dev)
    if n == 0:
        return out
    d_raw = torch.from_numpy(raw.reshape(-1)).to(dev)
    d_sf = None if amp_sf is None else torch.from_numpy(
        np.ascontiguousarray(amp_sf, dtype=np.float32)).to(dev)
    with torch.cuda.device(dev):
        rc = load_library().sarhip_decode_signal(
            ctypes.c_void_p(d_raw.data_ptr()), int(format_code), n, int(num_samples),
            ctypes.c_void_p(d_sf.data_ptr()) if d_sf is not None else None,
            ctypes.c_void_p(out.data_ptr()), _stream_handle(dev))
    check(rc, "sarhip_decode_signal")
    return out


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
        """``scat_posns`` is an [M, 3] host array, an [M, 3] float64 CUDA tensor
        on the plan's device (read once; the plan keeps its own copy), or a
        GridSarScene, whose positions are then generated on the device."""
        torch = _torch()
        L = load_library()
        if interp not in INTERP:
            raise ValueError(f"interp must be one of {sorted(INTERP)}")
        if precision not in PRECISION:
            raise ValueError(f"precision must be one of {sorted(PRECISION)}")
        tx = _f64(transmit_posns).reshape(-1)
        rx = _f64(receive_posns).reshape(-1)
        srp = _f64(stab_ref_posns).reshape(-1)
        W = _c128_interleaved(waveform_fft)
        w = _c128_interleaved(slow_time_weighting)
        if device is None:
            device = torch.device("cuda", torch.cuda.current_device())
        self.device = torch.device(device)
        grid = grid_args(scat_posns)
        if grid is not None:
            kind, M = "grid", grid.num_points
            self._pos_source = grid
        elif isinstance(scat_posns, torch.Tensor):
            t = scat_posns
            if not t.is_cuda or t.device != self.device:
                raise ValueError(f"device positions are on {t.device}, plan is on {self.device}")
            if t.dtype != torch.float64 or t.dim() != 2 or t.shape[1] != 3:
                raise ValueError("device positions must be an [M, 3] float64 tensor")
            kind, M = "device", int(t.shape[0])
            t = t.contiguous()
            self._pos_source = t
        else:
            pos = _f64(scat_posns).reshape(-1)
            if pos.shape[0] % 3:
                raise ValueError("scatterer positions must be a non-empty [M, 3] array")
            kind, M = "host", pos.shape[0] // 3
            self._pos_source = pos.reshape(-1, 3)
        self.position_source = kind
        self._geometry = {"tx": tx.reshape(-1, 3), "rx": rx.reshape(-1, 3),
                          "srp": srp.reshape(-1, 3)}
        self.num_slow_times = tx.shape[0] // 3
        self.num_fast_times = W.shape[0] // 2
        self.num_scatterers = M
        P, N = self.num_slow_times, self.num_fast_times
        if tx.shape[0] != 3 * P or rx.shape[0] != 3 * P or srp.shape[0] != 3 * P:
            raise ValueError("transmitter, receiver and SRP arrays must all be [P, 3]")
        if M == 0:
            raise ValueError("scatterer positions must be a non-empty [M, 3] array")
        if w.shape[0] != 2 * P:
            raise ValueError("slow_time_weighting must have one entry per pulse")
        self.interp, self.precision = interp, precision
        handle = ctypes.c_void_p()
        common = (_dptr(W), _dptr(w), float(centre_frequency), float(sample_frequency),
                  float(c_eff), float(upsample_ratio), float(sign_multiplier), INTERP[interp],
                  PRECISION[precision], self.device.index or 0)
        if kind == "grid":
            rc = L.sarhip_plan_create_grid(ctypes.byref(handle), N, P, grid.shape,
                                           _dptr(grid.dims), _dptr(grid.centre), _dptr(grid.rot),
                                           _dptr(tx), _dptr(rx), _dptr(srp), *common)
            check(rc, "sarhip_plan_create_grid")
        elif kind == "device":
            # the library reads the positions on the default stream
            torch.cuda.current_stream(self.device).synchronize()
            rc = L.sarhip_plan_create_device_positions(
                ctypes.byref(handle), N, P, M, _dptr(tx), _dptr(rx), _dptr(srp),
                ctypes.c_void_p(t.data_ptr()), *common)
            check(rc, "sarhip_plan_create_device_positions")
        else:
            rc = L.sarhip_plan_create(ctypes.byref(handle), N, P, M, _dptr(tx), _dptr(rx),
                                      _dptr(srp), _dptr(pos), *common)
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
        bp_ord = ctypes.c_int()
        bp_err = ctypes.c_double()
        fused = ctypes.c_int()
        check(L.sarhip_plan_layout(handle, *[ctypes.byref(v) for v in vals], ctypes.byref(err),
                                   ctypes.byref(reg), ctypes.byref(bp_ord), ctypes.byref(bp_err),
                                   ctypes.byref(fused)),
              "sarhip_plan_layout")
        inv = {v: k for k, v in PRECISION.items()}
        self.layout = {"precision": inv[vals[0].value], "monostatic": bool(vals[1].value),
                       "valid_pulses": vals[2].value, "fwd_window_bins": vals[3].value,
                       "fwd_pulses_per_group": vals[4].value, "series_err_bins": err.value,
                       "fwd_region_tiles": reg.value, "bp_series_order": bp_ord.value,
                       "bp_series_err_bins": bp_err.value, "fused_fft_ratio": fused.value}
        kappa = ctypes.c_double()
        check(L.sarhip_plan_forward_fp64_rows(handle, ctypes.byref(kappa), None),
              "sarhip_plan_forward_fp64_rows")
        self.layout["fwd_fft_kappa"] = kappa.value
        pairs = ctypes.c_int()
        check(L.sarhip_plan_gather_pairs(handle, ctypes.byref(pairs)), "sarhip_plan_gather_pairs")
        self.layout["bp_pairs"] = bool(pairs.value)
        self.effective_precision = self.layout["precision"]

    @property
    def geometry(self):
        """Host copies of the geometry: tx, rx, srp [P, 3] and pos [M, 3]
        (pos is brought to the host on first use for device and grid plans)."""
        if "pos" not in self._geometry:
            src = self._pos_source
            if self.position_source == "host":
                pos = src
            elif self.position_source == "device":
                pos = src.cpu().numpy()
            else:
                pos = grid_positions(src_scene_of(src), self.device).cpu().numpy()
            self._geometry["pos"] = pos
        return self._geometry

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

    def forward_fp64_rows(self) -> int:
        """Rows of the most recent forward whose spectral stage was redone in
        FP64 (cancellation factor above ``layout['fwd_fft_kappa']``; 0 when
        that stage is not adaptive).  Waits for the device."""
        rows = ctypes.c_int()
        check(load_library().sarhip_plan_forward_fp64_rows(self._handle, None, ctypes.byref(rows)),
              "sarhip_plan_forward_fp64_rows")
        return rows.value

    def timing(self, which: str):
        """Mean device ms per logged call of ``which`` ('forward' or 'adjoint'):
        dict with window, fft, kernel (gather/scatter), total and calls."""
        buf = (ctypes.c_float * 4)()
        n = ctypes.c_int()
        check(load_library().sarhip_plan_timing(self._handle, {"forward": 0, "adjoint": 1}[which],
                                                buf, ctypes.byref(n)), "sarhip_plan_timing")
        ms = ctypes.c_float()
        check(load_library().sarhip_plan_kernel_timing(
            self._handle, {"forward": 0, "adjoint": 1}[which], ctypes.byref(ms), None),
            "sarhip_plan_kernel_timing")
        return {"window": buf[0], "fft": buf[1], "kernel": buf[2], "total": buf[3],
                "main_kernel": ms.value, "calls": n.value}


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
