"""The C-ABI library and the host-side mirror of the reference API (CPU only)."""
import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

from conftest import ROOT


def header_functions():
    text = (ROOT / "include" / "sarsolver_hip.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|size_t|const char \*)\s*\**\s*(\w+)\s*\(", text, re.M)))


def test_header_declares_the_reference_abi():
    names = header_functions()
    assert "forward_evaluate" in names and "adjoint_evaluate" in names
    assert len(names) >= 10


def test_library_exports_every_declared_symbol():
    from sarsolver_amd import _native
    L = _native.load_library()
    for name in header_functions():
        assert hasattr(L, name), name
    assert set(header_functions()) == set(_native.EXPORTED_SYMBOLS)
    assert L.sarhip_version().startswith(b"sarsolver_hip")


def test_record_layout_is_128_bytes():
    from sarsolver_amd._native import SarCalculationInfo
    assert ctypes.sizeof(SarCalculationInfo) == 128
    offs = [getattr(SarCalculationInfo, f).offset for f, _ in SarCalculationInfo._fields_]
    assert offs == list(range(0, 128, 8))


def test_plan_create_rejects_bad_arguments_without_a_gpu():
    """Argument errors are reported as status codes, never crashes."""
    from sarsolver_amd import _native
    L = _native.load_library()
    h = ctypes.c_void_p()
    d = (ctypes.c_double * 6)()
    dp = ctypes.cast(d, ctypes.POINTER(ctypes.c_double))
    rc = L.sarhip_plan_create(ctypes.byref(h), 0, 1, 1, dp, dp, dp, dp, dp, dp, 1.0, 1.0, 1.0,
                              1.0, 1.0, 0, 0, 0)
    assert rc == -1 and b"positive" in L.sarhip_last_error()
    rc = L.sarhip_plan_create(ctypes.byref(h), 4, 1, 1, dp, dp, dp, dp, dp, dp, 1.0, 1.0, 1.0,
                              0.5, 1.0, 0, 0, 0)
    assert rc == -1
    rc = L.sarhip_plan_create(ctypes.byref(h), 4, 1, 1, dp, dp, dp, dp, dp, dp, 1.0, 1.0, 1.0,
                              2.0, 1.0, 7, 0, 0)
    assert rc == -1
    assert L.sarhip_forward(None, None, None, 0, None) == -1


def test_host_record_shape_checks():
    from sarsolver_amd.cxx_binding import SarCalculationInfo
    P, N, M = 3, 4, 5
    ok = dict(trans_posns=np.zeros((P, 3)), rec_posns=np.zeros((P, 3)), srps=np.zeros((P, 3)),
              scat_posns=np.zeros((M, 3)), phase_history=np.zeros((P, N), complex),
              scat_amplitudes=np.zeros(M, complex), waveform_fft=np.ones(N, complex),
              slow_time_weighting=np.ones(P, complex), centre_freq=1e9, sample_freq=1e8)
    info = SarCalculationInfo(**ok)
    assert (info.num_fast_times, info.num_slow_times, info.num_scatterers) == (N, P, M)
    for key, bad in [("trans_posns", np.zeros((P + 1, 3))), ("waveform_fft", np.ones(N + 1)),
                     ("scat_amplitudes", np.zeros(M - 1)), ("slow_time_weighting", np.ones(2))]:
        with pytest.raises(ValueError):
            SarCalculationInfo(**{**ok, key: bad})


def test_grid_positions_closed_form():
    """Spacing L/(n-1), C order over (z, y, x), columns (x, y, z), then R p + c."""
    from sarsolver_amd.numpy import GridSarScene
    R = np.array([[0.0, -1.0, 0.0], [1.0, 0.0, 0.0], [0.0, 0.0, 1.0]])
    c = np.array([10.0, -5.0, 2.0])
    sc = GridSarScene([2, 3, 4], [1.0, 2.0, 3.0], c, R)
    pos = sc.positions
    assert pos.shape == (24, 3) and sc.num_scatterers == 24
    k = 1 * 12 + 2 * 4 + 3  # z=1, y=2, x=3
    local = np.array([1.5, 1.0, 0.5])
    assert np.allclose(pos[k], R @ local + c, atol=1e-15)
    assert np.allclose(pos[0], R @ np.array([-1.5, -1.0, -0.5]) + c, atol=1e-15)


def test_blocking_helpers():
    from sarsolver_amd.base import block_hypothesis_vector, block_phase_history_array
    h = np.arange(10)
    parts = block_hypothesis_vector(h, 3)
    assert [len(p) for p in parts] == [3, 3, 4]
    assert np.array_equal(np.concatenate(parts), h)
    m = np.arange(14).reshape(7, 2)
    assert [p.shape[0] for p in block_phase_history_array(m, 2)] == [3, 4]


def test_aperture_blocks_drop_c_eff_like_the_reference():
    """The reference's to_blocks rebuilds apertures without c_eff, so blocked evaluation
    uses the vacuum speed of light (numpy/measurement.py:77-103); mirrored for parity."""
    from scipy.constants import c
    from sarsolver_amd.numpy import NumpySimpleSarAperture
    P = 6
    ap = NumpySimpleSarAperture(np.ones((P, 3)), np.ones((P, 3)), np.zeros((P, 3)), 1e9, 1e8,
                                num_freqs=4, c_eff=2.9e8)
    assert ap.c_eff == 2.9e8
    blocks = ap.to_blocks(2)
    assert [b.num_slow_times for b in blocks] == [3, 3]
    assert all(b.c_eff == c for b in blocks)
    assert ap.to_blocks(1)[0].c_eff == c
    assert ap.get_subset(slice(0, 2)).c_eff == 2.9e8


def test_freqs_and_bad_slow_times():
    from sarsolver_amd.numpy import NumpySimpleSarAperture
    tx = np.zeros((4, 3))
    tx[2, 1] = np.nan
    ap = NumpySimpleSarAperture(tx, np.zeros((4, 3)), np.zeros((4, 3)), 1e9, 4.0, num_freqs=5)
    assert np.allclose(ap.freqs, 1e9 + np.array([-1.6, -0.8, 0.0, 0.8, 1.6]))
    assert ap.bad_slow_times == [2]


def test_device_path_fails_loudly_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    from sarsolver_amd.cxx_binding import single_adjoint_evaluate
    from sarsolver_amd import synthetic
    ap = synthetic.circular_aperture(4, 8, 1.0)
    sc = synthetic.grid_scene(4, 0.5)
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        single_adjoint_evaluate(sc, ap, np.zeros((4, 8), complex))
