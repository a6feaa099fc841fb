"""Device plans for apertures and scenes, cached by content.

The reference rebuilds its native worker (and re-plans FFTW) on every call
(src/sarsolver/cxx_binding/sarsolver_cxx.cpp:219-252).  Here the geometry is
uploaded once per distinct (positions, aperture, settings) and reused, which
is what makes repeated matvecs inside a solver cheap.  The key is a digest of
every array and scalar that enters the plan, so mutated inputs get a new plan.
A GridSarScene enters by its grid description (shape, edge lengths, centre,
rotation): its positions are generated on the device and never formed,
hashed or uploaded on the host (numpy/hypothesis.py:191-206 builds them
anew on every call of the reference).
"""
from __future__ import annotations

import hashlib
import threading
from collections import OrderedDict

import numpy as np

from .._native import SarPlan, grid_args

_CACHE_SIZE = 4
_cache: "OrderedDict[bytes, SarPlan]" = OrderedDict()
_lock = threading.Lock()


def _digest(*items) -> bytes:
    h = hashlib.blake2b(digest_size=20)
    for it in items:
        if isinstance(it, np.ndarray):
            a = np.ascontiguousarray(it)
            h.update(str(a.dtype).encode() + str(a.shape).encode())
            h.update(a.data)
        else:
            h.update(repr(it).encode())
    return h.digest()


def device():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("sarsolver_amd needs a ROCm GPU (gfx950); none is visible. "
                           "There is no CPU fallback.")
    return torch.device("cuda", torch.cuda.current_device())


def plan_for(scene, aperture, upsample_ratio: float, interp: str = "nearest_ref",
             precision: str = "fp64", dev=None) -> SarPlan:
    """Plan for (scene, aperture); ``scene`` is a scene object or an [M, 3]
    position array.  The waveform is passed in FFT order (ifftshift of the
    aperture's centred spectrum, forward.py:39 / adjoint.py:38)."""
    dev = device() if dev is None else dev
    grid = grid_args(scene)
    if grid is not None:
        pos_key, pos_arg = ("grid",) + grid.key, scene
    else:
        pos = np.asarray(getattr(scene, "positions", scene), dtype=np.float64)
        pos_key, pos_arg = pos, pos
    tx = np.asarray(aperture.transmitter_positions, dtype=np.float64)
    rx = np.asarray(aperture.receiver_positions, dtype=np.float64)
    srp = np.asarray(aperture.stab_ref_posns, dtype=np.float64)
    W = np.fft.ifftshift(np.asarray(aperture.waveform_fft, dtype=np.complex128))
    w = np.asarray(aperture.slow_time_weighting, dtype=np.complex128)
    scal = (float(aperture.centre_frequency), float(aperture.sample_bandwidth),
            float(aperture.c_eff), float(upsample_ratio), float(aperture.sign_multiplier),
            interp, precision, str(dev))
    key = _digest(pos_key, tx, rx, srp, W, w, scal)
    with _lock:
        plan = _cache.get(key)
        if plan is not None:
            _cache.move_to_end(key)
            return plan
    plan = SarPlan(tx, rx, srp, pos_arg, W, w, scal[0], scal[1], scal[2], scal[3], scal[4],
                   interp=interp, precision=precision, device=dev)
    with _lock:
        _cache[key] = plan
        while len(_cache) > _CACHE_SIZE:
            _cache.popitem(last=False)[1].close()
    return plan


def clear_cache() -> None:
    with _lock:
        while _cache:
            _cache.popitem()[1].close()
