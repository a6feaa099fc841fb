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

_CACHE_SIZE = 4  # cache entries; an entry holds the plans of every pulse block of a call
_cache: "OrderedDict[bytes, list]" = OrderedDict()
_lock = threading.Lock()
stats = {"built": 0}  # plans created (a cache miss builds one plan per pulse block)


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


def _scene_key(scene):
    grid = grid_args(scene)
    if grid is not None:
        return ("grid",) + grid.key, scene
    pos = np.asarray(getattr(scene, "positions", scene), dtype=np.float64)
    return pos, pos


def _aperture_arrays(aperture):
    return (np.asarray(aperture.transmitter_positions, dtype=np.float64),
            np.asarray(aperture.receiver_positions, dtype=np.float64),
            np.asarray(aperture.stab_ref_posns, dtype=np.float64),
            np.fft.ifftshift(np.asarray(aperture.waveform_fft, dtype=np.complex128)),
            np.asarray(aperture.slow_time_weighting, dtype=np.complex128))


def _scalars(aperture, upsample_ratio, interp, precision, dev):
    return (float(aperture.centre_frequency), float(aperture.sample_bandwidth),
            float(aperture.c_eff), float(upsample_ratio), float(aperture.sign_multiplier),
            interp, precision, str(dev))


def _build(pos_arg, aperture, scal, interp, precision, dev) -> SarPlan:
    tx, rx, srp, W, w = _aperture_arrays(aperture)
    stats["built"] += 1
    return SarPlan(tx, rx, srp, pos_arg, W, w, scal[0], scal[1], scal[2], scal[3], scal[4],
                   interp=interp, precision=precision, device=dev)


def _cached(key, make):
    with _lock:
        plans = _cache.get(key)
        if plans is not None:
            _cache.move_to_end(key)
            return plans
    plans = make()
    with _lock:
        _cache[key] = plans
        while len(_cache) > _CACHE_SIZE:
            for pl in _cache.popitem(last=False)[1]:
                pl.close()
    return plans


def plan_for(scene, aperture, upsample_ratio: float, interp: str = "nearest_ref",
             precision: str = "fp64", dev=None) -> SarPlan:
    """Plan for (scene, aperture); ``scene`` is a scene object or an [M, 3]
    position array.  The waveform is passed in FFT order (ifftshift of the
    aperture's centred spectrum, forward.py:39 / adjoint.py:38)."""
    dev = device() if dev is None else dev
    pos_key, pos_arg = _scene_key(scene)
    scal = _scalars(aperture, upsample_ratio, interp, precision, dev)
    key = _digest(pos_key, *_aperture_arrays(aperture), scal)
    return _cached(key, lambda: [_build(pos_arg, aperture, scal, interp, precision, dev)])[0]


def plans_for_blocks(scene, aperture, num_blocks: int, upsample_ratio: float,
                     interp: str = "nearest_ref", precision: str = "fp64", dev=None):
    """One plan per pulse block of ``aperture.to_blocks(num_blocks)`` (the
    reference's multi_* decomposition, forward.py:77-96 / adjoint.py:75-94),
    cached as one entry keyed by the whole aperture and the block count, so a
    solver calling multi_* with any number of blocks rebuilds nothing."""
    dev = device() if dev is None else dev
    if int(num_blocks) < 1:
        raise ValueError("num_calculation_blocks must be >= 1")
    pos_key, pos_arg = _scene_key(scene)
    blocks = aperture.to_blocks(int(num_blocks))
    # every block is rebuilt by to_blocks with the same scalars (the vacuum
    # speed of light included, as the reference's blocks carry)
    scal = _scalars(blocks[0], upsample_ratio, interp, precision, dev)
    key = _digest("blocks", int(num_blocks), pos_key, *_aperture_arrays(aperture), scal)
    return _cached(key, lambda: [_build(pos_arg, b, scal, interp, precision, dev)
                                 for b in blocks])


def clear_cache() -> None:
    with _lock:
        while _cache:
            for pl in _cache.popitem()[1]:
                pl.close()
