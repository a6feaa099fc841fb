"""Forward projection A x through the gfx950 library (reference: cxx_binding/forward.py)."""
from __future__ import annotations

from typing import Optional

import numpy as np
from numpy.typing import NDArray

from ..base import BaseSarScene, BaseSimpleSarAperture
from ._plans import device, plan_for, plans_for_blocks


def _to_device_c64(x: NDArray, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(np.asarray(x).reshape(-1),
                                                 dtype=np.complex64)).to(dev)


def single_forward_evaluate(scene: BaseSarScene, measurements: BaseSimpleSarAperture,
                            scattering_amplitudes: NDArray, upsample_ratio: float = 128.0, *,
                            interp: str = "nearest_ref", precision: str = "fp64") -> NDArray:
    """Phase history [P, N] (complex128, centred frequency order) predicted from the
    scattering amplitudes (forward.py:13-44).  Device arithmetic is complex64 with
    FP64 ranges; inputs are cast at this boundary."""
    dev = device()
    plan = plan_for(scene, measurements, upsample_ratio, interp, precision, dev)
    y = plan.forward(_to_device_c64(scattering_amplitudes, dev), order="centred")
    return y.cpu().numpy().astype(np.complex128)


def multi_forward_evaluate(scene: BaseSarScene, measurements: BaseSimpleSarAperture,
                           scattering_amplitudes: NDArray, num_threads: Optional[int] = None,
                           upsample_ratio: float = 128.0, verbose: bool = False,
                           num_calculation_blocks: int = 1, *, interp: str = "nearest_ref",
                           precision: str = "fp64") -> NDArray:
    """Same signature as the reference (forward.py:47-96).  Pulse blocks
    (``num_calculation_blocks``) run sequentially and are concatenated; the scene is
    not split (``num_threads`` is accepted and has no effect: one device launch
    covers every pixel)."""
    dev = device()
    plans = plans_for_blocks(scene, measurements, num_calculation_blocks, upsample_ratio, interp,
                             precision, dev)
    x = _to_device_c64(scattering_amplitudes, dev)
    it = plans
    if verbose:
        from tqdm import tqdm
        it = tqdm(plans)
    chunks = []
    for plan in it:
        chunks.append(plan.forward(x, order="centred").cpu().numpy().astype(np.complex128))
    return np.concatenate(chunks, axis=0)
