"""Backprojection A^H y through the gfx950 library (reference: cxx_binding/adjoint.py)."""
from __future__ import annotations

from typing import Optional

import numpy as np
from numpy.typing import NDArray

from ..base import BaseSarScene, BaseSimpleSarAperture, block_phase_history_array
from ._plans import device, plan_for, plans_for_blocks


def _ph_to_device(y: NDArray, dev):
    import torch
    return torch.from_numpy(np.ascontiguousarray(np.asarray(y), dtype=np.complex64)).to(dev)


def single_adjoint_evaluate(scene: BaseSarScene, measurements: BaseSimpleSarAperture,
                            adjoint_amplitudes: NDArray, upsample_ratio: float = 128.0, *,
                            interp: str = "nearest_ref", precision: str = "fp64") -> NDArray:
    """Backprojected amplitudes [M] (complex128) of a phase history [P, N] in centred
    frequency order (adjoint.py:13-44)."""
    dev = device()
    plan = plan_for(scene, measurements, upsample_ratio, interp, precision, dev)
    x = plan.adjoint(_ph_to_device(adjoint_amplitudes, dev), order="centred")
    return x.cpu().numpy().astype(np.complex128)


def multi_adjoint_evaluate(scene: BaseSarScene, measurements: BaseSimpleSarAperture,
                           adjoint_amplitudes: NDArray, num_threads: Optional[int] = None,
                           upsample_ratio: float = 128.0, verbose: bool = False,
                           num_calculation_blocks: int = 1, *, interp: str = "nearest_ref",
                           precision: str = "fp64") -> NDArray:
    """Same signature as the reference (adjoint.py:47-94): pulse blocks run
    sequentially and their backprojections are summed; ``num_threads`` has no
    effect on the device path."""
    dev = device()
    plans = plans_for_blocks(scene, measurements, num_calculation_blocks, upsample_ratio, interp,
                             precision, dev)
    y_blocks = block_phase_history_array(np.asarray(adjoint_amplitudes), num_calculation_blocks)
    it = list(zip(plans, y_blocks))
    if verbose:
        from tqdm import tqdm
        it = tqdm(it)
    total = None
    for plan, yb in it:
        part = plan.adjoint(_ph_to_device(yb, dev), order="centred").cpu().numpy()
        part = part.astype(np.complex128)
        total = part if total is None else total + part
    return total
