"""Array-level SAR operator behind the CIL wrappers, without CIL.

The CIL operator of the reference (src/sarsolver/cil/operator.py:65-111)
flattens the image container to [M], calls multi_forward_evaluate, and puts
the [P, N] phase history in the range container; the adjoint reshapes the
[M] backprojection to the domain's grid shape (1, ny, nx).  `SarArrayOperator`
does exactly that on numpy arrays, so the CIL classes only move arrays in
and out of containers.

With ``bind_plans=True`` the operator takes its device plans (one per pulse
block) when it is built and keeps them, so a CIL solver's matvecs run on the
GPU without digesting the scene and aperture on every call.
"""
from __future__ import annotations

from typing import Optional, Tuple

import numpy as np

from ..base import block_phase_history_array
from ..cxx_binding import multi_adjoint_evaluate, multi_forward_evaluate


class SarArrayOperator:
    def __init__(self, scene, aperture, upsample_ratio: float = 128.0,
                 num_threads: Optional[int] = None, verbose: bool = False,
                 num_calculation_blocks: Optional[int] = 1, interp: str = "nearest_ref",
                 precision: str = "fp64", bind_plans: bool = False):
        self.scene, self.aperture = scene, aperture
        self.options = dict(num_threads=num_threads, upsample_ratio=upsample_ratio,
                            verbose=verbose, num_calculation_blocks=num_calculation_blocks or 1,
                            interp=interp, precision=precision)
        self._plans = None
        if bind_plans:
            from ..cxx_binding._plans import device, plans_for_blocks
            self._device = device()
            self._plans = plans_for_blocks(scene, aperture, num_calculation_blocks or 1,
                                           upsample_ratio, interp, precision, self._device)

    @property
    def domain_shape(self) -> Tuple[int, ...]:
        return tuple(int(s) for s in self.scene.grid_shape)

    @property
    def range_shape(self) -> Tuple[int, int]:
        return (int(self.aperture.num_slow_times), int(self.aperture.num_freqs))

    def direct_array(self, image: np.ndarray) -> np.ndarray:
        """Image (any shape with M elements) -> phase history [P, N] complex128."""
        flat = np.asarray(image).reshape(-1)
        if self._plans is None:
            return multi_forward_evaluate(self.scene, self.aperture, flat, **self.options)
        import torch
        x = torch.from_numpy(np.ascontiguousarray(flat, dtype=np.complex64)).to(self._device)
        rows = [pl.forward(x, order="centred").cpu().numpy() for pl in self._plans]
        return np.concatenate(rows, axis=0).astype(np.complex128)

    def adjoint_array(self, phase_history: np.ndarray) -> np.ndarray:
        """Phase history [P, N] -> backprojected image in the domain's shape."""
        ph = np.asarray(phase_history).reshape(self.range_shape)
        if self._plans is None:
            flat = multi_adjoint_evaluate(self.scene, self.aperture, ph, **self.options)
        else:
            import torch
            total = None
            blocks = block_phase_history_array(ph, len(self._plans))
            for pl, yb in zip(self._plans, blocks):
                y = torch.from_numpy(np.ascontiguousarray(yb, dtype=np.complex64)).to(self._device)
                part = pl.adjoint(y, order="centred").to(torch.complex128)
                total = part if total is None else total + part
            flat = total.cpu().numpy()
        return flat.reshape(self.domain_shape)
