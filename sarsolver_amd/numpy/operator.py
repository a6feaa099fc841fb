"""scipy LinearOperator over the gfx950 SAR operators (reference: numpy/operator.py)."""
from __future__ import annotations

from typing import Optional

import numpy as np
from numpy.typing import NDArray
from scipy.sparse.linalg import LinearOperator

from ..base import BaseSarScene, BaseSimpleSarAperture
from ..cxx_binding import multi_forward_evaluate, multi_adjoint_evaluate


class NumpySimpleSarOperator(LinearOperator):
    """Scalar Born SAR operator A: [M] scene -> [P*N] phase history, complex128
    (numpy/operator.py:11-56).  ``interp`` / ``precision`` select the device
    arithmetic; the defaults reproduce the reference's bin rules."""

    def __init__(self, scene: BaseSarScene, aperture: BaseSimpleSarAperture,
                 num_threads: Optional[int] = None, upsample_ratio: float = 128.0,
                 verbose: bool = False, num_calculation_blocks: int = 1, *,
                 interp: str = "nearest_ref", precision: str = "fp64"):
        super().__init__(dtype=np.complex128,
                         shape=[aperture.num_freqs * aperture.num_slow_times,
                                scene.num_scatterers])
        self._scene = scene
        self._aperture = aperture
        self._num_calculation_blocks = num_calculation_blocks
        self._upsample_ratio = upsample_ratio
        self._verbose = verbose
        self._num_threads = num_threads
        self._interp = interp
        self._precision = precision

    def _matvec(self, x: NDArray) -> NDArray:
        u = x.reshape([self._scene.num_scatterers])
        v = multi_forward_evaluate(scene=self._scene, measurements=self._aperture,
                                   scattering_amplitudes=u, num_threads=self._num_threads,
                                   num_calculation_blocks=self._num_calculation_blocks,
                                   upsample_ratio=self._upsample_ratio, verbose=self._verbose,
                                   interp=self._interp, precision=self._precision)
        return v.reshape([self._aperture.num_freqs * self._aperture.num_slow_times])

    def _rmatvec(self, y: NDArray) -> NDArray:
        v = y.reshape([self._aperture.num_slow_times, self._aperture.num_freqs])
        u = multi_adjoint_evaluate(scene=self._scene, measurements=self._aperture,
                                   adjoint_amplitudes=v, num_threads=self._num_threads,
                                   num_calculation_blocks=self._num_calculation_blocks,
                                   upsample_ratio=self._upsample_ratio, verbose=self._verbose,
                                   interp=self._interp, precision=self._precision)
        return u.reshape([self._scene.num_scatterers])
