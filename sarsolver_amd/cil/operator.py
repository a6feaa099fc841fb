"""CIL LinearOperator over the gfx950 SAR operators (reference: cil/operator.py:11-111)."""
from typing import Optional

from cil.framework import DataContainer
from cil.optimisation.operators import LinearOperator

from .hypothesis import CilSarGridGeometry
from .measurement import SimpleCilSarGeometry
from ..cxx_binding import multi_adjoint_evaluate, multi_forward_evaluate


class CpuSarOperator(LinearOperator):
    """Same name and constructor as the reference's CIL operator, so CIL scripts run
    unchanged; the work runs on the GPU.  ``interp`` / ``precision`` are extra
    keyword-only options."""

    def __init__(self, image_geometry: CilSarGridGeometry,
                 measurement_geometry: SimpleCilSarGeometry, upsample_ratio: float = 128.0,
                 num_threads: Optional[int] = None, verbose: bool = False,
                 num_calculation_blocks: Optional[int] = 1, *, interp: str = "nearest_ref",
                 precision: str = "fp64"):
        self._domain_geometry = image_geometry
        self._range_geometry = measurement_geometry
        self._norm = None
        self._upsample_ratio = upsample_ratio
        self._num_threads = num_threads
        self._num_calculation_blocks = num_calculation_blocks
        self._verbose = verbose
        self._interp = interp
        self._precision = precision

    @property
    def domain(self) -> CilSarGridGeometry:
        return self._domain_geometry

    @property
    def range(self) -> SimpleCilSarGeometry:
        return self._range_geometry

    def direct(self, x: DataContainer, out: Optional[DataContainer] = None) -> DataContainer:
        if out is None:
            out = self.range.allocate()
        out.array = multi_forward_evaluate(self.domain.scene, self.range.aperture,
                                           x.array.reshape([-1]), num_threads=self._num_threads,
                                           upsample_ratio=self._upsample_ratio,
                                           verbose=self._verbose,
                                           num_calculation_blocks=self._num_calculation_blocks,
                                           interp=self._interp, precision=self._precision)
        return out

    def adjoint(self, x: DataContainer, out: Optional[DataContainer] = None) -> DataContainer:
        if out is None:
            out = self.domain.allocate()
        h = multi_adjoint_evaluate(self.domain.scene, self.range.aperture, x.array,
                                   num_threads=self._num_threads,
                                   upsample_ratio=self._upsample_ratio, verbose=self._verbose,
                                   num_calculation_blocks=self._num_calculation_blocks,
                                   interp=self._interp, precision=self._precision)
        out.array = h.reshape(self.domain.shape)
        return out


GpuSarOperator = CpuSarOperator
