"""CIL LinearOperators over the gfx950 SAR operators (needs CIL).

`CpuSarOperator` keeps the reference's name and constructor
(src/sarsolver/cil/operator.py:11-111) so CIL scripts run unchanged; the
work runs on the GPU through multi_forward / multi_adjoint_evaluate.
`GpuSarOperator` takes the same arguments and binds its device plans once,
at construction, instead of looking them up by content on every call.
Both move arrays between CIL containers and `SarArrayOperator`, which holds
the shape logic: direct writes a [P, N] phase history into the range
container, adjoint a grid-shaped image into the domain container; `out`,
when given, receives the result array and is returned.
"""
from typing import Optional

from cil.framework import DataContainer
from cil.optimisation.operators import LinearOperator

from ._core import SarArrayOperator
from .geometry import CilSarGridGeometry, SimpleCilSarGeometry


class CpuSarOperator(LinearOperator):
    _bind_plans = False

    def __init__(self, image_geometry: CilSarGridGeometry,
                 measurement_geometry: SimpleCilSarGeometry, upsample_ratio: float = 128.0,
                 num_threads: Optional[int] = None, verbose: bool = False,
                 num_calculation_blocks: Optional[int] = 1, *, interp: str = "nearest_ref",
                 precision: str = "fp64"):
        # CIL's Operator keeps its geometries and cached norm in these fields
        self._domain_geometry = image_geometry
        self._range_geometry = measurement_geometry
        self._norm = None
        self._arrays = SarArrayOperator(image_geometry.scene, measurement_geometry.aperture,
                                        upsample_ratio, num_threads, verbose,
                                        num_calculation_blocks, interp, precision,
                                        bind_plans=self._bind_plans)

    @property
    def domain(self) -> CilSarGridGeometry:
        return self._domain_geometry

    @property
    def range(self) -> SimpleCilSarGeometry:
        return self._range_geometry

    @staticmethod
    def _deliver(result, out: Optional[DataContainer], geometry) -> DataContainer:
        if out is None:
            out = geometry.allocate(None)
        out.array = result
        return out

    def direct(self, x: DataContainer, out: Optional[DataContainer] = None) -> DataContainer:
        return self._deliver(self._arrays.direct_array(x.array), out, self.range)

    def adjoint(self, x: DataContainer, out: Optional[DataContainer] = None) -> DataContainer:
        return self._deliver(self._arrays.adjoint_array(x.array), out, self.domain)


class GpuSarOperator(CpuSarOperator):
    _bind_plans = True
