from typing import Sequence

import numpy as np

from .base import GenericCilGeometry
from ..numpy import GridSarScene


class CilSarGridGeometry(GenericCilGeometry):
    """CIL geometry of a GridSarScene; arrays are grid_shape, complex128 (cil/hypothesis.py:9-71)."""

    def __init__(self, scene: GridSarScene):
        self._scene = scene
        self._dtype = np.complex128

    @property
    def scene(self) -> GridSarScene:
        return self._scene

    @property
    def shape(self) -> Sequence[int]:
        return tuple(self._scene.grid_shape)

    @property
    def dtype(self) -> type:
        return self._dtype

    @dtype.setter
    def dtype(self, val: type):
        self._dtype = val
