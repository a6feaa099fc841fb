from typing import List

import numpy as np

from .base import GenericCilGeometry
from ..base import BaseSimpleSarAperture


class SimpleCilSarGeometry(GenericCilGeometry):
    """CIL geometry of an aperture; arrays are [P, N] complex128 (cil/measurement.py:9-71)."""

    def __init__(self, aperture: BaseSimpleSarAperture):
        self._aperture = aperture
        self._dtype = np.complex128

    @property
    def aperture(self) -> BaseSimpleSarAperture:
        return self._aperture

    @property
    def shape(self) -> List[int]:
        return [self.aperture.num_slow_times, self.aperture.num_freqs]

    @property
    def dtype(self) -> type:
        return self._dtype

    @dtype.setter
    def dtype(self, val: type):
        self._dtype = val
