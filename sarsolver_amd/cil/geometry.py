"""CIL geometries and container for complex SAR arrays (needs CIL).

CIL's solvers only ask a geometry for `allocate`, `shape`, `dtype` and
`copy`; its own image / acquisition geometries cannot describe a SAR scene or
aperture or hold complex data, so the reference defines its own
(src/sarsolver/cil/base.py, hypothesis.py, measurement.py).  Here one
`GenericCilGeometry` carries the whole allocate contract (the fill rules live
in `_arrays.initial_array`), and the two concrete geometries only say which
object they describe and its array shape:
  CilSarGridGeometry(scene)        grid_shape, e.g. (1, ny, nx)
  SimpleCilSarGeometry(aperture)   [num_slow_times, num_freqs]
both complex128 by default (settable: CIL sets dtype on geometries).
Copies are deep, so a solver that mutates a geometry it was given does not
reach the operator's.
"""
from __future__ import annotations

from copy import deepcopy
from typing import List, Optional, Sequence

import numpy as np
from cil.framework import DataContainer

from ._arrays import initial_array


class GenericCilGeometry:
    """A description of arrays of one shape and dtype that CIL can allocate."""

    _dtype = np.complex128

    @property
    def shape(self) -> Sequence[int]:
        raise NotImplementedError

    @property
    def dtype(self) -> type:
        return self._dtype

    @dtype.setter
    def dtype(self, val: type):
        self._dtype = val

    def allocate(self, value=0.0, dtype: Optional[type] = None, seed: Optional[int] = None,
                 max_value: int = 100) -> "ExtendedDataContainer":
        arr = initial_array(self.shape, self.dtype if dtype is None else dtype, value, seed,
                            max_value, geometry_dtype=self.dtype)
        return ExtendedDataContainer(array=arr, geometry=self.copy())

    def copy(self) -> "GenericCilGeometry":
        return deepcopy(self)


class ExtendedDataContainer(DataContainer):
    """A DataContainer whose geometry is a GenericCilGeometry (CIL's own
    container only accepts its image and acquisition geometries)."""

    def __init__(self, *args, **kwargs):
        # DataContainer.__init__ reads the label attribute before it sets it
        self._dimension_labels = None
        super().__init__(*args, **kwargs)

    @property
    def geometry(self) -> GenericCilGeometry:
        return self._geometry

    @geometry.setter
    def geometry(self, val: GenericCilGeometry):
        self._geometry = val


class CilSarGridGeometry(GenericCilGeometry):
    """Images of a GridSarScene: arrays of its grid_shape."""

    def __init__(self, scene):
        self._scene = scene
        self._dtype = np.complex128

    @property
    def scene(self):
        return self._scene

    @property
    def shape(self) -> Sequence[int]:
        return tuple(self._scene.grid_shape)


class SimpleCilSarGeometry(GenericCilGeometry):
    """Phase histories of an aperture: [num_slow_times, num_freqs] arrays."""

    def __init__(self, aperture):
        self._aperture = aperture
        self._dtype = np.complex128

    @property
    def aperture(self):
        return self._aperture

    @property
    def shape(self) -> List[int]:
        return [self._aperture.num_slow_times, self._aperture.num_freqs]
