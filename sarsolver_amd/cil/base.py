"""Complex-capable CIL geometry / container (reference: cil/base.py:11-136)."""
from __future__ import annotations

from abc import ABC, abstractmethod
from copy import deepcopy
from numbers import Number
from typing import Optional, Sequence

import numpy as np
from cil.framework import DataContainer
from cil.framework.labels import FillType


class GenericCilGeometry(ABC):
    """Allocates ExtendedDataContainers of a given shape and dtype; copies are deep."""

    @property
    @abstractmethod
    def shape(self) -> Sequence[int]:
        raise NotImplementedError

    @property
    def dtype(self) -> type:
        raise NotImplementedError

    def allocate(self, value: Number = 0.0, dtype: Optional[type] = None,
                 seed: Optional[int] = None, max_value: int = 100) -> "ExtendedDataContainer":
        dtype = self.dtype if dtype is None else dtype
        out = ExtendedDataContainer(array=np.full(shape=self.shape, fill_value=0, dtype=dtype),
                                    geometry=self.copy())
        if value is None:
            return out
        if value == FillType.RANDOM:
            if seed is not None:
                np.random.seed(seed)
            r = np.random.random_sample(self.shape)
            if np.iscomplexobj(out.array):
                r = r + 1j * np.random.random_sample(self.shape)
            out.fill(r)
            return out
        if value == FillType.RANDOM_INT:
            if seed is not None:
                np.random.seed(seed)
            r = np.random.randint(max_value, size=self.shape, dtype=np.int32)
            out.fill(np.asarray(r, dtype=self.dtype))
            return out
        out.fill(value)
        return out

    def copy(self) -> "GenericCilGeometry":
        return deepcopy(self)


class ExtendedDataContainer(DataContainer):
    """DataContainer whose geometry may be any GenericCilGeometry."""

    def __init__(self, *args, **kwargs):
        self._dimension_labels = None
        super().__init__(*args, **kwargs)

    @property
    def geometry(self) -> GenericCilGeometry:
        return self._geometry

    @geometry.setter
    def geometry(self, val: GenericCilGeometry):
        self._geometry = val
