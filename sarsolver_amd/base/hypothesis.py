"""Scene (hypothesis space) interface, as in the reference's base/hypothesis.py."""
from __future__ import annotations

from abc import ABC, abstractmethod
from typing import List

from numpy.typing import NDArray


def block_hypothesis_vector(h: NDArray, num_blocks: int = 1) -> List[NDArray]:
    """Split ``h`` along axis 0 into ``num_blocks`` pieces of length len(h)//num_blocks,
    the last piece taking the remainder (reference: base/hypothesis.py:7-31)."""
    if num_blocks == 1:
        return [h]
    step = h.shape[0] // num_blocks
    bounds = [k * step for k in range(num_blocks)] + [h.shape[0]]
    return [h[bounds[k]:bounds[k + 1]] for k in range(num_blocks)]


class BaseSarScene(ABC):
    """A set of point scatterers with Cartesian positions (base/hypothesis.py:34-79)."""

    @property
    @abstractmethod
    def positions(self) -> NDArray:
        """[num_scatterers, 3] float array of scatterer positions, metres."""
        raise NotImplementedError

    @property
    @abstractmethod
    def num_scatterers(self) -> int:
        raise NotImplementedError

    @abstractmethod
    def to_blocks(self, num_blocks: int = 1) -> List["BaseSarScene"]:
        """Split into ``num_blocks`` sub-scenes (used for blocked evaluation)."""
        raise NotImplementedError
