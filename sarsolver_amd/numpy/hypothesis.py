"""Scenes: point clouds and regular 3-D grids (reference: numpy/hypothesis.py)."""
from __future__ import annotations

from typing import List, Optional, Sequence

import numpy as np
from numpy.typing import NDArray

from ..base import BaseSarScene, BaseSimpleSarAperture
from ..utils import scene_params_classic_sar, scene_params_enu


def _slices(n: int, num_blocks: int):
    step = n // num_blocks
    bounds = [k * step for k in range(num_blocks)] + [n]
    return [slice(bounds[k], bounds[k + 1]) for k in range(num_blocks)]


class SimpleSarScene(BaseSarScene):
    """Scatterers at arbitrary positions (numpy/hypothesis.py:10-70)."""

    def __init__(self, scatterer_positions: NDArray):
        self._scatterer_positions = scatterer_positions
        self._num_scatterers = scatterer_positions.shape[0]

    @property
    def positions(self) -> NDArray:
        return self._scatterer_positions

    @property
    def num_scatterers(self) -> int:
        return self._num_scatterers

    def to_blocks(self, num_blocks: int = 1) -> List[BaseSarScene]:
        if num_blocks == 1:
            return [self]
        pos = self.positions
        return [SimpleSarScene(pos[s, :]) for s in _slices(self.num_scatterers, num_blocks)]


class GridSarScene(BaseSarScene):
    """A regular grid of scatterers filling a (rotated, translated) cuboid
    (numpy/hypothesis.py:73-267).  ``grid_shape`` and ``grid_dimensions`` are in
    (z, y, x) order; points reach the faces, so the spacing is L / (n - 1)."""

    def __init__(self, grid_shape: Sequence[int], grid_dimensions: Sequence[float],
                 grid_centre: Optional[NDArray] = None, rotation_matrix: Optional[NDArray] = None):
        self._grid_centre = np.zeros(3) if grid_centre is None else grid_centre
        self._rotation_matrix = np.identity(3) if rotation_matrix is None else rotation_matrix
        self._grid_shape = grid_shape
        self._grid_dimensions = grid_dimensions

    def to_blocks(self, num_blocks: int = 1) -> List[BaseSarScene]:
        if num_blocks == 1:
            return [self]
        pos = self.positions
        return [SimpleSarScene(pos[s, :]) for s in _slices(self.num_scatterers, num_blocks)]

    @property
    def num_scatterers(self) -> int:
        return int(np.prod(np.array(self.grid_shape)))

    @property
    def rotation_matrix(self) -> NDArray:
        return self._rotation_matrix

    @property
    def grid_centre(self) -> NDArray:
        return self._grid_centre

    @property
    def grid_dimensions(self) -> Sequence[float]:
        return self._grid_dimensions

    @property
    def grid_shape(self) -> Sequence[int]:
        return self._grid_shape

    @property
    def positions(self) -> NDArray:
        """[M, 3] positions, flattened in C order over (z, y, x), columns (x, y, z),
        then rotated and translated (numpy/hypothesis.py:191-206)."""
        axes = [np.linspace(-0.5 * L, 0.5 * L, n)
                for L, n in zip(self.grid_dimensions, self.grid_shape)]  # z, y, x
        zz, yy, xx = np.meshgrid(axes[0], axes[1], axes[2], indexing="ij")
        local = np.stack([xx.reshape(-1), yy.reshape(-1), zz.reshape(-1)], axis=-1)
        return (self.rotation_matrix @ local.T).T + np.expand_dims(self.grid_centre, axis=0)

    def device_positions(self, device=None):
        """``positions`` generated on the GPU: an [M, 3] float64 torch tensor,
        bit-identical to the host array (linspace steps as numpy takes them;
        R @ local as an FMA chain in k order)."""
        from .._native import grid_positions
        return grid_positions(self, device)

    @staticmethod
    def from_geodetic_enu(grid_shape: Sequence[int], grid_dimensions: Sequence[float],
                          scene_centre_geodetic: NDArray) -> "GridSarScene":
        """Grid aligned with local East/North/Up at a geodetic centre (numpy/hypothesis.py:208-232)."""
        centre_ecef, rotation = scene_params_enu(scene_centre_geodetic)
        return GridSarScene(grid_shape, grid_dimensions, centre_ecef, rotation)

    @staticmethod
    def from_aperture(aperture: BaseSimpleSarAperture, scene_centre_ecef: NDArray,
                      grid_dimensions: Sequence[float], grid_shape: Optional[Sequence[int]] = None,
                      safety_factor: float = 1.2) -> "GridSarScene":
        """Grid aligned with range / cross-range / out-of-plane of an aperture, sampled at
        safety_factor times the k-space Nyquist rate (numpy/hypothesis.py:235-267)."""
        rot, rates, _ = scene_params_classic_sar(scene_centre_ecef, aperture)
        if grid_shape is None:
            n = np.clip(np.array(grid_dimensions) * rates * safety_factor, a_min=1.0, a_max=None)
            grid_shape = np.rint(n).astype(int).tolist()
        return GridSarScene(grid_shape, grid_dimensions, scene_centre_ecef, rot)
