"""Oriented bounding boxes of k-space support and simple basebanding
(reference: utils/k_space.py:9-134)."""
from __future__ import annotations

from typing import Optional

import numpy as np
from numpy.typing import NDArray


def get_obb_basis(xs: NDArray):
    """Eigen-decomposition of the scatter matrix of the centred points (utils/k_space.py:9-25)."""
    zs = xs - np.mean(xs, axis=0, keepdims=True)
    return np.linalg.eigh(np.dot(zs.T, zs))


class OrientedBoundingBox:
    """PCA-aligned bounding box of a point cloud.  The axis whose box-centre
    coordinate is largest in magnitude is "range"; "cross range" is the widest
    other axis (or the middle-width axis when range is itself the widest); the
    remaining axis is "out of plane" (utils/k_space.py:28-74)."""

    def __init__(self, xs: NDArray):
        xs = xs[np.isfinite(xs).all(axis=1)]
        _, basis = get_obb_basis(xs)
        if np.linalg.det(basis) < 0.0:
            basis = -basis
        local = np.dot(basis.T, xs.T).T
        lo, hi = np.min(local, axis=0), np.max(local, axis=0)
        centre = 0.5 * (lo + hi)
        spans = np.abs(hi - lo)
        range_index = int(np.argmax(np.abs(centre)))
        if spans[range_index] < np.max(spans):
            cross_index = int(np.argmax(spans))
        else:
            cross_index = int(np.argsort(spans)[1])
        oop_index = int(np.delete(np.arange(3), [range_index, cross_index])[0])
        self.range_index = range_index
        self.cross_range_index = cross_index
        self.out_of_plane_index = oop_index
        self.range_span = spans[range_index]
        self.cross_range_span = spans[cross_index]
        self.out_of_plane_span = spans[oop_index]
        self.obb_basis = basis
        self.displacement = np.dot(basis, centre.T).T
        self.spans = spans

    def to_obb_coords(self, points: NDArray) -> NDArray:
        return np.dot(self.obb_basis.T, (points - np.expand_dims(self.displacement, 0)).T).T

    def from_obb_coords(self, points: NDArray) -> NDArray:
        return np.dot(self.obb_basis, points.T).T + self.displacement


def generate_simple_basebanding_values(scene, aperture, basebanding_spot: Optional[NDArray] = None,
                                       device=None):
    """Phases exp(-i sgn (x - spot) . d) that baseband an image about one point,
    d the displacement of the k-space box of the aperture's band-edge wave
    vectors seen from the spot (utils/k_space.py:109-134).  The spot defaults
    to the mean scene position.

    With ``device`` (a torch device), the phases are computed on the GPU and
    returned as an [M] complex128 torch tensor; a GridSarScene's positions are
    then generated there and never formed on the host (the default spot is
    their mean, taken on the device)."""
    if device is None:
        pos = scene.positions
        spot = np.mean(pos, axis=0) if basebanding_spot is None else np.asarray(basebanding_spot)
        box = OrientedBoundingBox(aperture.edge_k_vecs(spot))
        phases = (pos - np.expand_dims(spot, axis=0)) @ box.displacement
        return np.exp(-1.j * aperture.sign_multiplier * phases)
    import torch

    from .._native import baseband_values, grid_args, grid_positions
    dev = torch.device(device)
    if grid_args(scene) is not None:
        pos = grid_positions(scene, dev)
    else:
        pos = torch.as_tensor(np.ascontiguousarray(scene.positions, dtype=np.float64)).to(dev)
    if basebanding_spot is None:
        spot = pos.mean(dim=0).cpu().numpy()
    else:
        spot = np.asarray(basebanding_spot, dtype=np.float64)
    box = OrientedBoundingBox(aperture.edge_k_vecs(spot))
    return baseband_values(pos, spot, box.displacement, float(aperture.sign_multiplier))
