"""Scene orientation helpers (reference: utils/geodesy.py:99-172).

The reference converts between ECEF and geodetic coordinates with pymap3d,
which is not a dependency here; the WGS84 conversions below are this
module's own (closed-form geodetic->ECEF, fixed-point iteration for the
inverse)."""
from __future__ import annotations

import numpy as np
from numpy.typing import NDArray
from scipy.constants import pi

from .k_space import OrientedBoundingBox

WGS84_A = 6378137.0
WGS84_F = 1.0 / 298.257223563
WGS84_E2 = WGS84_F * (2.0 - WGS84_F)


def geodetic_to_ecef(lat_deg: float, lon_deg: float, alt: float) -> NDArray:
    lat, lon = np.radians(lat_deg), np.radians(lon_deg)
    n = WGS84_A / np.sqrt(1.0 - WGS84_E2 * np.sin(lat) ** 2)
    return np.array([(n + alt) * np.cos(lat) * np.cos(lon),
                     (n + alt) * np.cos(lat) * np.sin(lon),
                     (n * (1.0 - WGS84_E2) + alt) * np.sin(lat)])


def ecef_to_geodetic(x: NDArray):
    x = np.asarray(x, dtype=np.float64)
    lon = np.arctan2(x[1], x[0])
    p = np.hypot(x[0], x[1])
    lat = np.arctan2(x[2], p * (1.0 - WGS84_E2))
    alt = 0.0
    for _ in range(50):
        n = WGS84_A / np.sqrt(1.0 - WGS84_E2 * np.sin(lat) ** 2)
        alt = p / np.cos(lat) - n if abs(np.cos(lat)) > 1e-12 else abs(x[2]) - n * (1 - WGS84_E2)
        new = np.arctan2(x[2], p * (1.0 - WGS84_E2 * n / (n + alt)))
        if abs(new - lat) < 1e-15:
            lat = new
            break
        lat = new
    return np.degrees(lat), np.degrees(lon), alt


def enu_basis(lat_deg: float, lon_deg: float) -> NDArray:
    """Columns east, north, up at a geodetic location."""
    lat, lon = np.radians(lat_deg), np.radians(lon_deg)
    east = np.array([-np.sin(lon), np.cos(lon), 0.0])
    north = np.array([-np.sin(lat) * np.cos(lon), -np.sin(lat) * np.sin(lon), np.cos(lat)])
    up = np.array([np.cos(lat) * np.cos(lon), np.cos(lat) * np.sin(lon), np.sin(lat)])
    return np.stack([east, north, up], axis=-1)


def normalise(x: NDArray, tol: float = 1.0e-12) -> NDArray:
    """Unit vector; refuses vectors shorter than ``tol`` (utils/geodesy.py:11-30)."""
    norm = np.linalg.norm(x, axis=-1)
    if norm < tol:
        raise ValueError("Shouldn't normalise vectors of zero/negligible length!")
    return x / norm


def scene_params_enu(scene_centre_geodetic: NDArray):
    """ECEF centre and ENU-aligned rotation matrix (utils/geodesy.py:99-121)."""
    lat, lon, alt = (float(v) for v in scene_centre_geodetic)
    return geodetic_to_ecef(lat, lon, alt), enu_basis(lat, lon)


def scene_params_classic_sar(scene_centre_ecef: NDArray, aperture, tol: float = 1.0e-6):
    """Rotation to (range, cross range, out of plane) axes of an aperture at a scene
    centre, Nyquist spatial sample rates (k-space box spans / 2 pi) and ground-plane
    cosines (utils/geodesy.py:124-172)."""
    if np.linalg.norm(scene_centre_ecef) > tol:
        lat, lon, _ = ecef_to_geodetic(scene_centre_ecef)
        up = enu_basis(lat, lon)[:, 2]
    else:
        up = np.array([0.0, 0.0, 1.0])
    up = normalise(up)
    obb = OrientedBoundingBox(aperture.edge_k_vecs())
    axes = obb.obb_basis.T
    rng = normalise(axes[obb.range_index])
    cross = normalise(axes[obb.cross_range_index])
    oop = normalise(axes[obb.out_of_plane_index])
    if np.dot(oop, up) < 0.0:
        oop, cross = -oop, -cross
    rot = np.stack([rng, cross, oop], axis=0)
    ground = np.array([np.sqrt(1.0 - np.dot(rng, up) ** 2), np.sqrt(1.0 - np.dot(cross, up) ** 2),
                       np.dot(oop, up)]) / (2.0 * pi)
    return rot.T, obb.spans / (2.0 * pi), ground


def srp_range_change(old_srps: NDArray, new_srp: NDArray, trans_posns: NDArray,
                     rec_posns: NDArray) -> NDArray:
    """Per pulse, the bistatic range to `new_srp` minus that to the pulse's old
    SRP (utils/geodesy.py:64-66)."""
    new_srp = np.asarray(new_srp, dtype=np.float64).reshape(1, 3)
    new = (np.linalg.norm(trans_posns - new_srp, axis=-1)
           + np.linalg.norm(rec_posns - new_srp, axis=-1))
    old = (np.linalg.norm(trans_posns - old_srps, axis=-1)
           + np.linalg.norm(rec_posns - old_srps, axis=-1))
    return new - old


def move_srps(phase_history: NDArray, old_srps: NDArray, new_srp: NDArray, trans_posns: NDArray,
              rec_posns: NDArray, centre_freq: float, sample_bandwidth: float,
              sign_factor: float = 1.0, c_eff: float = 299792458.0) -> NDArray:
    """Rephase a centred-order phase history [P, N] to a new SRP: each sample is
    multiplied by exp(2 pi i dr (f_bb - sign fc) / c), dr from
    `srp_range_change`, f_bb the sample's baseband frequency
    (utils/geodesy.py:33-73)."""
    n = phase_history.shape[1]
    dr = srp_range_change(old_srps, new_srp, trans_posns, rec_posns)[:, None]
    f_bb = np.fft.fftshift(np.fft.fftfreq(n, 1.0 / sample_bandwidth))[None, :]
    return phase_history * np.exp((2.0j * pi / c_eff) * dr * (f_bb - sign_factor * centre_freq))

