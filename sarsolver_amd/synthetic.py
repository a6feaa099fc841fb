"""Synthetic spotlight-SAR workloads (circular flight path, grid scene).

Geometry: monostatic (Tx == Rx) platform on a circle of radius ``radius`` at
height ``height`` around the scene centre (the origin), pulses spread
uniformly over ``aperture_deg`` centred on azimuth ``start_deg`` (default 0);
SRP at the origin for every pulse; unit waveform spectrum and slow-time
weights; c = 299 792 458 m/s.  Grids are GridSarScene([1, n, n],
[0, s (n-1), s (n-1)]) centred on the origin.

Named workloads:
    c1_small      128^2 px at 0.25 m, 256 x 512, 2 deg, U = 8
    c2_large      1024^2 px at 0.10 m, 2048 x 4096, 4 deg, U = 8
    c3_precision  2048^2 px at 0.05 m, 4096 x 8192, 8 deg, U = 8
    c4_scene      512^2 px, 1024 x 2048, 3 deg, U = 8 (pixel spacing 0.1 m)
    c5_scene      512^2 px, 1024 x 2048, 3 deg, U = 8 (pixel spacing 0.1 m)
The pixel spacing of c4/c5 is not stated by the workload description; 0.1 m
(the c2 resolution cell) is used.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .numpy import GridSarScene, NumpySimpleSarAperture

C_LIGHT = 299792458.0


@dataclass(frozen=True)
class Workload:
    name: str
    n: int            # image side (pixels)
    spacing: float    # metres
    pulses: int
    samples: int
    aperture_deg: float
    upsample: float = 8.0
    fc: float = 9.6e9
    fs: float = 600e6
    radius: float = 10e3
    height: float = 5e3
    seed: int = 1000

    @property
    def pixel_pulses(self) -> int:
        return self.n * self.n * self.pulses


WORKLOADS = {
    "c1_small": Workload("c1_small", 128, 0.25, 256, 512, 2.0, seed=1001),
    "c2_large": Workload("c2_large", 1024, 0.10, 2048, 4096, 4.0, seed=1002),
    "c3_precision": Workload("c3_precision", 2048, 0.05, 4096, 8192, 8.0, seed=1003),
    "c4_scene": Workload("c4_scene", 512, 0.10, 1024, 2048, 3.0, seed=1004),
    "c5_scene": Workload("c5_scene", 512, 0.10, 1024, 2048, 3.0, seed=1005),
}


def circular_aperture(pulses: int, samples: int, aperture_deg: float, *, fc: float = 9.6e9,
                      fs: float = 600e6, radius: float = 10e3, height: float = 5e3,
                      start_deg: float = 0.0) -> NumpySimpleSarAperture:
    theta = np.deg2rad(start_deg + aperture_deg * (np.arange(pulses) / max(pulses - 1, 1) - 0.5))
    tx = np.stack([radius * np.cos(theta), radius * np.sin(theta),
                   np.full(pulses, float(height))], axis=1)
    return NumpySimpleSarAperture(tx, tx.copy(), np.zeros((pulses, 3)), fc, fs,
                                  num_freqs=samples, c_eff=C_LIGHT)


def grid_scene(n: int, spacing: float) -> GridSarScene:
    side = spacing * (n - 1)
    return GridSarScene([1, n, n], [0.0, side, side])


def build(w: Workload, start_deg: float = 0.0):
    """(aperture, scene) of a workload."""
    ap = circular_aperture(w.pulses, w.samples, w.aperture_deg, fc=w.fc, fs=w.fs,
                           radius=w.radius, height=w.height, start_deg=start_deg)
    return ap, grid_scene(w.n, w.spacing)


def workload_plan(w: Workload, ap: NumpySimpleSarAperture, pos, device, interp: str = "linear",
                  precision: str = "series"):
    """The device plan of a workload's aperture and pixel positions (as the
    benchmark builds it: linear interpolation and series ranges by default)."""
    from ._native import SarPlan
    return SarPlan(ap.transmitter_positions, ap.receiver_positions, ap.stab_ref_posns, pos,
                   np.fft.ifftshift(ap.waveform_fft), ap.slow_time_weighting,
                   ap.centre_frequency, ap.sample_bandwidth, ap.c_eff, w.upsample,
                   ap.sign_multiplier, interp=interp, precision=precision, device=device)


def dense_image(w: Workload, m: int) -> np.ndarray:
    """The dense complex Gaussian image the benchmark's forward is timed on
    (every pixel non-zero), seeded from the workload's seed."""
    g = np.random.default_rng(w.seed + 7)
    return (g.standard_normal(m) + 1j * g.standard_normal(m)) / np.sqrt(2)


def complex_normal(rng: np.random.Generator, shape) -> np.ndarray:
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)) / np.sqrt(2.0)


def point_scatterers(rng: np.random.Generator, n: int, count: int) -> np.ndarray:
    """n x n complex image with ``count`` CN(0,1) scatterers at distinct random pixels."""
    x = np.zeros(n * n, dtype=np.complex128)
    idx = rng.choice(n * n, size=count, replace=False)
    x[idx] = complex_normal(rng, count)
    return x.reshape(n, n)


def clutter_patch(rng: np.random.Generator, n: int, size: int) -> np.ndarray:
    """n x n image with a size x size patch of CN(0,1) clutter (Rayleigh magnitude,
    uniform phase) at a random location."""
    x = np.zeros((n, n), dtype=np.complex128)
    r0, c0 = rng.integers(0, n - size + 1, size=2)
    x[r0:r0 + size, c0:c0 + size] = complex_normal(rng, (size, size))
    return x


def smooth_field(rng: np.random.Generator, n: int, sigma_px: float = 3.0) -> np.ndarray:
    """CN(0,1) field convolved with a Gaussian of ``sigma_px`` pixels (periodic)."""
    f = complex_normal(rng, (n, n))
    k = np.fft.fftfreq(n)
    g = np.exp(-2.0 * (np.pi * sigma_px) ** 2 * (k[:, None] ** 2 + k[None, :] ** 2))
    out = np.fft.ifft2(np.fft.fft2(f) * g)
    return out / np.sqrt(np.mean(np.abs(out) ** 2))


def scene_amplitudes(w: Workload, rng: np.random.Generator) -> np.ndarray:
    """The workload's ground-truth image [n, n]."""
    if w.name == "c1_small":
        return point_scatterers(rng, w.n, 20)
    if w.name == "c2_large":
        return point_scatterers(rng, w.n, 500) + clutter_patch(rng, w.n, 64)
    if w.name == "c3_precision":
        return point_scatterers(rng, w.n, 2000)
    return smooth_field(rng, w.n, 3.0) + point_scatterers(rng, w.n, 50)


def add_noise(rng: np.random.Generator, y: np.ndarray, snr_db: float) -> np.ndarray:
    """y + CN noise with variance mean|y|^2 / 10^(snr/10)."""
    var = np.mean(np.abs(y) ** 2) / 10.0 ** (snr_db / 10.0)
    return y + np.sqrt(var) * complex_normal(rng, y.shape)
