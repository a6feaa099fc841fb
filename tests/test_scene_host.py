"""Pixel positions and basebanding on the host side (CPU only): the grid
generator's arithmetic against numpy itself, the grid descriptions the device
generator accepts, and generate_simple_basebanding_values pinned by what it is
for (moving a point target's image spectrum to zero frequency)."""
from decimal import Decimal

import numpy as np
import pytest


def _axis_steps(i, n, L):
    """The steps scene_kernels.hip's grid_axis takes, in FP64 numpy scalars."""
    f = np.float64
    start, stop = f(-0.5) * f(L), f(0.5) * f(L)
    delta = stop - start
    if n > 1:
        div = f(n - 1)
        step = delta / div
        y = (f(i) / div) * delta if step == 0.0 else f(i) * step
    else:
        y = f(i) * delta
    y = y + start
    if n > 1 and i == n - 1:
        y = stop
    return y


@pytest.mark.parametrize("n,L", [(1, 0.0), (1, 3.0), (2, 1.0), (5, 5e-324), (7, 0.3),
                                 (241, 119.99999), (1024, 102.3), (1000, 1e-300), (33, 0.0)])
def test_grid_axis_steps_equal_numpy_linspace(n, L):
    """numpy's own linspace (the reference's call, numpy/hypothesis.py:199-201)
    against the kernel's steps, bit for bit, including n = 1, zero length and
    a step that underflows to zero (numpy's denormal branch)."""
    want = np.linspace(-0.5 * L, 0.5 * L, n)
    got = np.array([_axis_steps(i, n, L) for i in range(n)])
    assert np.array_equal(got.view(np.int64), want.view(np.int64)) or np.array_equal(got, want)


def _fma(a, b, c):
    return float(Decimal(float(a)) * Decimal(float(b)) + Decimal(float(c)))


def test_rotation_rounding_is_an_fma_chain_on_this_host():
    """The device evaluates R @ local as fma(R2, z, fma(R1, y, R0 x)); that is
    how numpy's matmul (OpenBLAS FMA kernels) rounds it here."""
    from sarsolver_amd.numpy import GridSarScene
    sc = GridSarScene.from_geodetic_enu([1, 40, 50], [0.0, 39.0, 49.0],
                                        np.array([51.5, -1.2, 80.0]))
    pos = sc.positions
    R = np.asarray(sc.rotation_matrix)
    c = np.asarray(sc.grid_centre)
    xs = np.linspace(-24.5, 24.5, 50)
    ys = np.linspace(-19.5, 19.5, 40)
    for m in range(0, 2000, 7):
        lx, ly, lz = xs[m % 50], ys[m // 50], 0.0
        for r in range(3):
            v = _fma(R[r, 2], lz, _fma(R[r, 1], ly, float(R[r, 0]) * float(lx)))
            assert v + c[r] == pos[m, r]


def test_grid_args_accepts_exact_grids_only():
    from sarsolver_amd._native import grid_args
    from sarsolver_amd.numpy import GridSarScene, SimpleSarScene
    g = grid_args(GridSarScene((1, 5, 7), (0.0, 2, 3.5)))
    assert g is not None and g.num_points == 35 and list(g.shape) == [1, 5, 7]
    assert grid_args(GridSarScene((1, 5, 7), (0.0, np.float32(2.0), 3.0))) is None
    assert grid_args(GridSarScene((1, 5, 7), (0.0, np.float64(2.0), 3.0))) is not None
    assert grid_args(SimpleSarScene(np.zeros((4, 3)))) is None
    assert grid_args(np.zeros((4, 3))) is None
    a = grid_args(GridSarScene((1, 5, 7), (0.0, 2.0, 3.0), grid_centre=np.array([1.0, 2, 3])))
    b = grid_args(GridSarScene((1, 5, 7), (0.0, 2.0, 3.0), grid_centre=np.array([1.0, 2, 3.5])))
    assert a.key != b.key


def test_basebanding_moves_the_image_spectrum_to_zero(oracle_lib):
    """Physics pin of utils/k_space.py:109-134: the oracle backprojection of a
    point target at the spot has its 2-D spectrum centred at the k-space
    carrier (56.7 cycles/m here); times the basebanding phases it is centred
    at zero; with the conjugate phases (sign flipped) it moves the other way."""
    from sarsolver_amd import synthetic
    from sarsolver_amd.numpy import GridSarScene
    from sarsolver_amd.utils import generate_simple_basebanding_values
    ap = synthetic.circular_aperture(128, 128, 2.0)
    n, h = 64, 0.005
    centre = np.array([0.3, -0.2, 0.0])
    sc = GridSarScene((1, n, n), (0.0, h * (n - 1), h * (n - 1)), grid_centre=centre)
    tx = ap.transmitter_positions
    ones = np.ones(128, complex)
    Y = oracle_lib.forward(tx, tx, ap.stab_ref_posns, centre[None], np.ones(1, complex), ones,
                           ones, ap.centre_frequency, ap.sample_bandwidth, 128,
                           upsample_ratio=8.0, interp="linear")
    img = oracle_lib.adjoint(tx, tx, ap.stab_ref_posns, sc.positions, Y, ones, ones,
                             ap.centre_frequency, ap.sample_bandwidth, upsample_ratio=8.0,
                             interp="linear").reshape(n, n)

    def centroid(a):
        p = np.abs(np.fft.fftshift(np.fft.fft2(a))) ** 2
        f = np.fft.fftshift(np.fft.fftfreq(n, h))
        fy, fx = np.meshgrid(f, f, indexing="ij")
        return np.array([(p * fx).sum(), (p * fy).sum()]) / p.sum()

    b = generate_simple_basebanding_values(sc, ap).reshape(n, n)
    raw, based, flipped = centroid(img), centroid(img * b), centroid(img * np.conj(b))
    print(f"spectrum centroid (cycles/m): raw {raw}, basebanded {based}, flipped {flipped}")
    assert np.allclose(np.abs(b), 1.0, atol=1e-15)
    assert np.linalg.norm(raw) > 40.0
    assert np.linalg.norm(based) < 0.5
    assert np.linalg.norm(flipped) > 40.0
    # spot defaults to the mean position; an explicit spot at a pixel gives 1 there
    b2 = generate_simple_basebanding_values(sc, ap, basebanding_spot=sc.positions[100])
    assert abs(b2[100] - 1.0) < 1e-15
