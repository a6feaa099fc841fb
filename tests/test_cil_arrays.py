"""The CIL wrappers' array logic without CIL (CIL is not installed here, so the
container classes themselves are not exercised: their behaviour against CIL
is unpinned).  `initial_array` follows the allocate rules of the reference's
GenericCilGeometry (src/sarsolver/cil/base.py:44-89); the operator's shape
handling is checked on the GPU (tests/test_gpu_cil_arrays.py)."""
import enum

import numpy as np
import pytest

from sarsolver_amd.cil import fill_kind, initial_array


class _Fill(enum.Enum):  # stands for CIL's FillType (string-valued members)
    RANDOM = "random"
    RANDOM_INT = "random_int"


def test_fill_kinds():
    assert fill_kind(None) == "none"
    assert fill_kind(0.0) == "value" and fill_kind(np.ones(3)) == "value"
    assert fill_kind(_Fill.RANDOM) == "random" and fill_kind("random_int") == "random_int"
    assert fill_kind(_Fill.RANDOM_INT) == "random_int"


def test_zero_none_and_constant():
    a = initial_array((1, 3, 4), np.complex128)
    assert a.shape == (1, 3, 4) and a.dtype == np.complex128 and not a.any()
    assert not initial_array([2, 5], np.complex128, None).any()
    c = initial_array([2, 5], np.complex128, 2.5 - 1j)
    assert np.all(c == 2.5 - 1j)
    r = initial_array((2, 2), np.float32, np.arange(4).reshape(2, 2))
    assert r.dtype == np.float32 and np.array_equal(r, [[0, 1], [2, 3]])


def test_random_follows_the_seeded_global_generator():
    a = initial_array((3, 4), np.complex128, _Fill.RANDOM, seed=5)
    b = initial_array((3, 4), np.complex128, _Fill.RANDOM, seed=5)
    assert np.array_equal(a, b)
    # real and imaginary parts are two successive uniform draws in [0, 1)
    assert a.real.min() >= 0 and a.real.max() < 1 and a.imag.min() >= 0 and a.imag.max() < 1
    assert not np.array_equal(a.real, a.imag)
    r = initial_array((3, 4), np.float64, _Fill.RANDOM, seed=5)
    assert np.array_equal(r, a.real)  # same first draw
    assert not np.array_equal(initial_array((3, 4), np.complex128, _Fill.RANDOM, seed=6), a)


def test_random_int_range_and_geometry_dtype():
    a = initial_array((50, 40), np.complex128, _Fill.RANDOM_INT, seed=1, max_value=7,
                      geometry_dtype=np.complex128)
    assert a.dtype == np.complex128 and not a.imag.any()
    v = a.real
    assert v.min() >= 0 and v.max() <= 6 and np.all(v == np.round(v)) and len(np.unique(v)) == 7


def test_container_value_is_copied():
    class Box:
        def as_array(self):
            return np.full((2, 3), 4.0)
    assert np.all(initial_array((2, 3), np.complex128, Box()) == 4.0)
