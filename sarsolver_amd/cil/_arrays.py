"""Array initialisation of the CIL geometries, without CIL.

`initial_array` is what `GenericCilGeometry.allocate` fills a new container
with (reference: src/sarsolver/cil/base.py:44-89, CIL's allocate extended to
complex arrays).  The rules a CIL caller relies on:
  value None          zeros (no fill);
  "random"            uniform [0, 1) from numpy's global generator
                      (np.random.seed(seed) first when seed is given); complex
                      arrays take a second draw for the imaginary part;
  "random_int"        integers in [0, max_value) from the global generator
                      (int32 draw), cast to the geometry's own dtype;
  anything else       broadcast into the array (a number, an array, or a
                      container with `as_array()`).
The fill kinds are CIL's FillType members, compared by their string value,
so this module does not import CIL.
"""
from __future__ import annotations

from typing import Optional, Sequence

import numpy as np


def fill_kind(value) -> str:
    """'none', 'random', 'random_int' or 'value'."""
    if value is None:
        return "none"
    if isinstance(value, (str, bytes)) or hasattr(value, "value"):
        name = getattr(value, "value", value)
        if isinstance(name, bytes):
            name = name.decode()
        if isinstance(name, str) and name.lower() in ("random", "random_int"):
            return name.lower()
    return "value"


def initial_array(shape: Sequence[int], dtype, value=0.0, seed: Optional[int] = None,
                  max_value: int = 100, geometry_dtype=None) -> np.ndarray:
    shape = tuple(int(s) for s in shape)
    out = np.zeros(shape, dtype=dtype)
    kind = fill_kind(value)
    if kind == "none":
        return out
    if kind in ("random", "random_int") and seed is not None:
        np.random.seed(seed)
    if kind == "random":
        draw = np.random.random_sample(shape)
        if np.iscomplexobj(out):
            draw = draw + 1j * np.random.random_sample(shape)
        out[...] = draw
    elif kind == "random_int":
        draw = np.random.randint(max_value, size=shape, dtype=np.int32)
        out[...] = np.asarray(draw, dtype=geometry_dtype if geometry_dtype is not None else dtype)
    else:
        src = value.as_array() if hasattr(value, "as_array") else value
        out[...] = src
    return out
