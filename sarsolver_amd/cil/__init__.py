"""CIL wrappers (reference: src/sarsolver/cil).

The geometry, container and operator classes need CIL (``cil``), as the
reference's do; without it they raise ImportError when first used.  The
array-level parts (`SarArrayOperator`, `initial_array`) do not need CIL.
"""
from ._arrays import fill_kind, initial_array
from ._core import SarArrayOperator

_CIL_NAMES = {
    "GenericCilGeometry": "geometry", "ExtendedDataContainer": "geometry",
    "CilSarGridGeometry": "geometry", "SimpleCilSarGeometry": "geometry",
    "CpuSarOperator": "operator", "GpuSarOperator": "operator",
}


def __getattr__(name):
    mod = _CIL_NAMES.get(name)
    if mod is None:
        raise AttributeError(name)
    import importlib
    try:
        m = importlib.import_module(f".{mod}", __name__)
    except ImportError as e:
        raise ImportError(f"sarsolver_amd.cil.{name} needs CIL (the 'cil' package): {e}") from e
    return getattr(m, name)


__all__ = ["SarArrayOperator", "fill_kind", "initial_array", *_CIL_NAMES]
