"""Import path of the reference's cil/base.py: the generic geometry and container."""
from .geometry import ExtendedDataContainer, GenericCilGeometry

__all__ = ["ExtendedDataContainer", "GenericCilGeometry"]
