"""Import path of the reference's cil/measurement.py: the measurement geometry."""
from .geometry import SimpleCilSarGeometry

__all__ = ["SimpleCilSarGeometry"]
