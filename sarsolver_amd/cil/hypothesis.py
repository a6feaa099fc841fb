"""Import path of the reference's cil/hypothesis.py: the image geometry."""
from .geometry import CilSarGridGeometry

__all__ = ["CilSarGridGeometry"]
