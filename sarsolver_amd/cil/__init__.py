"""CIL wrappers (reference: src/sarsolver/cil).  Importing this subpackage
requires CIL (``cil``), exactly as the reference's does."""
from .base import ExtendedDataContainer, GenericCilGeometry
from .hypothesis import CilSarGridGeometry
from .measurement import SimpleCilSarGeometry
from .operator import CpuSarOperator, GpuSarOperator
