"""CPHD input (reference: src/sarsolver/io/): readers for CPHD 1.0.x and 0.3
without sarpy."""
from .cphd import CphdFile, open_cphd
from .cphd03 import Cphd03Dataset
from .cphd101 import Cphd101Dataset

__all__ = ["CphdFile", "open_cphd", "Cphd101Dataset", "Cphd03Dataset"]
