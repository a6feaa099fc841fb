"""One channel of a CPHD 1.0.x file (reference: io/cphd101.py:13-285)."""
from ._dataset import _CphdDataset


class Cphd101Dataset(_CphdDataset):
    """One channel of a CPHD 1.0.x file; (f0, df) = (SC0, SCSS) of the first
    vector, sign from Global/SGN."""
    _F0, _DF, _VERSION = "SC0", "SCSS", "1."
