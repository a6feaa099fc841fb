"""One channel of a CPHD 0.3 file (reference: io/cphd03.py:13-289)."""
from ._dataset import _CphdDataset


class Cphd03Dataset(_CphdDataset):
    """One channel of a CPHD 0.3 file; (f0, df) = (Fx0, Fx_SS) of the first
    vector, sign from Global/PhaseSGN."""
    _F0, _DF, _VERSION = "Fx0", "Fx_SS", "0.3"
