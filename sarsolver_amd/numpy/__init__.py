from .hypothesis import GridSarScene, SimpleSarScene
from .measurement import NumpySimpleSarAperture, NumpySimpleSarDataset
from .operator import NumpySimpleSarOperator
