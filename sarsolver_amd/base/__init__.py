from .hypothesis import block_hypothesis_vector, BaseSarScene
from .measurement import BaseSimpleSarAperture, BaseSimpleSarDataset, block_phase_history_array
