from .geodesy import scene_params_enu, scene_params_classic_sar
from .k_space import OrientedBoundingBox
from .misc import default_num_threads
