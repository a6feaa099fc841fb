from .geodesy import move_srps, scene_params_enu, scene_params_classic_sar, srp_range_change
from .k_space import OrientedBoundingBox, generate_simple_basebanding_values
from .misc import default_num_threads
from .signal import fast_freq_downsample, get_downsample_kernel, to_db_abs
