from .sarsolver_cxx import SarCalculationInfo, forward_evaluate, adjoint_evaluate
from .adjoint import single_adjoint_evaluate, multi_adjoint_evaluate
from .forward import single_forward_evaluate, multi_forward_evaluate
