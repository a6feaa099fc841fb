"""sarsolver_amd: MI355X (gfx950) SAR forward projection and backprojection.

Drop-in for the operator path of sarsolver: ``sarsolver_amd.cxx_binding``
(single_/multi_ forward and adjoint evaluation), ``sarsolver_amd.numpy``
(scenes, apertures, scipy LinearOperator) and ``sarsolver_amd.cil`` (CIL
operator; needs CIL).  The arithmetic runs in libsarsolver_hip.so through a C
ABI (include/sarsolver_hip.h); there is no CPU fallback.
"""
__version__ = "0.1.0"

from ._native import SarPlan, load_library  # noqa: F401
