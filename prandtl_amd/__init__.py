"""MI355X-native DGSEM right-hand side and SSP-RK3 stage update.

The compute path is libprandtl_hip.so (hand-written HIP for gfx950) behind the
C ABI of include/prandtl_hip.h; this package is its host-side mirror of the
reference's operator interface (DGSEMOperator::Mult / RK3SSPSolver::Step).
"""
from ._lib import PrandtlError, lib  # noqa: F401
from .operator import DGSEMOperator, RK3SSPSolver, lgl_nodes  # noqa: F401
