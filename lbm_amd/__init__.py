"""MI355X-native Inamuro two-phase D3Q15 lattice Boltzmann time step.

The compute path is lbm_amd/liblbm_hip.so (hand-written gfx950 kernels behind
the C ABI in include/lbm_hip.h); `lbm_amd.hip.InamuroHIP` binds it.
"""
from .hip import FP64, MIXED, InamuroHIP, SolverParams  # noqa: F401
