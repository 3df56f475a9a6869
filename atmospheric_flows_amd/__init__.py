"""MI355X-native IMEX-DG compressible-Euler hot path (drop-in for the reference's EULEROperator).

The compute path is libafl_hip.so (hand-written HIP kernels for gfx950) behind the C-ABI in
include/afl.h; `EulerOperator` mirrors the reference's operator interface over that ABI.
"""
from ._abi import (MeshDesc, SolverParams, StepInfo, HillState, Sponge, make_mesh, make_hill_state,
                   make_solver_params, make_sponge, n_per_field, AFL_NODES_GAUSS_LOBATTO, AFL_NODES_EQUIDISTANT)
from .euler_operator import EulerOperator, AflError, load_library, LIB_PATH

__all__ = ["MeshDesc", "SolverParams", "StepInfo", "HillState", "Sponge", "make_mesh", "make_hill_state",
           "make_solver_params", "make_sponge", "n_per_field", "EulerOperator", "AflError", "load_library",
           "LIB_PATH", "AFL_NODES_GAUSS_LOBATTO", "AFL_NODES_EQUIDISTANT"]
