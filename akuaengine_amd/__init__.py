"""akuaengine_amd -- an MI355X-native Position-Based Fluids solver.

The hot path (spatial hash, radix sort, cell table, neighbour lists, density/lambda and
position-correction iterations, velocity update with wall damping, vorticity confinement, XSPH)
runs as hand-written gfx950 HIP kernels in libakua.so behind the C ABI declared in
include/akua.h.  This package is the Python side of that ABI: the parameter structs
(``PBFConfig``, ``LambdaCorrParams``), the ``PBFSolver`` class and the dam-break scene generators.
"""
from .config import (DAM_BREAK_DT, LambdaCorrParams, PBFConfig, dam_break_params,  # noqa: F401
                     with_iterations)
from . import scenes  # noqa: F401


def __getattr__(name):
    # The solver classes need libakua.so; import them lazily so the parameter structs and the
    # scene generators work on a machine without the built library.
    if name == "PBFSolver":
        from .solver import PBFSolver
        return PBFSolver
    raise AttributeError(name)
