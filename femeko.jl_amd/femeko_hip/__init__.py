"""femeko_hip: MI355X-native (gfx950) P1 magnetostatics hot path of Femeko.jl.

Host-side mirror of the reference interface over the C ABI in
include/femeko_hip.h (libfemeko_hip.so). See fem3d for the Femeko-named
drop-ins, device for the handles, batch for batched solves, adjoint for the
field-target gradient, mesh/workloads for the synthetic meshes.
"""
from ._lib import FemekoHIPError, LIB_PATH  # noqa: F401
from .device import Context, DeviceMesh, SolveReport, context  # noqa: F401
from .fem3d import (BoundaryIntegral, device_mesh, invalidate, lagrange, lineSearch,  # noqa: F401
                    localStiffnessMatrix, magneticField, scalarPotential, solveDirichlet, solveSaddle,
                    stiffnessMatrix,
                    tangentialStiffnessMatrix)
from .mesh import MESH, MU0, kuhn_box_mesh  # noqa: F401
