"""MI355X-native explicit finite-volume 2-D Euler solver (the solve path of
GPUEularSolver2D): hand-written HIP kernels for gfx950 behind a C ABI
(include/fvm_abi.h, libfvm_hip.so), with a Python binding and a synthetic
NACA 0012 mesh generator."""
from .meshgen import Mesh, naca0012_mesh, read_dat, write_dat  # noqa: F401
from .solver import (Batch, FvmError, Solver, flow_params, load_library,  # noqa: F401
                     mesh_topology, scheme_opts)

__all__ = ["Mesh", "naca0012_mesh", "read_dat", "write_dat", "Batch", "FvmError", "Solver",
           "flow_params", "load_library", "mesh_topology", "scheme_opts"]
