"""pfgpu: MI355X-native nonlinear-elasticity assembly for PolyFEM.

The compute path is libpfgpu.so (hand-written HIP kernels for gfx950 behind
the C-ABI in include/pfgpu.h); this package holds its Python binding, the
mirror of PolyFEM's assembler/form/solver API, the synthetic-mesh generator
and the .msh reader.
"""
from . import _ffi, mesh, msh
from .assembler import ElasticForm, NLAssembler, make_assembler
from ._ffi import PfgpuError
from .mesh import tet_mesh
from .msh import load_msh, read_msh
from .solver import (BatchSolver, NewtonSolver, NLProblem, TransientSimulation, cg_solve, concatenate_meshes,
                     solve_static_adjoint)

__all__ = ["ElasticForm", "NLAssembler", "make_assembler", "PfgpuError", "mesh", "_ffi", "NewtonSolver",
           "NLProblem", "TransientSimulation", "cg_solve", "solve_static_adjoint", "BatchSolver", "concatenate_meshes",
           "msh", "tet_mesh", "load_msh", "read_msh"]
