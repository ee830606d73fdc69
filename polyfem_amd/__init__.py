"""pfgpu: MI355X-native nonlinear-elasticity assembly for PolyFEM.

The compute path is libpfgpu.so (hand-written HIP kernels for gfx950 behind
the C-ABI in include/pfgpu.h); this package holds its Python binding, the
mirror of PolyFEM's assembler/form API and the synthetic-mesh generator.
"""
from . import _ffi, mesh
from .assembler import ElasticForm, NLAssembler, make_assembler
from ._ffi import PfgpuError
from .solver import (BatchSolver, NewtonSolver, NLProblem, TransientSimulation, cg_solve, concatenate_meshes,
                     solve_static_adjoint)

__all__ = ["ElasticForm", "NLAssembler", "make_assembler", "PfgpuError", "mesh", "_ffi", "NewtonSolver",
           "NLProblem", "TransientSimulation", "cg_solve", "solve_static_adjoint", "BatchSolver", "concatenate_meshes"]
