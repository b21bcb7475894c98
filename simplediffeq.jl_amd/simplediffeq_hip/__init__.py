"""simplediffeq_hip: MI355X ensemble integrator for SimpleDiffEq.jl's
GPU-compatible out-of-place solvers (GPUSimpleTsit5, GPUSimpleATsit5,
GPUSimpleVern7, GPUSimpleAVern7, GPUSimpleVern9, GPUSimpleAVern9, GPUSimpleRK4,
GPUSimpleEuler).

The compute path is libsimplediffeq_hip.so (HIP kernels for gfx950 behind a C
ABI); this package is its Python host side.
"""
from . import _abi  # noqa: F401
from ._abi import load_library  # noqa: F401
from .problem import (DtLessThanMin, EnsembleHIP, EnsembleProblem, EnsembleSolution,  # noqa: F401
                      GPUSimpleATsit5, GPUSimpleAVern7, GPUSimpleAVern9, GPUSimpleEuler,
                      GPUSimpleRK4, GPUSimpleTsit5, GPUSimpleVern7, GPUSimpleVern9, ODEProblem, ODESolution, isadaptive,
                      remake, rhs, solve)
