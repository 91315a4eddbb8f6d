"""MI355X-native drop-in for DiffEqGPU's EnsembleGPUKernel path (GPUTsit5, GPUVern7, GPUVern9,
GPURosenbrock23, GPURodas4, GPURodas5P, GPUKvaerno3, GPUKvaerno5; SDEs: GPUEM, GPUSIEA), over the C ABI
of libdeghip.so (include/deghip.h).

Import under any module name (the directory name is not a Python identifier); the test
suite, the benchmark and the build entry point load it as ``diffeqgpu_amd``.
"""
from .algorithms import (EnsembleGPUKernel, EnsembleKernelAlgorithm, GPUEM, GPUODEAlgorithm, GPUODEImplicitAlgorithm,
                         GPUKvaerno3, GPUKvaerno5, GPURodas4, GPURodas5P, GPURosenbrock23, GPUSDEAlgorithm, GPUSIEA,
                         GPUTsit5, GPUVern7, GPUVern9, HIPDevice, alg_order)
from .dual import Dual, DualArray
from .callbacks import (CallbackSet, ContinuousCallback, DiscreteCallback, GPUContinuousCallback, GPUDiscreteCallback,
                        LeftRootFind, NoRootFind, RightRootFind, terminate, u_modified)
from .ensemble import EnsembleContext, EnsembleSolution, ODESolution, solve
from .lowerlevel import SolveResult, order_by_span, vectorized_asolve, vectorized_solve
from .problems import (DualODEFunction, EnsembleProblem, HIPProbVector, ImmutableODEProblem, ImmutableSDEProblem,
                       ODEFunction, ODEProblem, SDEFunction, SDEProblem, adapt, get_backend, is_diagonal_noise, make_prob_compatible, remake)
from .ranges import colon, linrange
from .sweeps import Draw, ParameterSweep
from . import autotune, costmodel, learned
from .tracer import SVector
from . import tracer as math  # sin, cos, exp, log, sqrt, tanh, fabs, pow, ifelse on traced values

__all__ = [
    "EnsembleGPUKernel", "EnsembleKernelAlgorithm", "GPUODEAlgorithm", "GPURosenbrock23", "GPUTsit5", "GPUVern7",
    "GPUVern9", "GPUODEImplicitAlgorithm", "GPURodas4", "GPURodas5P", "GPUKvaerno3", "GPUKvaerno5",
    "HIPDevice", "alg_order", "EnsembleContext", "EnsembleSolution", "ODESolution", "solve", "SolveResult", "order_by_span",
    "vectorized_asolve", "vectorized_solve", "EnsembleProblem", "HIPProbVector", "ImmutableODEProblem",
    "ODEFunction", "ODEProblem", "adapt", "get_backend", "make_prob_compatible", "remake", "colon", "linrange",
    "SVector", "math", "Draw", "ParameterSweep", "autotune", "costmodel", "learned", "CallbackSet", "ContinuousCallback",
    "DiscreteCallback", "GPUContinuousCallback", "GPUDiscreteCallback", "LeftRootFind", "NoRootFind", "RightRootFind",
    "terminate", "u_modified", "GPUSDEAlgorithm", "GPUEM", "GPUSIEA", "SDEFunction", "SDEProblem",
    "ImmutableSDEProblem", "is_diagonal_noise", "Dual", "DualArray", "DualODEFunction",
]
