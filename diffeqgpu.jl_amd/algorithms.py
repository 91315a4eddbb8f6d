"""Algorithm selectors and the ensemble algorithm.

Mirrors src/ensemblegpukernel/gpukernel_algorithms.jl (GPUTsit5 :30, GPUVern7 :57,
GPURosenbrock23 :112 with the ``autodiff`` keyword :214-220, GPUEM :243, GPUSIEA :266), alg_utils.jl:1-16
(``alg_order``) and src/algorithms.jl:192-211 (``EnsembleGPUKernel(dev, cpu_offload)``).
"""
from __future__ import annotations

from dataclasses import dataclass

from . import _lib


class GPUODEAlgorithm:
    method: int = -1
    order: int = 0

    def __repr__(self):
        return f"{type(self).__name__}()"

    def __eq__(self, other):
        return type(self) is type(other) and vars(self) == vars(other)

    def __hash__(self):
        return hash((type(self).__name__, tuple(sorted(vars(self).items()))))


class GPUTsit5(GPUODEAlgorithm):
    """Tsitouras 5(4) with FSAL (gpukernel_algorithms.jl:30)."""

    method = _lib.DEG_TSIT5
    order = 5


class GPUVern7(GPUODEAlgorithm):
    """Verner 7(6), not FSAL (gpukernel_algorithms.jl:57)."""

    method = _lib.DEG_VERN7
    order = 7


class GPUVern9(GPUODEAlgorithm):
    """Verner 9(8), not FSAL (gpukernel_algorithms.jl:83); dense output with ten extra stages."""

    method = _lib.DEG_VERN9
    order = 9


class GPUODEImplicitAlgorithm(GPUODEAlgorithm):
    """The methods that use the Jacobian (gpukernel_algorithms.jl: ``GPUODEImplicitAlgorithm{AD}``).
    The problem's ODEFunction ``jac``/``tgrad`` are used when given; otherwise ``autodiff=True``
    derives them by forward-mode differentiation of the traced RHS and ``autodiff=False`` by
    forward differences with step sqrt(eps(Float32)) (alg_utils.jl:17-27, nlsolve/type.jl:129-157)."""

    def __init__(self, autodiff: bool = True):
        self.autodiff = bool(autodiff)

    def __repr__(self):
        return f"{type(self).__name__}(autodiff={self.autodiff})"


class GPURosenbrock23(GPUODEImplicitAlgorithm):
    """Rosenbrock 2(3) (gpukernel_algorithms.jl:112, :214-220)."""

    method = _lib.DEG_RB23
    order = 2


class GPURodas4(GPUODEImplicitAlgorithm):
    """Rodas4, 6 stages, order 4 with a stiffly accurate embedded error (gpukernel_algorithms.jl:137)."""

    method = _lib.DEG_RODAS4
    order = 4


class GPURodas5P(GPUODEImplicitAlgorithm):
    """Rodas5P, 8 stages, order 5 (gpukernel_algorithms.jl:162)."""

    method = _lib.DEG_RODAS5P
    order = 5


class GPUKvaerno3(GPUODEImplicitAlgorithm):
    """Kvaerno3, ESDIRK order 3 with Newton-solved stages (gpukernel_algorithms.jl:187)."""

    method = _lib.DEG_KVAERNO3
    order = 3


class GPUKvaerno5(GPUODEImplicitAlgorithm):
    """Kvaerno5, ESDIRK order 5 with Newton-solved stages (gpukernel_algorithms.jl:212)."""

    method = _lib.DEG_KVAERNO5
    order = 5


class GPUSDEAlgorithm:
    """The fixed-step SDE methods of EnsembleGPUKernel (src/DiffEqGPU.jl:138-150)."""

    method: int = -1
    order: int = 0

    def __repr__(self):
        return f"{type(self).__name__}()"

    def __eq__(self, other):
        return type(self) is type(other)

    def __hash__(self):
        return hash(type(self).__name__)


class GPUEM(GPUSDEAlgorithm):
    """Euler-Maruyama, weak order 1, diagonal and general noise (gpukernel_algorithms.jl:243)."""

    method = _lib.DEG_EM
    order = 1


class GPUSIEA(GPUSDEAlgorithm):
    """SIEA, weak order 2 for Ito SDEs with diagonal noise (gpukernel_algorithms.jl:266)."""

    method = _lib.DEG_SIEA
    order = 2


def alg_order(alg) -> int:
    """alg_utils.jl:1-16."""
    if not isinstance(alg, (GPUODEAlgorithm, GPUSDEAlgorithm)) or alg.order == 0:
        raise ValueError("Order is not defined for this algorithm")
    return alg.order


@dataclass(frozen=True)
class HIPDevice:
    """The backend object handed to ``EnsembleGPUKernel`` (the role of ``ROCBackend()``).

    Knobs (all optional):
      device      HIP device ordinal.
      contract    "off": no FP contraction (the reference's unfused Float32 sequence; bitwise
                  comparable with the CPU restatement). "fast": FMA contraction. None: "off" for
                  fixed-step solves, "fast" for adaptive solves.
      powmode     controller powers: "fast" (default: the PI controller evaluated in log2 space with
                  hardware log2/exp2 on the squared error norm), "ocml" (device powf, the same
                  library call AMDGPU.jl makes for Float32 ^, with the reference's operation
                  sequence) or "det" (deterministic polynomial, bitwise comparable with the oracle).
      persistent  persistent grid with a device work counter (lanes refill as trajectories end);
                  None picks it for adaptive solves.
      kahan       compensated accumulation of t.
      block_size  threads per block (64..256); None: 256 or the autotuned value.
      max_attempts per-trajectory attempt cap (0: 10^7). The reference has none and loops forever
                  when t + dt rounds to t; here such a lane ends with retcode MaxIters.
      refill      persistent grids: a wave takes new trajectories once this many of its lanes
                  are idle (1..64; 0: 4).
      chunk       persistent grids: trajectory ids a wave claims per queue atomic (0: 16).
      waves       register cap: compile the solver for at least this many resident waves per
                  SIMD (1..7; 0: the compiler's choice).
      blocks_per_cu persistent grids: launch this many blocks per compute unit (0: as many as
                  fit). Fewer resident lanes means more trajectories per lane, which shortens
                  the tail where lanes sit idle.
      batch_saves adaptive GPUTsit5 (n <= 4) with saveat: stash steps that cross save points in
                  LDS and interpolate for many lanes at once (same results).
      compact     persistent adaptive grids: once the queue is empty, waves with few
                  trajectories left move them to other waves of the workgroup (through LDS)
                  and leave, so fewer, fuller waves finish the tail.
      start_order persistent adaptive grids: "cost" starts the trajectories in descending order of
                  span / initial step estimate (deg_start_order, on the device before every solve),
                  "model" in descending order of the attempt count ``cost_model`` predicts
                  (deg_start_order_model), "none" in index order; None: "none" (the autotuner
                  measures them). An explicit ``order=`` overrides it. Results do not depend on the order.
      cost_model  the 3 + n + np weights of a start-order cost model (costmodel.fit), for
                  start_order="model".
      ordered_inputs with a cost or model start order: the ordering pass also copies u0 and p into
                  the order in which the queues hand the trajectories out, so lanes read their
                  inputs contiguously instead of gathering them (deg_solve_args.inputs_ordered;
                  same results, (n + np) floats per trajectory of device memory). It cuts the
                  solver's input reads to about their algorithmic size but costs the scatter about
                  as much time as the gathers it removes, so it is off by default.
    """

    device: int = 0
    contract: str | None = None
    powmode: str = "fast"
    persistent: bool | None = None
    kahan: bool = False
    block_size: int | None = None
    max_attempts: int = 0
    refill: int = 0
    chunk: int = 0
    waves: int = 0
    blocks_per_cu: int = 0
    compact: bool = False
    batch_saves: bool = False
    start_order: str | None = None
    cost_model: tuple | None = None
    ordered_inputs: bool = False

    def cost_order(self, adaptive: bool) -> bool:
        """Whether solves on this device start their trajectories in cost order (by the span / initial
        step key or by the cost model)."""
        mode = self.start_order or "none"
        if mode not in ("cost", "model", "none"):
            raise ValueError("start_order must be 'cost', 'model' or 'none'")
        if mode == "model" and not self.cost_model:
            raise ValueError("start_order='model' needs cost_model weights (costmodel.fit)")
        persistent = self.persistent if self.persistent is not None else adaptive
        return mode in ("cost", "model") and adaptive and persistent

    def sde_flags(self) -> int:
        """Compile flags of the fixed-step SDE kernels: contraction (default "off") and the
        register cap; the ODE-only knobs (persistent grids, kahan, batched saves, compaction,
        controller powers) do not apply."""
        contract = self.contract or "off"
        if contract not in ("off", "fast"):
            raise ValueError("contract must be 'off' or 'fast'")
        if self.kahan:
            raise ValueError("kahan summation of t is not provided by the SDE kernels")
        if not 0 <= self.waves <= 7:
            raise ValueError("waves must be in 0..7")
        return (_lib.F_PARITY if contract == "off" else 0) | (self.waves << _lib.F_WAVES_SHIFT)

    def flags(self, adaptive: bool) -> int:
        f = 0
        if adaptive:
            f |= _lib.F_ADAPTIVE
        contract = self.contract or ("fast" if adaptive else "off")
        if contract not in ("off", "fast"):
            raise ValueError("contract must be 'off' or 'fast'")
        if contract == "off":
            f |= _lib.F_PARITY
        if self.powmode not in ("ocml", "fast", "det"):
            raise ValueError("powmode must be 'ocml', 'fast' or 'det'")
        if adaptive and self.powmode == "det":
            f |= _lib.F_DETPOW
        elif adaptive and self.powmode == "fast":
            f |= _lib.F_FASTPOW
        if self.kahan:
            f |= _lib.F_KAHAN
        persistent = self.persistent if self.persistent is not None else adaptive
        if persistent:
            f |= _lib.F_PERSISTENT
        if self.compact and adaptive and persistent:
            f |= _lib.F_COMPACT
        if self.batch_saves and adaptive:
            f |= _lib.F_BATCH_SAVES
        if self.ordered_inputs and adaptive and persistent and (self.start_order or "none") in ("cost", "model"):
            f |= _lib.F_ORDERED_INPUTS
        if not 0 <= self.waves <= 7:
            raise ValueError("waves must be in 0..7")
        f |= self.waves << _lib.F_WAVES_SHIFT
        return f


class EnsembleKernelAlgorithm:
    """Abstract type of the kernel ensemble algorithms (src/DiffEqGPU.jl:107)."""


class EnsembleGPUKernel(EnsembleKernelAlgorithm):
    """``EnsembleGPUKernel(dev, cpu_offload = 0.0)`` (src/algorithms.jl:192-211).

    ``cpu_offload`` must be 0.0: this library has no CPU solver path (the reference offloads a
    fraction of trajectories to SimpleDiffEq on host threads, solve.jl:37-81)."""

    def __init__(self, dev: HIPDevice | None = None, cpu_offload: float = 0.0):
        self.dev = dev if dev is not None else HIPDevice()
        if cpu_offload != 0.0:
            raise NotImplementedError("cpu_offload > 0 is not provided: every trajectory runs on the GPU")
        self.cpu_offload = 0.0

    def __repr__(self):
        return f"EnsembleGPUKernel({self.dev!r}, {self.cpu_offload})"
