"""High-level ``solve(EnsembleProblem, alg, EnsembleGPUKernel(dev); kwargs...)``.

Mirrors the kernel branch of DiffEqGPU's ensemble front end: src/solve.jl:1-95
(``__solve``: trajectories, batch_size, adaptive, seeding), :174-287 (``batch_solve``:
prob_func expansion, per-trajectory tspan/saveat rules, build_solution with trimming and
return codes) and :382-419 (``batch_solve_up_kernel``: adapt, dispatch to
vectorized_solve / vectorized_asolve, copy back).
"""
from __future__ import annotations

import copy
import time
from dataclasses import dataclass, field
from typing import Any, List, Optional

import numpy as np

from . import _lib
from .algorithms import EnsembleGPUKernel, GPUODEAlgorithm, GPUSDEAlgorithm
from .callbacks import merged_callback
from .dual import split_rows
from .lowerlevel import check_mass, vectorized_asolve, vectorized_solve
from .problems import EnsembleProblem, adapt, make_prob_compatible


@dataclass
class EnsembleContext:
    """What prob_func / output_func receive as ``ctx``: 1-based ``sim_id`` and an ``rng``."""

    sim_id: int
    rng: Optional[np.random.Generator] = None
    seed: Optional[int] = None


@dataclass
class ODESolution:
    t: np.ndarray
    u: np.ndarray          # (len(t), n); (len(t), n, 1 + P) for dual inputs (values, then partials)
    retcode: str
    prob: Any = None
    alg: Any = None
    stats: Any = None

    def __getitem__(self, i):
        return self.u[i]

    def __call__(self, tval):
        """Linear interpolation between saved points (SciMLBase's LinearInterpolation of a
        solution built without dense output): the first saved entry at tval if there is one,
        otherwise the line between its neighbours. No extrapolation."""
        t = np.asarray(self.t, dtype=np.float64)
        u = np.asarray(self.u, dtype=np.float64)
        x = float(tval)
        if len(t) == 0:
            raise ValueError("empty solution")
        if x > t[-1] or x < t[0]:
            raise ValueError("Solution interpolation cannot extrapolate outside the saved time span.")
        i = int(np.searchsorted(t, x, side="left"))
        if t[i] == x:
            return u[i]
        th = (x - t[i - 1]) / (t[i] - t[i - 1])
        return (1.0 - th) * u[i - 1] + th * u[i]


@dataclass
class EnsembleSolution:
    u: List[Any]
    elapsedTime: float
    converged: bool = True
    stats: dict = field(default_factory=dict)

    def __len__(self):
        return len(self.u)

    def __getitem__(self, i):
        return self.u[i]


def _ctx(i: int, seed: Optional[int]) -> EnsembleContext:
    if seed is None:
        return EnsembleContext(sim_id=i)
    s = (int(seed) * 1000003 + i) % (2 ** 63)
    return EnsembleContext(sim_id=i, rng=np.random.default_rng(s), seed=s)


def solve(ensembleprob: EnsembleProblem, alg: GPUODEAlgorithm, ensemblealg: EnsembleGPUKernel, *,
          trajectories: int, batch_size: Optional[int] = None, adaptive: bool = True, seed=None,
          unstable_check=None, **kwargs) -> EnsembleSolution:
    """solve.jl:1-95. ``adaptive`` selects vectorized_asolve (default) or vectorized_solve. The SDE
    methods (GPUEM, GPUSIEA) are fixed-step only: they need ``adaptive=False`` and ``dt``, as in the
    reference, whose adaptive SDE entry point raises."""
    if not isinstance(ensemblealg, EnsembleGPUKernel):
        raise TypeError("only EnsembleGPUKernel is provided")
    if not isinstance(alg, (GPUODEAlgorithm, GPUSDEAlgorithm)):
        raise TypeError("alg must be a GPU ODE algorithm (GPUTsit5, GPUVern7, GPUVern9, GPURosenbrock23, GPURodas4, "
                        "GPURodas5P, GPUKvaerno3, GPUKvaerno5) or a GPU SDE algorithm (GPUEM, GPUSIEA)")
    if trajectories < 1:
        raise ValueError("trajectories must be >= 1")
    batch_size = trajectories if batch_size is None else int(batch_size)
    t_start = time.perf_counter()
    sols: List[Any] = []
    for lo in range(0, trajectories, batch_size):
        I = range(lo + 1, min(lo + batch_size, trajectories) + 1)  # 1-based sim ids like the reference
        sols.extend(batch_solve(ensembleprob, alg, ensemblealg, I, adaptive, seed=seed, **kwargs))
    return EnsembleSolution(sols, time.perf_counter() - t_start, True)


def batch_solve(ensembleprob, alg, ensemblealg, I, adaptive, seed=None, **kwargs):
    """solve.jl:174-287 (kernel branch)."""
    if len(I) == 0:
        raise ValueError("empty batch")
    probs = []
    for i in I:
        base = copy.deepcopy(ensembleprob.prob) if ensembleprob.safetycopy else ensembleprob.prob
        probs.append(make_prob_compatible(ensembleprob.prob_func(base, _ctx(i, seed))))
    # per-trajectory time spans (solve.jl:206-245)
    same_tspan = all(p.tspan == probs[0].tspan for p in probs)
    inner_saveat = probs[0].kwargs.get("saveat")
    if not same_tspan:
        if inner_saveat is None and not adaptive and kwargs.get("save_everystep", True):
            raise ValueError("Using different time-spans require either turning off save_everystep or using "
                             "saveat. If using saveat, it should be of same length across the ensemble.")
        lens = {np.size(p.kwargs.get("saveat")) if p.kwargs.get("saveat") is not None else None for p in probs}
        if len(lens) > 1:
            raise ValueError("Using different saveat in EnsembleGPUKernel requires all of them to be of same "
                             "length. Use saveats of same size only.")
    saveat = inner_saveat if inner_saveat is not None else kwargs.pop("saveat", None)
    kwargs.pop("saveat", None)
    if isinstance(alg, GPUSDEAlgorithm):
        kwargs["traj_base"] = I[0] - 1  # noise streams numbered by ensemble index, whatever the batching
    solts, solus, rc = batch_solve_up_kernel(ensembleprob, probs, alg, ensemblealg, I, adaptive,
                                             saveat=saveat, **kwargs)
    out = []
    for k, i in enumerate(I):
        ts = solts[:, k]
        us = solus[:, k, :]
        nz = np.nonzero(ts != probs[k].tspan[0])[0]
        if len(nz) == 0:
            raise RuntimeError(f"Batch solve failed: no solution found for trajectory {i} (tspan[1] = "
                               f"{probs[k].tspan[0]})")
        sol_idx = int(nz[-1])
        retcode = "Terminated" if sol_idx != len(ts) - 1 else "Success"
        if rc[k] != _lib.RET_SUCCESS:
            retcode = _lib.RETCODE_NAMES.get(int(rc[k]), "Failure")
        rows = us[: sol_idx + 1].copy()
        dual_dims = getattr(probs[k], "dual", None)  # SDE problems carry none
        if dual_dims is not None:
            # dual inputs: each saved state as (n, 1 + P), values in column 0, partials after
            n, _np, P = dual_dims
            rows = split_rows(rows, n, P)
        sol = ODESolution(ts[: sol_idx + 1].copy(), rows, retcode, probs[k], alg)
        out.append(ensembleprob.output_func(sol, _ctx(i, seed))[0])
    return out


def batch_solve_up_kernel(ensembleprob, probs, alg, ensemblealg, I, adaptive, **kwargs):
    """solve.jl:382-419: adapt to the device, run the lower-level solve, copy back. The kernel's
    callbacks are the problem's ``callback`` keyword plus the solve's ``callback`` when
    ``merge_callbacks`` is set (ensemblegpuarray/callbacks.jl:61-74), converted to GPU callbacks."""
    kwargs["callback"] = merged_callback(probs[0].kwargs.get("callback"), kwargs.get("callback"),
                                         bool(kwargs.pop("merge_callbacks", False)))
    if kwargs["callback"] is None:
        kwargs.pop("callback")
    if isinstance(alg, GPUODEAlgorithm):
        check_mass(probs[0].f, alg)
        if probs[0].dual is not None and kwargs.get("callback") is not None:
            raise NotImplementedError("callbacks with dual-number inputs are not supported")
    if adaptive and isinstance(alg, GPUSDEAlgorithm):
        # the adaptive SDE entry point raises (lowerlevel_solve.jl:349-356); say so before any device work
        raise NotImplementedError("Adaptive time-stepping is not supported yet with GPUEM.")
    dev = ensemblealg.dev
    gprobs = adapt(dev, probs)
    if adaptive:
        res = vectorized_asolve(gprobs, ensembleprob.prob, alg, **kwargs)
    else:
        if "dt" not in kwargs:
            raise ValueError("a fixed-step solve needs dt")
        res = vectorized_solve(gprobs, ensembleprob.prob, alg, **kwargs)
    ts, us = res
    return ts.cpu().numpy(), us.cpu().numpy(), res.retcode.cpu().numpy()
