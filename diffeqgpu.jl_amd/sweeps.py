"""Graph-captured parameter sweeps.

A sweep re-draws some rows of the ensemble's parameters (and optionally per-trajectory end
times or initial states) on the device and solves the whole ensemble again. ``ParameterSweep``
records ``sweeps`` such rounds into one HIP graph (``deg_graph_*``). A replay runs every draw and
solve with no host work between launches. Each sweep writes its own output tensors, so all
results are available after one replay.

The reference has no counterpart. There, a sweep is a fresh ``solve(EnsembleProblem, ...)``
per parameter set, with the draws in ``prob_func`` on the host (src/solve.jl:187-202).

With ``longest_first`` every sweep also orders its trajectories by descending span on the
device (``deg_order_by_span``) so that the persistent queue starts the longest ones first.

Draw j of sweep k fills its target row with ``lo + scale * U(seed + k, i * rows + j)``, where
``rows = len(draws)``. This is the same table as ``rng.uniform_table(seed + k, rows, N)``, so
sweep 0 with ``seed`` reproduces a host-drawn workload built from that table.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import _lib
from .algorithms import GPUODEAlgorithm
from .lowerlevel import vectorized_asolve, vectorized_solve
from .problems import HIPProbVector, ODEProblem


@dataclass(frozen=True)
class Draw:
    """One drawn row: ``target`` is "p", "u0" or "tfs"; ``row`` indexes p/u0 (ignored for tfs)."""

    target: str
    row: int
    lo: float
    scale: float


class ParameterSweep:
    """``sweeps`` device-side re-draws + solves of ``probs``, captured into one HIP graph.

    ``solve_kw`` are the keywords of ``vectorized_asolve`` (``adaptive=True``) or
    ``vectorized_solve`` (``adaptive=False``). Construction runs every sweep once eagerly
    (this compiles the kernel and allocates the outputs). Then it captures the graph on a
    private stream."""

    def __init__(self, probs: HIPProbVector, prob: ODEProblem, alg: GPUODEAlgorithm, *, sweeps: int, seed: int,
                 draws, adaptive: bool = True, longest_first: bool = False, **solve_kw):
        import torch

        if sweeps < 1:
            raise ValueError("sweeps must be >= 1")
        self.draws = [d if isinstance(d, Draw) else Draw(*d) for d in draws]
        for d in self.draws:
            if d.target not in ("p", "u0", "tfs"):
                raise ValueError(f"unknown draw target {d.target!r}")
            if d.target == "tfs" and probs.tfs is None:
                raise ValueError("drawing tfs needs per-trajectory end times in probs")
            if d.target in ("p", "u0"):
                rows = probs.p.shape[0] if d.target == "p" else probs.n
                if not 0 <= d.row < rows:
                    raise ValueError(f"draw row {d.row} outside {d.target}")
        self.probs, self.prob, self.alg = probs, prob, alg
        self.sweeps, self.seed, self.adaptive = sweeps, seed, adaptive
        self.solve_kw = dict(solve_kw)
        self.device = probs.u0.device
        self.stream = torch.cuda.Stream(device=self.device)
        self.order = self.work = None
        if longest_first:
            # re-ordered on the device after every draw: longest spans start first
            self.order = torch.empty(probs.N, dtype=torch.int32, device=self.device)
            self.work = torch.empty(_lib.ORDER_WORK_INTS, dtype=torch.int32, device=self.device)
        # eager pass: compiles the kernel, allocates every sweep's outputs
        self.results = []
        with torch.cuda.stream(self.stream):
            for k in range(sweeps):
                self._draw(k)
                self.results.append(self._solve(None))
        self.stream.synchronize()
        self.outputs = [dict(ts=r.ts, us_soa=r.us_soa, retcode=r.retcode, naccept=r.naccept, nreject=r.nreject,
                             saveat=r.saveat, order_key=r.order_key, order_auto=r.order_auto,
                             order_work=r.order_work, cost_model=r.cost_model, inputs_ordered=r.inputs_ordered) for r in self.results]
        # capture
        sp = self.stream.cuda_stream
        _lib.Graph.begin(sp)
        try:
            for k in range(sweeps):
                self._draw(k)
                self._solve(self.outputs[k])
        except BaseException:
            try:
                _lib.Graph.end(sp)
            finally:
                raise
        self.graph = _lib.Graph.end(sp)

    def _draw(self, k: int):
        pv = self.probs
        N = pv.N
        rows = len(self.draws)
        for j, d in enumerate(self.draws):
            if d.target == "tfs":
                base = pv.tfs.data_ptr()
            else:
                t = pv.p if d.target == "p" else pv.u0
                base = t.data_ptr() + d.row * N * 4
            _lib.draw_uniform(base, N, self.seed + k, rows, j, float(np.float32(d.lo)), float(np.float32(d.scale)),
                              self.stream.cuda_stream)

    def _solve(self, outputs):
        fn = vectorized_asolve if self.adaptive else vectorized_solve
        if self.order is not None:
            pv = self.probs
            _lib.order_by_span(pv.t0s.data_ptr() if pv.t0s is not None else 0,
                               pv.tfs.data_ptr() if pv.tfs is not None else 0, float(pv.tspan[0]),
                               float(pv.tspan[1]), pv.N, self.order.data_ptr(), self.work.data_ptr(),
                               self.stream.cuda_stream)
        return fn(self.probs, self.prob, self.alg, stream=self.stream, outputs=outputs, order=self.order,
                  **self.solve_kw)

    def replay(self, stream=None):
        """Launch the captured sweeps (asynchronous) on ``stream`` (default: the sweep's own stream)."""
        s = stream if stream is not None else self.stream
        self.graph.launch(s.cuda_stream)

    def synchronize(self):
        self.stream.synchronize()

    def num_nodes(self) -> int:
        return self.graph.num_nodes()

