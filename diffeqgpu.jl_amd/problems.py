"""Problem types and their device packing.

Mirrors the SciMLBase objects the reference path consumes: ``ODEFunction`` (with optional
``jac``/``tgrad``), out-of-place ``ODEProblem``, ``remake``, ``EnsembleProblem``, and
DiffEqGPU's ``make_prob_compatible`` (src/utils.jl:36-40), which turns a problem into an
isbits ``ImmutableODEProblem``. ``adapt(HIPDevice(), probs)`` plays the role of
``adapt(backend, probs)`` (solve.jl:399-400): it packs the batch into structure-of-arrays
device tensors (``HIPProbVector``), the layout the HIP kernels read coalesced.
"""
from __future__ import annotations

import copy
from dataclasses import dataclass, field
from typing import Any, Callable, Optional, Sequence

import numpy as np

from . import tracer
from .algorithms import HIPDevice


class ODEFunction:
    """Out-of-place RHS ``f(u, p, t) -> SVector`` with optional analytic ``jac`` / ``tgrad``.

    ``jac(u, p, t)`` returns n rows of n entries (row r, column c = d f_r / d u_c), or a flat
    row-major sequence. Only the identity mass matrix is supported.
    """

    def __init__(self, f: Callable, jac: Optional[Callable] = None, tgrad: Optional[Callable] = None,
                 mass_matrix=None):
        if isinstance(f, ODEFunction):
            jac = jac or f.jac
            tgrad = tgrad or f.tgrad
            f = f.f
        if mass_matrix is not None:
            m = np.asarray(mass_matrix, dtype=np.float64)
            if not (m.ndim == 2 and m.shape[0] == m.shape[1] and np.array_equal(m, np.eye(m.shape[0]))):
                raise NotImplementedError("only the identity mass matrix is supported")
        self.f = f
        self.jac = jac
        self.tgrad = tgrad
        self._src_cache = {}

    def __call__(self, u, p, t):
        return self.f(u, p, t)

    def has_jac(self):
        return self.jac is not None

    def hip_source(self, n: int, np_: int, with_jac: bool) -> tuple[str, int]:
        """Traced HIP C for deg_rhs (+ deg_jac/deg_tgrad), and the RHS flop count."""
        key = (n, np_, with_jac)
        if key not in self._src_cache:
            self._src_cache[key] = tracer.emit_rhs(self.f, n, np_, self.jac, self.tgrad, with_jac)
        return self._src_cache[key]


def _as_f32_vec(x) -> np.ndarray:
    if x is None:
        return np.zeros(0, np.float32)
    a = np.asarray(x, dtype=np.float32).reshape(-1)
    return a


def _as_tspan(tspan) -> tuple:
    t0, tf = tspan
    return (np.float32(t0), np.float32(tf))


class ODEProblem:
    """``ODEProblem{false}(f, u0, tspan, p; kwargs...)`` (out-of-place only)."""

    def __init__(self, f, u0, tspan, p=None, **kwargs):
        self.f = f if isinstance(f, ODEFunction) else ODEFunction(f)
        self.u0 = _as_f32_vec(u0)
        self.tspan = _as_tspan(tspan)
        self.p = _as_f32_vec(p)
        self.kwargs = dict(kwargs)

    def __repr__(self):
        return f"ODEProblem(u0={self.u0}, tspan={self.tspan}, p={self.p})"


def remake(prob: ODEProblem, *, u0=None, p=None, tspan=None, f=None, **kwargs) -> ODEProblem:
    """SciMLBase.remake for the fields the ensemble path uses."""
    kw = dict(prob.kwargs)
    kw.update(kwargs)
    return ODEProblem(f if f is not None else prob.f, prob.u0 if u0 is None else u0,
                      prob.tspan if tspan is None else tspan, prob.p if p is None else p, **kw)


@dataclass(frozen=True)
class ImmutableODEProblem:
    """Isbits problem as produced by make_prob_compatible (src/utils.jl:36-40)."""

    f: ODEFunction
    u0: np.ndarray
    tspan: tuple
    p: np.ndarray
    kwargs: dict = field(default_factory=dict)


def make_prob_compatible(prob) -> ImmutableODEProblem:
    """src/utils.jl:36-40."""
    if isinstance(prob, ImmutableODEProblem):
        return prob
    if not isinstance(prob, ODEProblem):
        raise TypeError("make_prob_compatible expects an ODEProblem")
    return ImmutableODEProblem(prob.f, prob.u0.copy(), prob.tspan, prob.p.copy(), dict(prob.kwargs))


def _default_prob_func(prob, ctx):
    return prob


def _default_output_func(sol, ctx):
    return sol, False


class EnsembleProblem:
    """SciMLBase.EnsembleProblem (prob_func(prob, ctx), output_func(sol, ctx), safetycopy)."""

    def __init__(self, prob: ODEProblem, prob_func: Callable = _default_prob_func,
                 output_func: Callable = _default_output_func, reduction=None, u_init=None,
                 safetycopy: bool = True):
        self.prob = prob
        self.prob_func = prob_func
        self.output_func = output_func
        if reduction is not None:
            raise NotImplementedError("EnsembleProblem reductions are not provided by the HIP path")
        self.reduction = None
        self.u_init = u_init
        self.safetycopy = safetycopy


class HIPProbVector:
    """A batch of compatible problems packed as structure-of-arrays device tensors.

    u0[c, i] and p[j, i] are float32 tensors of shape (n, N) and (np, N) on the HIP device;
    t0s/tfs are (N,) tensors when the time spans differ, otherwise None and the shared span
    is ``tspan``. ``dev`` is the HIPDevice the batch lives on (``get_backend(probs)``).
    """

    def __init__(self, f: ODEFunction, u0, p, tspan, dev: HIPDevice, t0s=None, tfs=None):
        import torch

        self.f = f if isinstance(f, ODEFunction) else ODEFunction(f)
        self.dev = dev
        device = torch.device("cuda", dev.device)

        def put(x, shape_hint):
            if x is None:
                return None
            t = torch.as_tensor(x, dtype=torch.float32)
            if t.device != device:
                t = t.to(device)
            return t.contiguous()

        self.u0 = put(u0, None)
        if self.u0.dim() != 2:
            raise ValueError("u0 must have shape (n, N)")
        self.n, self.N = int(self.u0.shape[0]), int(self.u0.shape[1])
        pp = put(p, None) if p is not None else None
        if pp is None or pp.numel() == 0:
            pp = torch.zeros((0, self.N), dtype=torch.float32, device=device)
        if pp.dim() != 2 or pp.shape[1] != self.N:
            raise ValueError("p must have shape (np, N)")
        self.p = pp
        self.np = int(pp.shape[0])
        self.tspan = _as_tspan(tspan)
        self.t0s = put(t0s, None)
        self.tfs = put(tfs, None)
        for name, a in (("t0s", self.t0s), ("tfs", self.tfs)):
            if a is not None and (a.dim() != 1 or a.shape[0] != self.N):
                raise ValueError(f"{name} must have shape (N,)")

    def __len__(self):
        return self.N

    @classmethod
    def from_arrays(cls, f, u0, p, tspan, dev: HIPDevice | None = None, t0s=None, tfs=None):
        return cls(f, u0, p, tspan, dev or HIPDevice(), t0s, tfs)


def pack_host(probs: Sequence[ImmutableODEProblem]):
    """Host half of adapt(): SoA numpy arrays (u0 (n,N), p (np,N), t0s, tfs or None)."""
    if len(probs) == 0:
        raise ValueError("empty batch")
    n = probs[0].u0.shape[0]
    np_ = probs[0].p.shape[0]
    N = len(probs)
    u0 = np.empty((n, N), np.float32)
    p = np.empty((np_, N), np.float32)
    t0s = np.empty(N, np.float32)
    tfs = np.empty(N, np.float32)
    for i, pr in enumerate(probs):
        if pr.u0.shape[0] != n or pr.p.shape[0] != np_:
            raise ValueError("all problems of a batch must have the same state and parameter sizes")
        u0[:, i] = pr.u0
        p[:, i] = pr.p
        t0s[i], tfs[i] = pr.tspan
    same = bool(np.all(t0s == t0s[0]) and np.all(tfs == tfs[0]))
    return u0, p, (None if same else t0s), (None if same else tfs), (t0s[0], tfs[0])


def adapt(dev: HIPDevice, probs) -> HIPProbVector:
    """``adapt(backend, probs)``: move a vector of compatible problems to the device."""
    if isinstance(probs, HIPProbVector):
        return probs
    probs = [make_prob_compatible(p) for p in probs]
    u0, p, t0s, tfs, tspan = pack_host(probs)
    return HIPProbVector(probs[0].f, u0, p, tspan, dev, t0s, tfs)


def get_backend(probs) -> HIPDevice:
    if isinstance(probs, HIPProbVector):
        return probs.dev
    raise TypeError("probs is not on a HIP device; call adapt(HIPDevice(), probs) first")


def deepcopy_prob(prob):
    return copy.deepcopy(prob)
