"""Problem types and their device packing.

Mirrors the SciMLBase objects the reference path consumes: ``ODEFunction`` (with optional
``jac``/``tgrad``), out-of-place ``ODEProblem``, ``SDEFunction`` / ``SDEProblem`` (drift, diffusion,
``seed``, ``noise_rate_prototype``), ``remake``, ``EnsembleProblem``, and
DiffEqGPU's ``make_prob_compatible`` (src/utils.jl:36-40), which turns a problem into an
isbits ``ImmutableODEProblem`` (an ``ImmutableSDEProblem`` for an SDE). ``adapt(HIPDevice(), probs)`` plays the role of
``adapt(backend, probs)`` (solve.jl:399-400): it packs the batch into structure-of-arrays
device tensors (``HIPProbVector``), the layout the HIP kernels read coalesced.
"""
from __future__ import annotations

import copy
from dataclasses import dataclass, field
from typing import Any, Callable, Optional, Sequence

import numpy as np

from . import dual, tracer
from .algorithms import HIPDevice


class ODEFunction:
    """Out-of-place RHS ``f(u, p, t) -> SVector`` with optional analytic ``jac`` / ``tgrad`` and a
    constant ``mass_matrix`` M (M u' = f(u, p, t)).

    ``jac(u, p, t)`` returns n rows of n entries (row r, column c = d f_r / d u_c), or a flat
    row-major sequence. ``mass_matrix`` None (or the string "I") is the identity, SciMLBase's
    UniformScaling default; an n x n array (even the identity matrix) is applied as a matrix by
    the implicit methods, as the reference does with an SMatrix (utils.jl:42-57). Singular M
    gives a DAE (algebraic rows).
    """

    def __init__(self, f: Callable, jac: Optional[Callable] = None, tgrad: Optional[Callable] = None,
                 mass_matrix=None):
        if isinstance(f, ODEFunction):
            jac = jac or f.jac
            tgrad = tgrad or f.tgrad
            if mass_matrix is None:
                mass_matrix = f.mass_matrix
            f = f.f
        if isinstance(mass_matrix, str):
            if mass_matrix != "I":
                raise ValueError("mass_matrix must be None, 'I' or an n x n matrix")
            mass_matrix = None
        if mass_matrix is not None:
            m = np.asarray(mass_matrix, dtype=np.float64)
            if m.ndim != 2 or m.shape[0] != m.shape[1]:
                raise ValueError("mass_matrix must be square")
            if not np.all(np.isfinite(m)):
                raise ValueError("mass_matrix must be finite")
            mass_matrix = m.astype(np.float32)
        self.f = f
        self.jac = jac
        self.tgrad = tgrad
        self.mass_matrix = mass_matrix
        self._src_cache = {}

    def __call__(self, u, p, t):
        return self.f(u, p, t)

    def has_jac(self):
        return self.jac is not None

    def hip_source(self, n: int, np_: int, with_jac: bool, fd: bool = False) -> tuple[str, int]:
        """Traced HIP C for deg_rhs (+ deg_jac/deg_tgrad and the mass matrix), and the RHS flop count.
        ``fd``: a Jacobian / time gradient not given here comes from forward differences
        (``autodiff=False``) instead of forward-mode differentiation."""
        key = (n, np_, with_jac, bool(fd) and with_jac)
        if key not in self._src_cache:
            src, flops = tracer.emit_rhs(self.f, n, np_, self.jac, self.tgrad, with_jac, fd=bool(fd))
            if self.mass_matrix is not None:
                src = tracer.emit_mass(self.mass_matrix, n) + src
            self._src_cache[key] = (src, flops)
        return self._src_cache[key]


class SDEFunction:
    """Out-of-place drift ``f(u, p, t)`` and diffusion ``g(u, p, t)`` of an Ito SDE
    du = f dt + g dW. For diagonal noise g returns n entries; for general noise the n x m
    matrix (n rows of m entries, or flat row-major)."""

    def __init__(self, f: Callable, g: Optional[Callable] = None):
        if isinstance(f, SDEFunction):
            g = g or f.g
            f = f.f
        if g is None:
            raise TypeError("SDEFunction needs a diffusion g")
        self.f = f
        self.g = g
        self._src_cache = {}

    def __call__(self, u, p, t):
        return self.f(u, p, t)

    def hip_source(self, n: int, np_: int, m: Optional[int]) -> tuple[str, int]:
        """Traced HIP C for deg_rhs and deg_noise (with DEG_NM / DEG_DIAG_NOISE); m None is
        diagonal noise."""
        key = (n, np_, m)
        if key not in self._src_cache:
            self._src_cache[key] = tracer.emit_noise(self.f, self.g, n, np_, m)
        return self._src_cache[key]


def _as_f32_vec(x) -> np.ndarray:
    if x is None:
        return np.zeros(0, np.float32)
    a = np.asarray(x, dtype=np.float32).reshape(-1)
    return a


def _as_tspan(tspan) -> tuple:
    t0, tf = tspan
    return (np.float32(t0), np.float32(tf))


class DualODEFunction(ODEFunction):
    """The augmented right-hand side of a problem whose u0 / p hold dual numbers: state and
    parameters are values followed by partials (:mod:`diffeqgpu_amd.dual`), and deg_rhs is the
    traced dual evaluation of ``f``; for the implicit methods the source adds the dual Jacobian over
    the value columns and the augmented time gradient (:func:`diffeqgpu_amd.tracer.emit_dual_rhs`),
    and the kernel factors W over the values and solves each partial with that factor."""

    def __init__(self, f: ODEFunction, n: int, np_: int, P: int):
        if f.mass_matrix is not None:
            raise NotImplementedError("dual numbers with a mass matrix are not supported")
        super().__init__(f.f)
        self.base = f
        self.n, self.np, self.P = int(n), int(np_), int(P)

    def hip_source(self, n: int, np_: int, with_jac: bool, fd: bool = False) -> tuple[str, int]:
        if with_jac and (self.base.jac is not None or self.base.tgrad is not None):
            raise NotImplementedError("dual-number problems take the Jacobian from the traced f: an ODEFunction "
                                      "jac / tgrad is not evaluated on dual numbers here")
        if (n, np_) != (self.n * (1 + self.P), self.np * (1 + self.P)):
            raise ValueError("augmented sizes do not match the dual problem")
        key = ("dual", n, np_, bool(with_jac), bool(fd) and bool(with_jac))
        if key not in self._src_cache:
            self._src_cache[key] = tracer.emit_dual_rhs(self.f, self.n, self.np, self.P, with_jac=with_jac,
                                                        fd=bool(fd) and bool(with_jac))
        return self._src_cache[key]


class ODEProblem:
    """``ODEProblem{false}(f, u0, tspan, p; kwargs...)`` (out-of-place only). ``u0`` and ``p`` may
    hold :class:`~diffeqgpu_amd.dual.Dual` numbers (forward-mode sensitivities, as ForwardDiff
    duals do in the reference); they are then kept as :class:`~diffeqgpu_amd.dual.DualArray`."""

    def __init__(self, f, u0, tspan, p=None, **kwargs):
        self.f = f if isinstance(f, ODEFunction) else ODEFunction(f)
        du, dp = dual.as_dual_array(u0), dual.as_dual_array(p)
        if du is not None or dp is not None:
            P = (du if du is not None else dp).npartials
            self.u0 = dual.with_partials(du if du is not None else _as_f32_vec(u0), P)
            self.p = dual.with_partials(dp if dp is not None else _as_f32_vec(p), P)
        else:
            self.u0 = _as_f32_vec(u0)
            self.p = _as_f32_vec(p)
        self.tspan = _as_tspan(tspan)
        self.kwargs = dict(kwargs)

    @property
    def is_dual(self) -> bool:
        return isinstance(self.u0, dual.DualArray)

    def __repr__(self):
        return f"ODEProblem(u0={self.u0}, tspan={self.tspan}, p={self.p})"


class SDEProblem:
    """``SDEProblem(f, g, u0, tspan, p; seed = 0, noise_rate_prototype = nothing, kwargs...)``
    (out-of-place). ``noise_rate_prototype`` None is diagonal noise; an n x m array selects general
    noise with m Wiener processes (SciMLBase.is_diagonal_noise). ``seed`` keys the trajectory's
    noise streams (the reference calls Random.seed!(prob.seed) in its kernels,
    gpu_em_perform_step.jl:8)."""

    def __init__(self, f, g=None, u0=None, tspan=None, p=None, *, seed: int = 0, noise_rate_prototype=None,
                 **kwargs):
        self.f = f if isinstance(f, SDEFunction) else SDEFunction(f, g)
        if g is not None and isinstance(f, SDEFunction):
            self.f = SDEFunction(f.f, g)
        self.u0 = _as_f32_vec(u0)
        self.tspan = _as_tspan(tspan)
        self.p = _as_f32_vec(p)
        self.seed = int(seed) & 0xFFFFFFFFFFFFFFFF
        if noise_rate_prototype is not None:
            proto = np.asarray(noise_rate_prototype)
            if proto.ndim != 2 or proto.shape[0] != self.u0.shape[0]:
                raise ValueError("noise_rate_prototype must be an n x m matrix")
            self.noise_rate_prototype = proto
        else:
            self.noise_rate_prototype = None
        self.kwargs = dict(kwargs)

    @property
    def g(self):
        return self.f.g

    @property
    def noise_m(self) -> Optional[int]:
        """Columns of the noise matrix, or None for diagonal noise."""
        return None if self.noise_rate_prototype is None else int(self.noise_rate_prototype.shape[1])

    def __repr__(self):
        return f"SDEProblem(u0={self.u0}, tspan={self.tspan}, p={self.p}, seed={self.seed})"


def is_diagonal_noise(prob) -> bool:
    """SciMLBase.is_diagonal_noise: no noise_rate_prototype."""
    return getattr(prob, "noise_rate_prototype", None) is None


def remake(prob, *, u0=None, p=None, tspan=None, f=None, seed=None, **kwargs):
    """SciMLBase.remake for the fields the ensemble path uses."""
    kw = dict(prob.kwargs)
    kw.update(kwargs)
    if isinstance(prob, (SDEProblem, ImmutableSDEProblem)):
        return SDEProblem(f if f is not None else prob.f, None, prob.u0 if u0 is None else u0,
                          prob.tspan if tspan is None else tspan, prob.p if p is None else p,
                          seed=prob.seed if seed is None else seed, noise_rate_prototype=prob.noise_rate_prototype,
                          **kw)
    if seed is not None:
        raise TypeError("an ODEProblem has no seed")
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
    dual: Optional[tuple] = None  # (n, np, P) of a problem with dual inputs (u0 / p augmented)


@dataclass(frozen=True)
class ImmutableSDEProblem:
    """The isbits form of an SDEProblem a kernel receives (one per trajectory)."""

    f: SDEFunction
    u0: np.ndarray
    tspan: tuple
    p: np.ndarray
    seed: int = 0
    noise_rate_prototype: Any = None
    kwargs: dict = field(default_factory=dict)

    @property
    def noise_m(self) -> Optional[int]:
        return None if self.noise_rate_prototype is None else int(np.shape(self.noise_rate_prototype)[1])


def make_prob_compatible(prob):
    """src/utils.jl:36-40 (an SDEProblem is passed through in its isbits form)."""
    if isinstance(prob, (ImmutableODEProblem, ImmutableSDEProblem)):
        return prob
    if isinstance(prob, SDEProblem):
        return ImmutableSDEProblem(prob.f, prob.u0.copy(), prob.tspan, prob.p.copy(), prob.seed,
                                   prob.noise_rate_prototype, dict(prob.kwargs))
    if not isinstance(prob, ODEProblem):
        raise TypeError("make_prob_compatible expects an ODEProblem or SDEProblem")
    if prob.is_dual:
        # the augmented Float32 problem: values, then partials
        P = prob.u0.npartials
        f = DualODEFunction(prob.f, len(prob.u0), len(prob.p), P)
        return ImmutableODEProblem(f, prob.u0.augmented(), prob.tspan, prob.p.augmented(), dict(prob.kwargs),
                                   dual=(len(prob.u0), len(prob.p), P))
    return ImmutableODEProblem(prob.f, prob.u0.copy(), prob.tspan, prob.p.copy(), dict(prob.kwargs))


def _default_prob_func(prob, ctx):
    return prob


def _default_output_func(sol, ctx):
    return sol, False


class EnsembleProblem:
    """SciMLBase.EnsembleProblem (prob_func(prob, ctx), output_func(sol, ctx), safetycopy)."""

    def __init__(self, prob, prob_func: Callable = _default_prob_func,
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
    An SDE batch (``f`` an SDEFunction) also carries ``seed`` (shared) or ``seeds`` (an (N,) int64
    tensor holding the 64-bit seeds' bits) and ``noise_m`` (None: diagonal noise).
    """

    def __init__(self, f, u0, p, tspan, dev: HIPDevice, t0s=None, tfs=None, *, seed: int = 0, seeds=None,
                 noise_m: Optional[int] = None):
        import torch

        self.f = f if isinstance(f, (ODEFunction, SDEFunction)) else ODEFunction(f)
        self.is_sde = isinstance(self.f, SDEFunction)
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
        self.seed = int(seed) & 0xFFFFFFFFFFFFFFFF
        self.noise_m = None if noise_m is None else int(noise_m)
        self.seeds = None
        if seeds is not None:
            if not self.is_sde:
                raise ValueError("seeds belong to an SDE batch")
            a = np.asarray(seeds)
            if a.shape != (self.N,):
                raise ValueError("seeds must have shape (N,)")
            bits = np.ascontiguousarray(a.astype(np.uint64)).view(np.int64)
            self.seeds = torch.as_tensor(bits).to(device)

    def __len__(self):
        return self.N

    @classmethod
    def from_arrays(cls, f, u0, p, tspan, dev: HIPDevice | None = None, t0s=None, tfs=None, **sde):
        """``sde``: seed / seeds / noise_m for an SDE batch (f an SDEFunction)."""
        return cls(f, u0, p, tspan, dev or HIPDevice(), t0s, tfs, **sde)


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
    if isinstance(probs[0], ImmutableSDEProblem):
        if not all(isinstance(q, ImmutableSDEProblem) for q in probs):
            raise TypeError("a batch mixes ODE and SDE problems")
        ms = {q.noise_m for q in probs}
        if len(ms) > 1:
            raise ValueError("all problems of a batch must have the same noise shape")
        seeds = np.array([q.seed for q in probs], dtype=np.uint64)
        same = bool(np.all(seeds == seeds[0]))
        return HIPProbVector(probs[0].f, u0, p, tspan, dev, t0s, tfs, seed=int(seeds[0]),
                             seeds=None if same else seeds, noise_m=probs[0].noise_m)
    return HIPProbVector(probs[0].f, u0, p, tspan, dev, t0s, tfs)


def get_backend(probs) -> HIPDevice:
    if isinstance(probs, HIPProbVector):
        return probs.dev
    raise TypeError("probs is not on a HIP device; call adapt(HIPDevice(), probs) first")


def deepcopy_prob(prob):
    return copy.deepcopy(prob)
