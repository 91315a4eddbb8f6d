"""Callbacks for the kernel path: ``GPUDiscreteCallback``, ``GPUContinuousCallback`` and the
SciML ``DiscreteCallback`` / ``ContinuousCallback`` / ``CallbackSet`` they are converted from.

Reference: src/ensemblegpukernel/callbacks.jl:1-124 (types, defaults, conversions, the
``save_positions`` check) and src/solve.jl:382-397 with ensemblegpuarray/callbacks.jl:42-74
(which callbacks reach the kernel). The kernel applies them after every accepted step exactly as
integrators/integrator_utils.jl:52-479 does (handle_callbacks!, find_callback_time, the ITP root
finder, change_t_via_interpolation!).

A condition is ``condition(u, t, integrator)`` and an affect ``affect(integrator)``, written as
Python functions over traced values the way the right-hand side is (diffeqgpu_amd.tracer):

* a discrete condition returns a comparison, e.g. ``t == 2.4`` (Python floats are Float32
  constants, so this is ``t == 2.4f0``); combine comparisons with ``&``, ``|`` and ``~``;
* a continuous condition returns a value whose sign change marks the event, e.g. ``u[0]``;
* an affect assigns ``integrator.u`` (and may assign ``integrator.p``), e.g.
  ``integrator.u += SVector(10.0)`` or ``integrator.u = integrator.u + SVector(0.0, -2.0) * integrator.u``
  (vector ``*`` is elementwise, Julia's ``.*``), and may call ``terminate(integrator)`` or
  ``u_modified(integrator, False)``.
Indices are 0-based (``u[0]`` is Julia's ``u[1]``).
"""
from __future__ import annotations

from fractions import Fraction
from typing import List, Tuple

import numpy as np

from . import tracer
from .tracer import Graph, TraceError, Vec

_SAVE_POS_MSG = ("Callback `save_positions` are incompatible with kernel-based GPU ODE solvers due requiring "
                 "static sizing. Please ensure `save_positions = (false,false)` is set in all callback definitions "
                 "used with such solvers.")

EPS32 = float(np.finfo(np.float32).eps)


class RootfindOpt:
    """SciMLBase.RootfindOpt."""

    def __init__(self, name: str, code: int):
        self.name = name
        self.code = code

    def __repr__(self):
        return self.name


NoRootFind = RootfindOpt("NoRootFind", 0)
LeftRootFind = RootfindOpt("LeftRootFind", 1)
RightRootFind = RootfindOpt("RightRootFind", 2)

_SAME = object()  # affect_neg defaults to affect


def _save_positions(sp) -> Tuple[bool, bool]:
    sp = tuple(bool(x) for x in sp)
    if len(sp) != 2:
        raise ValueError("save_positions must have two entries")
    return sp


class DiscreteCallback:
    """SciMLBase.DiscreteCallback(condition, affect!; initialize, finalize, save_positions = (true, true))."""

    def __init__(self, condition, affect, *, initialize=None, finalize=None, save_positions=(True, True)):
        self.condition = condition
        self.affect = affect
        self.initialize = initialize
        self.finalize = finalize
        self.save_positions = _save_positions(save_positions)


class ContinuousCallback:
    """SciMLBase.ContinuousCallback(condition, affect!, affect_neg! = affect!; rootfind = LeftRootFind,
    save_positions = (true, true), interp_points = 10, abstol = 10eps(), reltol = 0,
    repeat_nudge = 1//100, idxs = nothing, dtrelax = 1)."""

    def __init__(self, condition, affect, affect_neg=_SAME, *, rootfind=LeftRootFind, save_positions=(True, True),
                 interp_points=10, abstol=10 * float(np.finfo(np.float64).eps), reltol=0, repeat_nudge=Fraction(1, 100),
                 idxs=None, dtrelax=1, initialize=None, finalize=None):
        self.condition = condition
        self.affect = affect
        self.affect_neg = affect if affect_neg is _SAME else affect_neg
        self.rootfind = rootfind
        self.save_positions = _save_positions(save_positions)
        self.interp_points = int(interp_points)
        self.abstol = abstol
        self.reltol = reltol
        self.repeat_nudge = repeat_nudge
        self.idxs = idxs
        self.dtrelax = dtrelax
        self.initialize = initialize
        self.finalize = finalize


class GPUDiscreteCallback(DiscreteCallback):
    """callbacks.jl:1-35. ``save_positions`` must be (false, false)."""

    def __init__(self, condition, affect, *, initialize=None, finalize=None, save_positions=(False, False)):
        if _save_positions(save_positions) != (False, False):
            raise ValueError(_SAVE_POS_MSG)
        super().__init__(condition, affect, initialize=initialize, finalize=finalize, save_positions=save_positions)


class GPUContinuousCallback(ContinuousCallback):
    """callbacks.jl:37-124: defaults abstol = 10eps(Float32), reltol = 0, repeat_nudge = 1//100,
    rootfind = LeftRootFind, interp_points = 10, dtrelax = 1. ``save_positions`` must be
    (false, false). ``idxs``, ``interp_points`` and ``reltol`` are kept but, as in the
    reference kernel, not used."""

    def __init__(self, condition, affect, affect_neg=_SAME, *, rootfind=LeftRootFind, save_positions=(False, False),
                 interp_points=10, abstol=10 * EPS32, reltol=0, repeat_nudge=Fraction(1, 100), idxs=None, dtrelax=1,
                 initialize=None, finalize=None):
        if _save_positions(save_positions) != (False, False):
            raise ValueError(_SAVE_POS_MSG)
        super().__init__(condition, affect, affect_neg, rootfind=rootfind, save_positions=save_positions,
                         interp_points=interp_points, abstol=abstol, reltol=reltol, repeat_nudge=repeat_nudge,
                         idxs=idxs, dtrelax=dtrelax, initialize=initialize, finalize=finalize)


class CallbackSet:
    """SciMLBase.CallbackSet: discrete and continuous callbacks, each kept in argument order
    (nested sets are flattened, ``None`` entries dropped)."""

    def __init__(self, *cbs):
        self.discrete_callbacks: List[DiscreteCallback] = []
        self.continuous_callbacks: List[ContinuousCallback] = []
        for cb in cbs:
            self._add(cb)

    def _add(self, cb):
        if cb is None:
            return
        if isinstance(cb, CallbackSet):
            self.continuous_callbacks.extend(cb.continuous_callbacks)
            self.discrete_callbacks.extend(cb.discrete_callbacks)
        elif isinstance(cb, (list, tuple)):
            for c in cb:
                self._add(c)
        elif isinstance(cb, DiscreteCallback):
            self.discrete_callbacks.append(cb)
        elif isinstance(cb, ContinuousCallback):
            self.continuous_callbacks.append(cb)
        else:
            raise TypeError("Callback unsupported")  # ensemblegpuarray/callbacks.jl:56-59

    def __len__(self):
        return len(self.discrete_callbacks) + len(self.continuous_callbacks)


def to_gpu(cb) -> GPUDiscreteCallback | GPUContinuousCallback:
    """``convert(GPUDiscreteCallback, x)`` / ``convert(GPUContinuousCallback, x)`` (callbacks.jl:30-35,
    :117-124). A SciML ContinuousCallback gets abstol = 100 eps(Float32); GPU callbacks pass through."""
    if isinstance(cb, (GPUDiscreteCallback, GPUContinuousCallback)):
        return cb
    if isinstance(cb, DiscreteCallback):
        return GPUDiscreteCallback(cb.condition, cb.affect, initialize=cb.initialize, finalize=cb.finalize,
                                   save_positions=cb.save_positions)
    if isinstance(cb, ContinuousCallback):
        return GPUContinuousCallback(cb.condition, cb.affect, cb.affect_neg, rootfind=cb.rootfind,
                                     save_positions=cb.save_positions, interp_points=cb.interp_points,
                                     abstol=100 * EPS32, reltol=cb.reltol, repeat_nudge=cb.repeat_nudge,
                                     idxs=cb.idxs, dtrelax=cb.dtrelax, initialize=cb.initialize, finalize=cb.finalize)
    raise TypeError("Callback unsupported")


def gpu_callback_set(callback) -> CallbackSet:
    """The CallbackSet of GPU callbacks the kernel runs (solve.jl:386-397)."""
    cs = callback if isinstance(callback, CallbackSet) else CallbackSet(callback)
    out = CallbackSet()
    out.discrete_callbacks = [to_gpu(c) for c in cs.discrete_callbacks]
    out.continuous_callbacks = [to_gpu(c) for c in cs.continuous_callbacks]
    return out


def merged_callback(prob_callback, kwarg_callback, merge_callbacks: bool):
    """ensemblegpuarray/callbacks.jl:61-74: the problem's callback, plus the keyword callback only
    when ``merge_callbacks`` is set. None when there is nothing."""
    kw = kwarg_callback if merge_callbacks else None
    cs = CallbackSet(prob_callback, kw)
    return cs if len(cs) else None


# ------------------------------------------------------------------------------------------
# integrator functions an affect may call
# ------------------------------------------------------------------------------------------
def terminate(integrator, retcode=None):
    """DiffEqBase.terminate!(integrator) (integrator_utils.jl:52-67): the trajectory ends after this step
    with ReturnCode.Terminated."""
    integrator._flags |= 1


def u_modified(integrator, flag: bool):
    """DiffEqBase.u_modified!(integrator, flag) (nonstiff/types.jl:46-48): ``False`` lets the next step
    reuse the last stage (FSAL) although an affect ran."""
    if flag:
        integrator._flags &= ~2
    else:
        integrator._flags |= 2


class TracedIntegrator:
    """What a traced condition or affect receives as ``integrator``: ``u``, ``p`` (vectors) and ``t``.
    In an affect ``u`` and ``p`` may be assigned."""

    def __init__(self, g: Graph, u, p, t, mutable: bool):
        self._g = g
        self._u = list(u)
        self._p = list(p)
        self._t = t
        self._mutable = mutable
        self._flags = 0

    def _set(self, name, v, n):
        if not self._mutable:
            raise TraceError(f"a condition cannot assign integrator.{name}")
        items = list(v)
        if len(items) != n:
            raise TraceError(f"integrator.{name} needs {n} entries, got {len(items)}")
        return [self._g.lift(x) for x in items]

    @property
    def u(self):
        return Vec(self._g, self._u)

    @u.setter
    def u(self, v):
        self._u = self._set("u", v, len(self._u))

    @property
    def p(self):
        return Vec(self._g, self._p)

    @p.setter
    def p(self, v):
        self._p = self._set("p", v, len(self._p))

    @property
    def t(self):
        return self._t

    @t.setter
    def t(self, v):
        raise TraceError("an affect cannot assign integrator.t (the kernel path changes t only through tstops "
                         "and continuous events)")


# ------------------------------------------------------------------------------------------
# HIP C emission (the callback part of the traced source; the contract is in include/deghip.h)
# ------------------------------------------------------------------------------------------
def _inputs(n, np_, cmp_ok):
    g = Graph(cmp_ok=cmp_ok)
    u = [g.mk("u", (), i) for i in range(n)]
    p = [g.mk("p", (), j) for j in range(np_)]
    t = g.mk("t", (), None)
    return g, u, p, t


def _f32_lit(x) -> str:
    return tracer._lit(float(np.float32(x)))


def _trace_condition(fn, n, np_, boolean: bool):
    g, u, p, t = _inputs(n, np_, True)
    integ = TracedIntegrator(g, u, p, t, mutable=False)
    out = fn(tracer.SymVec(u, "u"), t, integ)
    node = g.lift_bool(out) if boolean else g.lift(out)
    lines, names, _ = tracer._emit_values(g, [node], "        ")
    return lines + [f"        return {names[node.idx]};"]


def _trace_affect(fn, n, np_):
    g, u, p, t = _inputs(n, np_, True)
    integ = TracedIntegrator(g, u, p, t, mutable=True)
    fn(integ)
    outs, targets = [], []
    for i, x in enumerate(integ._u):
        if x is not u[i]:
            outs.append(x)
            targets.append(f"u[{i}]")
    for j, x in enumerate(integ._p):
        if x is not p[j]:
            outs.append(x)
            targets.append(f"p[{j}]")
    lines, names, _ = tracer._emit_values(g, outs, "        ")
    # every value is computed before any entry is overwritten
    lines += [f"        {tg} = {names[o.idx]};" for tg, o in zip(targets, outs)]
    return lines + [f"        return {integ._flags};"]


def _switch(sig: str, cases: List[List[str]], default: str) -> List[str]:
    out = [f"__device__ __forceinline__ {sig} {{", "    (void)j;", "    switch (j) {"]
    for j, body in enumerate(cases):
        out += [f"    case {j}: {{", *body, "    }"]
    out += ["    default: break;", "    }", f"    return {default};", "}"]
    return out


def emit_callbacks(cbset: CallbackSet, n: int, np_: int) -> str:
    """HIP C defining DEG_NDCB / DEG_NCCB and the deg_dcb_* / deg_ccb_* functions for a set of GPU
    callbacks (discrete and continuous, each in set order)."""
    dcs = [to_gpu(c) for c in cbset.discrete_callbacks]
    ccs = [to_gpu(c) for c in cbset.continuous_callbacks]
    src = ["// ---- traced callbacks ----", f"#define DEG_NDCB {len(dcs)}", f"#define DEG_NCCB {len(ccs)}"]
    if dcs:
        src += _switch("bool deg_dcb_cond(int j, const float* __restrict__ u, const float* __restrict__ p, float t)",
                       [["        (void)u; (void)p; (void)t;"] + _trace_condition(c.condition, n, np_, True) for c in dcs],
                       "false")
        src += _switch("int deg_dcb_affect(int j, float* __restrict__ u, float* __restrict__ p, float t)",
                       [["        (void)u; (void)p; (void)t;"] + _trace_affect(c.affect, n, np_) for c in dcs], "0")
    if ccs:
        src += _switch("float deg_ccb_cond(int j, const float* __restrict__ u, const float* __restrict__ p, float t)",
                       [["        (void)u; (void)p; (void)t;"] + _trace_condition(c.condition, n, np_, False)
                        for c in ccs], "0.0f")
        cases = []
        for c in ccs:
            body = ["        (void)u; (void)p; (void)t;"]
            pos = _trace_affect(c.affect, n, np_) if c.affect is not None else ["        return 0;"]
            neg = _trace_affect(c.affect_neg, n, np_) if c.affect_neg is not None else ["        return 0;"]
            body += ["        if (!neg) {", *["    " + x for x in pos], "        }", *neg]
            cases.append(body)
        src += _switch("int deg_ccb_affect(int j, int neg, float* __restrict__ u, float* __restrict__ p, float t)",
                       cases, "0")
        modes, atol, nudge, relax = [], [], [], []
        for c in ccs:
            rf = c.rootfind if isinstance(c.rootfind, RootfindOpt) else {"NoRootFind": NoRootFind,
                                                                         "LeftRootFind": LeftRootFind,
                                                                         "RightRootFind": RightRootFind}[str(c.rootfind)]
            modes.append([f"        return {(1 if c.affect is not None else 0) | (2 if c.affect_neg is not None else 0) | (rf.code << 2)};"])
            atol.append([f"        return {_f32_lit(c.abstol)};"])
            nudge.append([f"        return {_f32_lit(float(Fraction(c.repeat_nudge)))};"])
            relax.append([f"        return {_f32_lit(c.dtrelax)};"])
        src += _switch("int deg_ccb_mode(int j)", modes, "0")
        src += _switch("float deg_ccb_abstol(int j)", atol, "0.0f")
        src += _switch("float deg_ccb_nudge(int j)", nudge, "0.0f")
        src += _switch("float deg_ccb_dtrelax(int j)", relax, "1.0f")
    return "\n".join(src) + "\n"
