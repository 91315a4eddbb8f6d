"""Trace an out-of-place right-hand side ``f(u, p, t)`` into HIP C.

The reference compiles the user's ``f(u::SVector, p, t)::SVector`` straight into its
GPU kernel (src/ensemblegpukernel/kernels.jl:1-152, the RHS is ``prob.f``). Here the
Python function is called once with symbolic arguments; every arithmetic operation it
performs is recorded in evaluation order and emitted as Float32 HIP C, so the kernel
performs the same IEEE operations with the same association as the source expression.

Python float constants are Float32 constants (``(float)<double literal>``, i.e.
``convert(Float32, lit)``); integer constants are exact. Elementary functions come from
:mod:`diffeqgpu_amd.math` (``sin``, ``cos``, ``exp``, ``log``, ``sqrt``, ``tanh``, ``abs``,
``pow``), which accept both numbers and traced values.

For Rosenbrock23 the Jacobian ``d f / d u`` and ``d f / d t`` are produced by symbolic
forward-mode differentiation of the recorded expression graph (the role ForwardDiff plays
in nlsolve/type.jl:129-157) unless the problem supplies ``jac``/``tgrad`` functions,
which are traced like ``f``.
"""
from __future__ import annotations

import math as _pymath
import struct
from typing import Callable, Dict, List, Optional, Sequence, Tuple


class TraceError(TypeError):
    pass


def _f32(x: float) -> float:
    return struct.unpack("f", struct.pack("f", float(x)))[0]


class Node:
    """A traced scalar. Nodes are hash-consed per :class:`Graph`."""

    __slots__ = ("g", "op", "args", "val", "idx", "_key")

    def __init__(self, g: "Graph", op: str, args: Tuple["Node", ...], val=None):
        self.g = g
        self.op = op
        self.args = args
        self.val = val
        self.idx = -1
        self._key = None

    # -- arithmetic (left operand first, exactly as written) --
    def _bin(self, op, other, rev=False):
        if isinstance(other, DualNode):
            return NotImplemented  # a real times a dual: the dual's reflected operator
        o = self.g.lift(other)
        return self.g.mk(op, (o, self) if rev else (self, o))

    def __add__(self, o): return self._bin("+", o)
    def __radd__(self, o): return self._bin("+", o, True)
    def __sub__(self, o): return self._bin("-", o)
    def __rsub__(self, o): return self._bin("-", o, True)
    def __mul__(self, o): return self._bin("*", o)
    def __rmul__(self, o): return self._bin("*", o, True)
    def __truediv__(self, o): return self._bin("/", o)
    def __rtruediv__(self, o): return self._bin("/", o, True)
    def __neg__(self): return self.g.mk("neg", (self,))
    def __pos__(self): return self

    def __pow__(self, e):
        # Julia lowers x^2, x^3 with literal integer exponents to repeated multiplication
        # (Base.literal_pow); other exponents call pow.
        if isinstance(e, DualNode):
            return NotImplemented
        if isinstance(e, int) and not isinstance(e, bool):
            if e == 0:
                return self.g.const(1.0)
            if e == 1:
                return self
            if e == 2:
                return self * self
            if e == 3:
                return (self * self) * self
            if e == -1:
                return self.g.const(1.0) / self
        return self.g.mk("pow", (self, self.g.lift(e)))

    def __rpow__(self, b):
        return self.g.mk("pow", (self.g.lift(b), self))

    def __bool__(self):
        raise TraceError("a traced value has no truth value: use diffeqgpu_amd.math.ifelse for branches, and "
                         "& | ~ (not and/or/not) to combine the comparisons of a callback condition")

    # -- comparisons: recorded only while tracing a callback condition (Graph.cmp_ok) --
    def _cmp(self, other, op, rev=False):
        g = self.g
        if not g.cmp_ok:
            raise TraceError("comparisons of traced values are recorded only in callback conditions; "
                             "use diffeqgpu_amd.math.ifelse in a right-hand side")
        o = g.lift(other)
        return g.mk("cmp", (o, self) if rev else (self, o), op)

    def __lt__(self, o): return self._cmp(o, "<")
    def __le__(self, o): return self._cmp(o, "<=")
    def __gt__(self, o): return self._cmp(o, ">")
    def __ge__(self, o): return self._cmp(o, ">=")

    def __eq__(self, o):
        if self.g.cmp_ok and _is_scalar(o):
            return self._cmp(o, "==")
        return self is o

    def __ne__(self, o):
        if self.g.cmp_ok and _is_scalar(o):
            return self._cmp(o, "!=")
        return self is not o

    __hash__ = object.__hash__

    # -- boolean combinations of comparisons (Julia's & | ! on Bool) --
    def _bool_bin(self, other, op):
        g = self.g
        return g.mk(op, (g.lift_bool(self), g.lift_bool(other)))

    def __and__(self, o): return self._bool_bin(o, "and")
    def __rand__(self, o): return self.g.mk("and", (self.g.lift_bool(o), self.g.lift_bool(self)))
    def __or__(self, o): return self._bool_bin(o, "or")
    def __ror__(self, o): return self.g.mk("or", (self.g.lift_bool(o), self.g.lift_bool(self)))
    def __invert__(self): return self.g.mk("not", (self.g.lift_bool(self),))

    @property
    def is_bool(self):
        return self.op in _BOOL_OPS

    def __float__(self):
        raise TraceError("cannot convert a traced value to float inside a traced RHS")

    def __repr__(self):
        return f"Node({self.op}, #{self.idx})"


_UNARY = {"neg", "sin", "cos", "exp", "log", "sqrt", "tanh", "abs"}
_BOOL_OPS = {"cmp", "and", "or", "not", "bconst"}


def _is_scalar(x) -> bool:
    if isinstance(x, Node):
        return True
    if isinstance(x, bool):
        return False
    if isinstance(x, (int, float)):
        return True
    try:
        import numpy as np

        return isinstance(x, np.generic) and not isinstance(x, np.bool_)
    except ImportError:  # pragma: no cover
        return False


class Graph:
    def __init__(self, cmp_ok: bool = False):
        self.nodes: List[Node] = []
        self.table: Dict[tuple, Node] = {}
        self.cmp_ok = cmp_ok

    def _key(self, op, args, val):
        return (op, tuple(a.idx for a in args), val)

    def mk(self, op, args, val=None):
        for a in args:
            if not isinstance(a, Node) or a.g is not self:
                raise TraceError("mixing values from different traces")
        key = self._key(op, args, val)
        n = self.table.get(key)
        if n is None:
            n = Node(self, op, tuple(args), val)
            n.idx = len(self.nodes)
            self.nodes.append(n)
            self.table[key] = n
        return n

    def const(self, x: float) -> Node:
        return self.mk("const", (), float(x))

    def lift_bool(self, x) -> Node:
        """A boolean node: a traced comparison (or combination) or a Python bool constant."""
        if isinstance(x, Node):
            if not x.is_bool:
                raise TraceError("& | ~ combine comparisons, not arithmetic values")
            return x
        if isinstance(x, bool):
            return self.mk("bconst", (), bool(x))
        try:
            import numpy as np

            if isinstance(x, np.bool_):
                return self.mk("bconst", (), bool(x))
        except ImportError:  # pragma: no cover
            pass
        raise TraceError(f"expected a condition, got {type(x).__name__}")

    def lift(self, x) -> Node:
        if isinstance(x, Node):
            if x.is_bool:
                raise TraceError("a comparison is not an arithmetic value")
            return x
        if isinstance(x, bool):
            raise TraceError("booleans are not arithmetic values")
        if isinstance(x, (int, float)):
            return self.const(x)
        try:
            import numpy as np

            if isinstance(x, np.generic):
                return self.const(float(x))
        except ImportError:  # pragma: no cover
            pass
        raise TraceError(f"unsupported value in traced RHS: {type(x).__name__}")


class SymVec:
    """Read-only indexable vector of traced inputs (u or p), 0-based like a Python list."""

    def __init__(self, items: Sequence[Node], name: str):
        self._items = list(items)
        self._name = name

    def __len__(self):
        return len(self._items)

    def __getitem__(self, i):
        if isinstance(i, slice):
            return list(self._items[i])
        return self._items[i]

    def __iter__(self):
        return iter(self._items)


class SVector(tuple):
    """Out-of-place RHS return type, the counterpart of StaticArrays' ``SVector``."""

    def __new__(cls, *xs):
        if len(xs) == 1 and isinstance(xs[0], (list, tuple)):
            xs = tuple(xs[0])
        return super().__new__(cls, xs)


class Vec:
    """The state or parameter vector an affect sees: elementwise ``+ - * /`` with another vector
    (an SVector, list, tuple or Vec of the same length) or a scalar, the way Julia's
    ``integrator.u + @SVector[...]`` and ``@SVector[...] .* integrator.u`` combine SVectors.
    Items may also be assigned (``integrator.u[0] = ...``)."""

    def __init__(self, g: "Graph", items: List[Node]):
        self._g = g
        self._items = items

    def __len__(self):
        return len(self._items)

    def __getitem__(self, i):
        if isinstance(i, slice):
            return list(self._items[i])
        return self._items[i]

    def __setitem__(self, i, v):
        self._items[i] = self._g.lift(v)

    def __iter__(self):
        return iter(self._items)

    def _zip(self, o, fn, rev=False):
        g = self._g
        if isinstance(o, (Vec, SymVec, list, tuple)):
            if len(o) != len(self._items):
                raise TraceError(f"vector lengths differ: {len(self._items)} and {len(o)}")
            pairs = [(g.lift(a), g.lift(b)) for a, b in zip(self._items, o)]
        else:
            bb = g.lift(o)
            pairs = [(g.lift(a), bb) for a in self._items]
        return Vec(g, [fn(b, a) if rev else fn(a, b) for a, b in pairs])

    def __add__(self, o): return self._zip(o, lambda a, b: a + b)
    def __radd__(self, o): return self._zip(o, lambda a, b: a + b, True)
    def __sub__(self, o): return self._zip(o, lambda a, b: a - b)
    def __rsub__(self, o): return self._zip(o, lambda a, b: a - b, True)
    def __mul__(self, o): return self._zip(o, lambda a, b: a * b)
    def __rmul__(self, o): return self._zip(o, lambda a, b: a * b, True)
    def __truediv__(self, o): return self._zip(o, lambda a, b: a / b)
    def __rtruediv__(self, o): return self._zip(o, lambda a, b: a / b, True)
    def __neg__(self): return Vec(self._g, [-a for a in self._items])


# ------------------------------------------------------------------------------------------
# elementary functions usable on numbers and traced values
# ------------------------------------------------------------------------------------------
def _unary(name, pyfn):
    def fn(x):
        if isinstance(x, DualNode):
            return x.unary(name)
        if isinstance(x, Node):
            return x.g.mk(name, (x,))
        return pyfn(x)

    fn.__name__ = name
    return fn


sin = _unary("sin", _pymath.sin)
cos = _unary("cos", _pymath.cos)
exp = _unary("exp", _pymath.exp)
log = _unary("log", _pymath.log)
sqrt = _unary("sqrt", _pymath.sqrt)
tanh = _unary("tanh", _pymath.tanh)
fabs = _unary("abs", abs)


def pow(x, y):  # noqa: A001 - mirrors the math function name
    if isinstance(x, DualNode) or isinstance(y, DualNode):
        return _dual_pow(x, y)
    if isinstance(x, Node) or isinstance(y, Node):
        g = x.g if isinstance(x, Node) else y.g
        return g.mk("pow", (g.lift(x), g.lift(y)))
    return _pymath.pow(x, y)


def ifelse(cond_lhs, cmp, cond_rhs=None, a=None, b=None):
    """``cond_lhs <cmp> cond_rhs ? a : b`` with cmp in <, <=, >, >=, ==, != (both branches evaluated).
    Inside a callback condition or affect, ``ifelse(cond, a, b)`` also takes a traced comparison."""
    if b is None and a is None and cond_rhs is not None:
        cond, a, b = cond_lhs, cmp, cond_rhs
        if isinstance(cond, Node):
            g = cond.g
            return g.mk("bselect", (g.lift_bool(cond), g.lift(a), g.lift(b)))
        if isinstance(cond, bool):
            return a if cond else b
        raise TraceError("ifelse(cond, a, b) needs a traced comparison or a bool")
    if cmp not in ("<", "<=", ">", ">=", "==", "!="):
        raise TraceError(f"bad comparison {cmp!r}")
    if any(isinstance(v, DualNode) for v in (cond_lhs, cond_rhs, a, b)):
        # comparisons of dual numbers compare their values; the branch selects value and partials
        lhs = cond_lhs.v if isinstance(cond_lhs, DualNode) else cond_lhs
        rhs = cond_rhs.v if isinstance(cond_rhs, DualNode) else cond_rhs
        P = next(v.P for v in (cond_lhs, cond_rhs, a, b) if isinstance(v, DualNode))
        g = next(v.g for v in (cond_lhs, cond_rhs, a, b) if isinstance(v, DualNode))
        da, db = _as_dual(g, a, P), _as_dual(g, b, P)
        return DualNode(ifelse(lhs, cmp, rhs, da.v, db.v), [ifelse(lhs, cmp, rhs, x, y) for x, y in zip(da.d, db.d)])
    vals = [cond_lhs, cond_rhs, a, b]
    g = next((v.g for v in vals if isinstance(v, Node)), None)
    if g is None:
        ok = {"<": cond_lhs < cond_rhs, "<=": cond_lhs <= cond_rhs, ">": cond_lhs > cond_rhs,
              ">=": cond_lhs >= cond_rhs, "==": cond_lhs == cond_rhs, "!=": cond_lhs != cond_rhs}[cmp]
        return a if ok else b
    return g.mk("select", tuple(g.lift(v) for v in vals), cmp)


# ------------------------------------------------------------------------------------------
# dual numbers (forward mode, the arithmetic of ForwardDiff.Dual) over traced values
# ------------------------------------------------------------------------------------------
class DualNode:
    """A traced dual number: a value node ``v`` and P partial nodes ``d``. Each operation emits
    the value and the partials as separate Float32 operations, in these forms (x, y dual, r real):
    x +- y: (xv +- yv, xd +- yd); x * y: (xv yv, xd yv + yd xv); x / y: (xv / yv,
    xd (1 / yv) + yd (-(xv / (yv yv)))); r * x: (r xv, r xd); x / r: (xv / r, xd / r);
    r / y: (r / yv, yd (-(r / (yv yv)))); f(x): (f(xv), f'(xv) xd) with f' = cos, -sin, exp,
    1 / x, 1 / (2 sqrt x), 1 - tanh^2, and for abs the sign of xv (negative partials where xv < 0)."""

    __slots__ = ("v", "d")

    def __init__(self, v: Node, d: Sequence[Node]):
        self.v = v
        self.d = tuple(d)

    @property
    def g(self) -> "Graph":
        return self.v.g

    @property
    def P(self) -> int:
        return len(self.d)

    def _other(self, o):
        if isinstance(o, DualNode):
            if o.P != self.P:
                raise TraceError("dual numbers with different numbers of partials")
            return o
        return self.g.lift(o)

    def __add__(self, o):
        o = self._other(o)
        if isinstance(o, DualNode):
            return DualNode(self.v + o.v, [a + b for a, b in zip(self.d, o.d)])
        return DualNode(self.v + o, self.d)

    def __radd__(self, o):
        o = self.g.lift(o)
        return DualNode(o + self.v, self.d)

    def __sub__(self, o):
        o = self._other(o)
        if isinstance(o, DualNode):
            return DualNode(self.v - o.v, [a - b for a, b in zip(self.d, o.d)])
        return DualNode(self.v - o, self.d)

    def __rsub__(self, o):
        o = self.g.lift(o)
        return DualNode(o - self.v, [-a for a in self.d])

    def __mul__(self, o):
        o = self._other(o)
        if isinstance(o, DualNode):
            return DualNode(self.v * o.v, [a * o.v + b * self.v for a, b in zip(self.d, o.d)])
        return DualNode(self.v * o, [a * o for a in self.d])

    def __rmul__(self, o):
        o = self.g.lift(o)
        return DualNode(o * self.v, [o * a for a in self.d])

    def __truediv__(self, o):
        o = self._other(o)
        if isinstance(o, DualNode):
            inv = 1.0 / o.v
            c = -(self.v / (o.v * o.v))
            return DualNode(self.v / o.v, [a * inv + b * c for a, b in zip(self.d, o.d)])
        return DualNode(self.v / o, [a / o for a in self.d])

    def __rtruediv__(self, o):
        o = self.g.lift(o)
        c = -(o / (self.v * self.v))
        return DualNode(o / self.v, [b * c for b in self.d])

    def __neg__(self):
        return DualNode(-self.v, [-a for a in self.d])

    def __pos__(self):
        return self

    def __pow__(self, e):
        if isinstance(e, int) and not isinstance(e, bool):
            if e == 0:
                return _as_dual(self.g, 1.0, self.P)
            if e == 1:
                return self
            if e == 2:
                return self * self
            if e == 3:
                return (self * self) * self
            if e == -1:
                return 1.0 / self
        return _dual_pow(self, e)

    def __rpow__(self, b):
        return _dual_pow(b, self)

    def __bool__(self):
        raise TraceError("a traced dual number has no truth value: use diffeqgpu_amd.math.ifelse")

    def unary(self, name: str) -> "DualNode":
        g, x = self.g, self.v
        if name == "abs":
            # DiffRules: abs'(x) = signbit(x) ? -1 : 1 (so -0.0 negates the partials too)
            return DualNode(g.mk("abs", (x,)), [g.mk("sbsel", (x, -a, a)) for a in self.d])
        fx = g.mk(name, (x,))
        if name == "sin":
            fp = g.mk("cos", (x,))
        elif name == "cos":
            fp = -g.mk("sin", (x,))
        elif name == "exp":
            fp = fx
        elif name == "log":
            fp = 1.0 / x
        elif name == "sqrt":
            fp = 1.0 / (2.0 * fx)
        elif name == "tanh":
            fp = 1.0 - fx * fx
        else:  # pragma: no cover
            raise TraceError(f"no dual rule for {name}")
        return DualNode(fx, [fp * a for a in self.d])


def _as_dual(g: "Graph", x, P: int) -> DualNode:
    if isinstance(x, DualNode):
        return x
    zero = g.const(0.0)
    return DualNode(g.lift(x), [zero] * P)


def _all_zero(g: "Graph", parts) -> Node:
    """A boolean node: every partial is zero (ForwardDiff's iszero(partials) / isconstant)."""
    zero = g.const(0.0)
    acc = None
    for d in parts:
        c = g.mk("cmp", (d, zero), "==")
        acc = c if acc is None else g.mk("and", (acc, c))
    return acc if acc is not None else g.lift_bool(True)


def _dual_pow(x, y) -> DualNode:
    """x^y with a dual base and/or exponent, by ForwardDiff's rules for ^ (ForwardDiff 1.3.3, not
    vendored in the reference), whose guards keep log(x) out where it would give NaN:
      dual ^ real:  partials 0 if y == 0 or x's partials are all 0, else (x' y) x^(y - 1);
      real ^ dual:  partials (x == 0 and y > 0 ? 0 : x^y log x) y';
      dual ^ dual:  x' (y x^(y - 1)) + y' L with L = 1 if y's partials are all 0, else 0 if x == 0 and
                    y > 0, else x^y log x."""
    ref = x if isinstance(x, DualNode) else y
    g, P = ref.g, ref.P
    zero = g.const(0.0)
    if isinstance(x, DualNode) and not isinstance(y, DualNode):
        if isinstance(y, (int, float)) and not isinstance(y, bool) and y == 0:
            return DualNode(g.mk("pow", (x.v, g.lift(y))), [zero] * P)
        yv = g.lift(y)
        pw = g.mk("pow", (x.v, yv - 1.0))
        const_x = _all_zero(g, x.d)
        ynz = g.mk("cmp", (yv, zero), "==")
        dead = g.mk("or", (ynz, const_x))
        return DualNode(g.mk("pow", (x.v, yv)), [g.mk("bselect", (dead, zero, (a * yv) * pw)) for a in x.d])
    if not isinstance(x, DualNode):
        xv = g.lift(x)
        val = g.mk("pow", (xv, y.v))
        lg = val * g.mk("log", (xv,))
        guard = g.mk("and", (g.mk("cmp", (xv, zero), "=="), g.mk("cmp", (y.v, zero), ">")))
        fp = g.mk("bselect", (guard, zero, lg))
        return DualNode(val, [fp * b for b in y.d])
    val = g.mk("pow", (x.v, y.v))
    fa = y.v * g.mk("pow", (x.v, y.v - 1.0))
    lg = val * g.mk("log", (x.v,))
    guard = g.mk("and", (g.mk("cmp", (x.v, zero), "=="), g.mk("cmp", (y.v, zero), ">")))
    fb = g.mk("bselect", (_all_zero(g, y.d), g.const(1.0), g.mk("bselect", (guard, zero, lg))))
    return DualNode(val, [a * fa + b * fb for a, b in zip(x.d, y.d)])


def trace_dual(f: Callable, n: int, np_: int, P: int) -> "Trace":
    """Trace f on dual inputs: a Trace over the n (1 + P) state and np (1 + P) parameter entries
    (values, then the partials of each component) whose outputs are the n values followed by the
    partials of each output component."""
    if P < 1:
        raise TraceError("a dual problem needs at least one partial")
    na, npa = n * (1 + P), np_ * (1 + P)
    g, ua, pa, t = _new_inputs(na, npa)
    u = [DualNode(ua[i], [ua[n + i * P + j] for j in range(P)]) for i in range(n)]
    p = [DualNode(pa[k], [pa[np_ + k * P + j] for j in range(P)]) for k in range(np_)]
    out = f(SymVec(u, "u"), SymVec(p, "p"), t)
    if isinstance(out, (int, float, Node, DualNode)):
        out = (out,)
    out = list(out)
    if len(out) != n:
        raise TraceError(f"f must return {n} entries, got {len(out)}")
    duals = [_as_dual(g, x, P) for x in out]
    outs = [x.v for x in duals] + [x.d[j] for x in duals for j in range(P)]
    return Trace(g, ua, pa, t, outs)


def emit_dual_rhs(f: Callable, n: int, np_: int, P: int, with_jac: bool = False,
                  fd: bool = False) -> Tuple[str, int]:
    """HIP C for the augmented right-hand side of a problem with dual inputs: deg_rhs over the
    n (1 + P) state entries (values, then the partials of each component) with the np (1 + P)
    parameters laid out the same way, and the macros DEG_DUAL_NV (n) and DEG_DUAL_P (P) that
    select the dual error norm and the dual W factor. Returns the source and the flop count of the
    value part.

    ``with_jac`` (the implicit methods) adds the dual Jacobian over the value columns, deg_jac
    writing J[r n + c] = d F_r / d u_c for the n (1 + P) augmented rows r and the n values c (rows n
    + i P + j are partial j of the Jacobian's row i), and the augmented time gradient deg_tgrad: by
    forward-mode differentiation of the traced dual graph, or with ``fd`` (``autodiff=False``) by
    forward differences of the dual deg_rhs that perturb a value alone, as finite_diff_jac does on a
    dual u (alg_utils.jl:17-27)."""
    tr = trace_dual(f, n, np_, P)
    body, _ = _emit_body(tr.g, tr.outs, "du")
    # the value part alone, for the flop count
    vtr = trace(f, n, np_, nout=n)
    _vb, flops = _emit_body(vtr.g, vtr.outs, "du")
    src = [f"#define DEG_DUAL_NV {n}", f"#define DEG_DUAL_P {P}",
           "__device__ __forceinline__ void deg_rhs(const float* __restrict__ u, const float* __restrict__ p, float t,",
           "                                        float* __restrict__ du) {",
           "    (void)u; (void)p; (void)t;", *body, "}"]
    if with_jac:
        N = n * (1 + P)
        if fd:
            src += _fd_jac_lines(N, cols=n) + _fd_tgrad_lines(N)
        else:
            g = tr.g
            zero = g.const(0.0)
            J = []
            for o in tr.outs:
                for c in range(n):
                    dd = _Diff(g, tr.u[c]).d(o)
                    J.append(zero if dd is None else dd)
            T = []
            for o in tr.outs:
                dd = _Diff(g, tr.t).d(o)
                T.append(zero if dd is None else dd)
            jbody, _ = _emit_body(g, J, "J")
            tbody, _ = _emit_body(g, T, "dT")
            src += ["__device__ __forceinline__ void deg_jac(const float* __restrict__ u, const float* __restrict__ p, float t,",
                    "                                        float* __restrict__ J) {",
                    "    (void)u; (void)p; (void)t;", *jbody, "}",
                    "__device__ __forceinline__ void deg_tgrad(const float* __restrict__ u, const float* __restrict__ p, float t,",
                    "                                          float* __restrict__ dT) {",
                    "    (void)u; (void)p; (void)t;", *tbody, "}"]
    return "\n".join(src) + "\n", flops


# ------------------------------------------------------------------------------------------
# tracing
# ------------------------------------------------------------------------------------------
class Trace:
    """Result of tracing one function: inputs, the graph and the output nodes."""

    def __init__(self, g: Graph, u: List[Node], p: List[Node], t: Node, outs: List[Node]):
        self.g = g
        self.u = u
        self.p = p
        self.t = t
        self.outs = outs


def _new_inputs(n: int, np_: int):
    g = Graph()
    u = [g.mk("u", (), i) for i in range(n)]
    p = [g.mk("p", (), j) for j in range(np_)]
    t = g.mk("t", (), None)
    return g, u, p, t


def trace(f: Callable, n: int, np_: int, nout: int | None = None, g: Graph | None = None,
          inputs=None) -> Trace:
    if g is None:
        g, u, p, t = _new_inputs(n, np_)
    else:
        u, p, t = inputs
    out = f(SymVec(u, "u"), SymVec(p, "p"), t)
    if isinstance(out, (int, float, Node)):
        out = (out,)
    try:
        outs = [g.lift(x) for x in out]
    except TypeError as e:
        raise TraceError(f"RHS must return a sequence of scalars: {e}") from e
    if nout is not None and len(outs) != nout:
        raise TraceError(f"expected {nout} outputs, got {len(outs)}")
    return Trace(g, u, p, t, outs)


# ------------------------------------------------------------------------------------------
# symbolic forward-mode differentiation
# ------------------------------------------------------------------------------------------
class _Diff:
    """d(node)/d(var) over the graph, with zero/one folding."""

    def __init__(self, g: Graph, var: Node):
        self.g = g
        self.var = var
        self.memo: Dict[int, Node | None] = {}

    def zero(self, x):
        return x is None

    def add(self, a, b):
        if a is None:
            return b
        if b is None:
            return a
        return a + b

    def sub(self, a, b):
        if b is None:
            return a
        if a is None:
            return -b
        return a - b

    def mul(self, a, b):  # a, b may be None (zero)
        if a is None or b is None:
            return None
        if a.op == "const" and a.val == 1.0:
            return b
        if b.op == "const" and b.val == 1.0:
            return a
        return a * b

    def d(self, x: Node):
        if x.idx in self.memo:
            return self.memo[x.idx]
        r = self._d(x)
        self.memo[x.idx] = r
        return r

    def _d(self, x: Node):
        g = self.g
        if x is self.var:
            return g.const(1.0)
        op = x.op
        if op in ("const", "u", "p", "t"):
            return None
        a = x.args
        if op == "+":
            return self.add(self.d(a[0]), self.d(a[1]))
        if op == "-":
            return self.sub(self.d(a[0]), self.d(a[1]))
        if op == "*":
            return self.add(self.mul(self.d(a[0]), a[1]), self.mul(a[0], self.d(a[1])))
        if op == "/":
            da, db = self.d(a[0]), self.d(a[1])
            t1 = None if da is None else da / a[1]
            t2 = None if db is None else (x * db) / a[1]
            return self.sub(t1, t2)
        if op == "neg":
            da = self.d(a[0])
            return None if da is None else -da
        if op == "sin":
            return self.mul(g.mk("cos", (a[0],)), self.d(a[0]))
        if op == "cos":
            da = self.d(a[0])
            return None if da is None else -(g.mk("sin", (a[0],)) * da)
        if op == "exp":
            return self.mul(x, self.d(a[0]))
        if op == "log":
            da = self.d(a[0])
            return None if da is None else da / a[0]
        if op == "sqrt":
            da = self.d(a[0])
            return None if da is None else da / (g.const(2.0) * x)
        if op == "tanh":
            da = self.d(a[0])
            return None if da is None else (g.const(1.0) - x * x) * da
        if op == "abs":
            # DiffRules: abs'(x) = signbit(x) ? -1 : 1
            da = self.d(a[0])
            if da is None:
                return None
            return g.mk("sbsel", (a[0], -da, da))
        if op == "pow":
            # ForwardDiff's ^ rules with their guards: the base term vanishes for a zero exponent;
            # the log term is 0 where the base is 0 and the exponent positive
            base, ex = a
            db, de = self.d(base), self.d(ex)
            zero = g.const(0.0)
            r = None
            if db is not None:
                bt = (ex * g.mk("pow", (base, ex - g.const(1.0)))) * db
                if de is None:
                    bt = g.mk("bselect", (g.mk("cmp", (ex, zero), "=="), zero, bt))
                r = bt
            if de is not None:
                guard = g.mk("and", (g.mk("cmp", (base, zero), "=="), g.mk("cmp", (ex, zero), ">")))
                r = self.add(r, g.mk("bselect", (guard, zero, x * g.mk("log", (base,)))) * de)
            return r
        if op == "select":
            da, db = self.d(a[2]), self.d(a[3])
            if da is None and db is None:
                return None
            return g.mk("select", (a[0], a[1], da if da is not None else g.const(0.0),
                                   db if db is not None else g.const(0.0)), x.val)
        raise TraceError(f"cannot differentiate op {op}")


def jacobian_nodes(tr: Trace) -> Tuple[List[List[Node]], List[Node]]:
    """Rows J[r][c] = d out_r / d u_c and tgrad[r] = d out_r / d t (zeros as const 0)."""
    g = tr.g
    zero = g.const(0.0)
    J = []
    for r, o in enumerate(tr.outs):
        row = []
        for c, uc in enumerate(tr.u):
            dd = _Diff(g, uc).d(o)
            row.append(zero if dd is None else dd)
        J.append(row)
    T = []
    for o in tr.outs:
        dd = _Diff(g, tr.t).d(o)
        T.append(zero if dd is None else dd)
    return J, T


# ------------------------------------------------------------------------------------------
# HIP C emission
# ------------------------------------------------------------------------------------------
_FN = {"sin": "sinf", "cos": "cosf", "exp": "expf", "log": "logf", "sqrt": "__builtin_sqrtf",
       "tanh": "tanhf", "abs": "__builtin_fabsf"}


def _lit(v: float) -> str:
    if v == int(v) and abs(v) < 2 ** 24:
        return f"{int(v)}.0f"
    return f"(float){v!r}"


def _emit_body(g: Graph, outs: List[Node], out_name: str, indent="    ") -> Tuple[List[str], int]:
    """Emit statements computing `outs` into out_name[i]; returns lines and the number of flops."""
    lines, names, flops = _emit_values(g, outs, indent)
    for i, o in enumerate(outs):
        lines.append(f"{indent}{out_name}[{i}] = {names[o.idx]};")
    return lines, flops


def _emit_values(g: Graph, outs: List[Node], indent="    "):
    """Emit one const statement per node `outs` depend on (in recording order); returns the lines,
    the C expression naming each node, and the number of flops."""
    need = set()
    stack = list(outs)
    while stack:
        x = stack.pop()
        if x.idx in need:
            continue
        need.add(x.idx)
        stack.extend(x.args)
    names = {}
    lines = []
    flops = 0
    for x in g.nodes:
        if x.idx not in need:
            continue
        op = x.op
        if op == "const":
            names[x.idx] = _lit(x.val)
            continue
        if op == "u":
            names[x.idx] = f"u[{x.val}]"
            continue
        if op == "p":
            names[x.idx] = f"p[{x.val}]"
            continue
        if op == "t":
            names[x.idx] = "t"
            continue
        if op == "bconst":
            names[x.idx] = "true" if x.val else "false"
            continue
        a = [names[y.idx] for y in x.args]
        if op in _BOOL_OPS:
            if op == "cmp":
                expr = f"({a[0]} {x.val} {a[1]})"
            elif op == "and":
                expr = f"({a[0]} && {a[1]})"
            elif op == "or":
                expr = f"({a[0]} || {a[1]})"
            else:  # not
                expr = f"!({a[0]})"
            v = f"v{x.idx}"
            lines.append(f"{indent}const bool {v} = {expr};")
            names[x.idx] = v
            continue
        if op in ("+", "-", "*", "/"):
            expr = f"{a[0]} {op} {a[1]}"
            flops += 1
        elif op == "neg":
            expr = f"-({a[0]})"
        elif op in _FN:
            expr = f"{_FN[op]}({a[0]})"
            if op == "sqrt":
                flops += 1
        elif op == "pow":
            expr = f"powf({a[0]}, {a[1]})"
        elif op == "select":
            expr = f"({a[0]} {x.val} {a[1]}) ? {a[2]} : {a[3]}"
        elif op == "bselect":
            expr = f"{a[0]} ? {a[1]} : {a[2]}"
        elif op == "sbsel":
            expr = f"__builtin_signbit({a[0]}) ? {a[1]} : {a[2]}"
        else:  # pragma: no cover
            raise TraceError(f"unknown op {op}")
        v = f"v{x.idx}"
        lines.append(f"{indent}const float {v} = {expr};")
        names[x.idx] = v
    return lines, names, flops


def count_flops(tr: Trace) -> int:
    return _emit_body(tr.g, tr.outs, "du")[1]


# sqrt(eps(Float32)) = 2^-11.5 rounded to Float32: the finite-difference step of alg_utils.jl:18
# and nlsolve/type.jl:152
FD_STEP_BITS = 0x39B504F3


def _fd_jac_lines(n: int, cols: int | None = None):
    """deg_jac by forward differences (finite_diff_jac, alg_utils.jl:17-27): column i is
    (f(u + dx e_i) - f(u)) / dx, with the perturbed entry rounded to Float32 and a true division.
    ``cols``: only the first cols columns (the values of a dual-number state), J[r * cols + i]."""
    m = n if cols is None else cols
    return [
        "__device__ __forceinline__ void deg_jac(const float* __restrict__ u, const float* __restrict__ p, float t,",
        "                                        float* __restrict__ J) {",
        f"    const float dx = __builtin_bit_cast(float, 0x{FD_STEP_BITS:08x}u);",
        f"    float f0[{n}], f1[{n}], ux[{n}];",
        "    deg_rhs(u, p, t, f0);",
        f"    _Pragma(\"unroll\") for (int i = 0; i < {m}; ++i) {{",
        f"        _Pragma(\"unroll\") for (int r = 0; r < {n}; ++r) ux[r] = u[r];",
        "        ux[i] = u[i] + dx;",
        "        deg_rhs(ux, p, t, f1);",
        f"        _Pragma(\"unroll\") for (int r = 0; r < {n}; ++r) J[r * {m} + i] = (f1[r] - f0[r]) / dx;",
        "    }",
        "}"]


def _fd_tgrad_lines(n: int):
    """deg_tgrad by a forward difference in t (nlsolve/type.jl:150-154): (f(u, p, t + dt) - f(u, p, t)) / dt
    with dt = sqrt(eps(Float32))."""
    return [
        "__device__ __forceinline__ void deg_tgrad(const float* __restrict__ u, const float* __restrict__ p, float t,",
        "                                          float* __restrict__ dT) {",
        f"    const float dx = __builtin_bit_cast(float, 0x{FD_STEP_BITS:08x}u);",
        f"    float f0[{n}], f1[{n}];",
        "    deg_rhs(u, p, t, f0);",
        "    deg_rhs(u, p, t + dx, f1);",
        f"    _Pragma(\"unroll\") for (int r = 0; r < {n}; ++r) dT[r] = (f1[r] - f0[r]) / dx;",
        "}"]


def emit_rhs(f: Callable, n: int, np_: int, jac: Callable | None = None, tgrad: Callable | None = None,
             with_jac: bool = False, fd: bool = False) -> Tuple[str, int]:
    """HIP C source defining deg_rhs (and deg_jac/deg_tgrad when with_jac) plus the RHS flop count.

    The Jacobian and time gradient come from ``jac``/``tgrad`` when given, else from forward-mode
    differentiation of the traced DAG, or with ``fd`` (``autodiff=False``) from forward differences
    of deg_rhs, in the order build_J_W / build_tgrad choose them (nlsolve/type.jl:129-157)."""
    tr = trace(f, n, np_, nout=n)
    body, flops = _emit_body(tr.g, tr.outs, "du")
    src = ["__device__ __forceinline__ void deg_rhs(const float* __restrict__ u, const float* __restrict__ p, float t,",
           "                                        float* __restrict__ du) {",
           "    (void)u; (void)p; (void)t;", *body, "}"]
    if with_jac and fd:
        if jac is not None:
            g, u, p, t = _new_inputs(n, np_)
            out = jac(SymVec(u, "u"), SymVec(p, "p"), t)
            rows = _flatten_matrix(out, n)
            jt = Trace(g, u, p, t, [g.lift(x) for x in rows])
            jbody, _ = _emit_body(jt.g, jt.outs, "J")
            src += ["__device__ __forceinline__ void deg_jac(const float* __restrict__ u, const float* __restrict__ p, float t,",
                    "                                        float* __restrict__ J) {",
                    "    (void)u; (void)p; (void)t;", *jbody, "}"]
        else:
            src += _fd_jac_lines(n)
        if tgrad is not None:
            tt = trace(tgrad, n, np_, nout=n)
            tbody, _ = _emit_body(tt.g, tt.outs, "dT")
            src += ["__device__ __forceinline__ void deg_tgrad(const float* __restrict__ u, const float* __restrict__ p, float t,",
                    "                                          float* __restrict__ dT) {",
                    "    (void)u; (void)p; (void)t;", *tbody, "}"]
        else:
            src += _fd_tgrad_lines(n)
    elif with_jac:
        if jac is not None:
            g, u, p, t = _new_inputs(n, np_)
            out = jac(SymVec(u, "u"), SymVec(p, "p"), t)
            rows = _flatten_matrix(out, n)
            jt = Trace(g, u, p, t, [g.lift(x) for x in rows])
            jbody, _ = _emit_body(jt.g, jt.outs, "J")
        else:
            J, _T = jacobian_nodes(tr)
            jbody, _ = _emit_body(tr.g, [x for row in J for x in row], "J")
        if tgrad is not None:
            tt = trace(tgrad, n, np_, nout=n)
            tbody, _ = _emit_body(tt.g, tt.outs, "dT")
        else:
            _J, T = jacobian_nodes(tr)
            tbody, _ = _emit_body(tr.g, T, "dT")
        src += ["__device__ __forceinline__ void deg_jac(const float* __restrict__ u, const float* __restrict__ p, float t,",
                "                                        float* __restrict__ J) {",
                "    (void)u; (void)p; (void)t;", *jbody, "}",
                "__device__ __forceinline__ void deg_tgrad(const float* __restrict__ u, const float* __restrict__ p, float t,",
                "                                          float* __restrict__ dT) {",
                "    (void)u; (void)p; (void)t;", *tbody, "}"]
    return "\n".join(src) + "\n", flops


def emit_noise(f: Callable, g: Callable, n: int, np_: int, m: Optional[int] = None) -> Tuple[str, int]:
    """HIP C source for an SDE: the drift deg_rhs, the diffusion deg_noise and the noise-shape
    macros DEG_NM / DEG_DIAG_NOISE, plus the flop count of one f and one g evaluation.

    ``m`` None is diagonal noise: g returns n entries (multiplied elementwise with the n Wiener
    increments). Otherwise g returns the n x m noise matrix, as n rows of m entries or a flat
    row-major sequence (the product with the m increments sums over columns in order)."""
    src, flops = emit_rhs(f, n, np_)
    diag = m is None
    mm = n if diag else int(m)
    if mm < 1:
        raise TraceError("the noise matrix needs at least one column")
    g_, u, p, t = _new_inputs(n, np_)
    out = g(SymVec(u, "u"), SymVec(p, "p"), t)
    if isinstance(out, (int, float, Node)):
        out = (out,)
    if diag:
        vals = list(out)
        if len(vals) != n:
            raise TraceError(f"diagonal noise: g must return {n} entries, got {len(vals)}")
    else:
        if len(out) == n and all(isinstance(r, (list, tuple)) for r in out):
            if any(len(r) != mm for r in out):
                raise TraceError(f"g must return {n} rows of {mm} entries")
            vals = [x for r in out for x in r]
        else:
            vals = list(out)
        if len(vals) != n * mm:
            raise TraceError(f"g must return {n}x{mm} entries, got {len(vals)}")
    try:
        gouts = [g_.lift(x) for x in vals]
    except TypeError as e:
        raise TraceError(f"g must return scalars: {e}") from e
    gbody, gflops = _emit_body(g_, gouts, "g")
    head = [f"#define DEG_NM {mm}", f"#define DEG_DIAG_NOISE {1 if diag else 0}"]
    noise = ["__device__ __forceinline__ void deg_noise(const float* __restrict__ u, const float* __restrict__ p, float t,",
             "                                          float* __restrict__ g) {",
             "    (void)u; (void)p; (void)t;", *gbody, "}"]
    return "\n".join(head) + "\n" + src + "\n".join(noise) + "\n", flops + gflops


def emit_mass(M, n: int) -> str:
    """The constant mass matrix as compile-time entries: DEG_MASS and DEG_MASS_ENTRIES (row-major
    Float32 literals, bit-exact)."""
    import numpy as np

    m = np.asarray(M, dtype=np.float32)
    if m.shape != (n, n):
        raise TraceError(f"mass_matrix must be {n}x{n}, got {m.shape}")
    ents = ", ".join(f"__builtin_bit_cast(float, 0x{int(v):08x}u)" for v in m.reshape(-1).view(np.uint32))
    return f"#define DEG_MASS 1\n#define DEG_MASS_ENTRIES {ents}\n"


def _flatten_matrix(out, n):
    """Accept a nested row list [[J11, J12, ...], ...] or a flat row-major sequence."""
    if len(out) == n and all(isinstance(r, (list, tuple)) for r in out):
        rows = [x for r in out for x in r]
    else:
        rows = list(out)
    if len(rows) != n * n:
        raise TraceError(f"jac must return {n}x{n} entries, got {len(rows)}")
    return rows


def evaluate(f: Callable, u, p, t):
    """Evaluate an RHS written for the tracer on plain numbers (float64), for host checks."""
    out = f(list(u), list(p), t)
    return [float(x) for x in out]


def flop_counts(f: Callable, n: int, np_: int, jac: Callable | None = None, tgrad: Callable | None = None) -> dict:
    """Flops of one RHS evaluation (F), of the Jacobian (FJ) and of the time gradient (FT), as
    emitted (common subexpressions counted once, products of the Jacobian's own DAG only)."""
    tr = trace(f, n, np_, nout=n)
    F = _emit_body(tr.g, tr.outs, "du")[1]
    J, T = jacobian_nodes(tr)
    if jac is not None:
        g, u, p, t = _new_inputs(n, np_)
        rows = _flatten_matrix(jac(SymVec(u, "u"), SymVec(p, "p"), t), n)
        jt = Trace(g, u, p, t, [g.lift(x) for x in rows])
        FJ = _emit_body(jt.g, jt.outs, "J")[1]
    else:
        FJ = _emit_body(tr.g, [x for row in J for x in row], "J")[1]
    if tgrad is not None:
        tt = trace(tgrad, n, np_, nout=n)
        FT = _emit_body(tt.g, tt.outs, "dT")[1]
    else:
        FT = _emit_body(tr.g, T, "dT")[1]
    return dict(F=F, FJ=FJ, FT=FT)
