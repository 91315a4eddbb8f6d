"""Forward-mode dual numbers in the initial state and parameters of an ensemble problem.

The reference's kernels accept an ``ODEProblem`` whose ``u0`` and ``p`` hold
``ForwardDiff.Dual`` numbers (test/gpu_kernel_de/forward_diff.jl): every operation of the step
then carries the value and the partial derivatives with respect to the seeded inputs, so the
solution holds d u(t) / d (u0, p) of the discrete method alongside u(t).

Here a problem with dual inputs is solved as an augmented Float32 system. Its state is the
values followed by the partials, row-major: ``[v_0 .. v_{n-1}, d_{0,0} .. d_{0,P-1}, d_{1,0} ..]``,
and likewise for the parameters. The right-hand side is the traced dual evaluation of ``f``
(:func:`diffeqgpu_amd.tracer.emit_dual_rhs`). Linear step arithmetic acts on values and partials
alike. The adaptive error norm is the dual one: the scaled error ``e ./ (abstol .+ max.(|uprev|,
|u|) * reltol)`` is formed in dual arithmetic, and the norm is the root mean square over all
values and partials (DiffEqBase's ``ODE_DEFAULT_NORM`` for dual arrays).
"""
from __future__ import annotations

from typing import Optional, Sequence

import numpy as np


class Dual:
    """``ForwardDiff.Dual(value, partials...)`` on the host, in Float32."""

    __slots__ = ("value", "partials")

    def __init__(self, value, partials: Sequence[float] = ()):
        self.value = np.float32(value)
        self.partials = np.asarray(partials, dtype=np.float32).reshape(-1)

    def __repr__(self):
        return f"Dual({self.value}, {tuple(float(x) for x in self.partials)})"

    def __eq__(self, other):
        return (isinstance(other, Dual) and self.value == other.value
                and np.array_equal(self.partials, other.partials))

    __hash__ = None


class DualArray:
    """A vector of dual numbers with a common number of partials: ``values`` (n,) and
    ``partials`` (n, P), Float32."""

    def __init__(self, values, partials):
        self.values = np.asarray(values, dtype=np.float32).reshape(-1)
        self.partials = np.asarray(partials, dtype=np.float32).reshape(len(self.values), -1)

    @property
    def npartials(self) -> int:
        return self.partials.shape[1]

    def __len__(self):
        return len(self.values)

    def __getitem__(self, i):
        return Dual(self.values[i], self.partials[i])

    def augmented(self) -> np.ndarray:
        """Values, then the partials of each entry in turn."""
        return np.concatenate([self.values, self.partials.reshape(-1)]).astype(np.float32)

    def copy(self):
        return DualArray(self.values.copy(), self.partials.copy())

    def __repr__(self):
        return f"DualArray(values={self.values}, partials={self.partials.tolist()})"


def as_dual_array(x) -> Optional[DualArray]:
    """A DualArray for a DualArray or a sequence holding at least one Dual (the plain numbers
    get zero partials); None for anything else."""
    if isinstance(x, DualArray):
        return x
    if x is None or isinstance(x, (str, bytes, np.ndarray)):
        return None
    try:
        items = list(x)
    except TypeError:
        return None
    duals = [v for v in items if isinstance(v, Dual)]
    if not duals:
        return None
    P = len(duals[0].partials)
    if any(len(v.partials) != P for v in duals):
        raise ValueError("all dual numbers of a problem need the same number of partials")
    vals = np.array([v.value if isinstance(v, Dual) else np.float32(v) for v in items], np.float32)
    parts = np.zeros((len(items), P), np.float32)
    for i, v in enumerate(items):
        if isinstance(v, Dual):
            parts[i] = v.partials
    return DualArray(vals, parts)


def with_partials(x: DualArray | np.ndarray, P: int) -> DualArray:
    """x as a DualArray with P partials (plain numbers get zero partials)."""
    if isinstance(x, DualArray):
        if x.npartials != P:
            raise ValueError(f"u0 and p carry different numbers of partials ({x.npartials} and {P})")
        return x
    v = np.asarray(x, np.float32).reshape(-1)
    return DualArray(v, np.zeros((len(v), P), np.float32))


def split_rows(us: np.ndarray, n: int, P: int) -> np.ndarray:
    """Saved rows of an augmented state, (..., n (1 + P)), as (..., n, 1 + P): [..., 0] the values,
    [..., 1:] the partials."""
    vals = us[..., :n]
    parts = us[..., n:].reshape(us.shape[:-1] + (n, P))
    return np.concatenate([vals[..., None], parts], axis=-1)
