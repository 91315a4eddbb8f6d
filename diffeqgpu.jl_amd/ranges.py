"""Julia range semantics needed by the ensemble front end.

``vectorized_solve`` sizes its output with ``length(t0:dt:tf)`` (lowerlevel_solve.jl:64-73)
and both lower-level entry points normalise ``saveat`` through ``range(t0, tf, length=k)``
(lowerlevel_solve.jl:89-102, :271-299). Those lengths and values follow Julia Base's float
ranges, which look for exact rational representations of the endpoints (``Base.rat``)
before falling back to the literal step. This module restates that logic; elements are
the correctly rounded values of the exact rational sequence.
"""
from __future__ import annotations

import math
from fractions import Fraction

import numpy as np

_DT = {np.float32: (np.float32, 2048, 16777216), np.float64: (np.float64, 16777216, 9007199254740992)}


def _cfg(T):
    T = np.dtype(T).type
    if T not in _DT:
        raise TypeError("float32 or float64 ranges only")
    return _DT[T]


def rat(x, T=np.float32):
    """Base.rat: continued-fraction rational approximation computed in type T."""
    T, m_narrow, _ = _cfg(T)
    x = T(x)
    y = x
    a = d = 1
    b = c = 0
    m = m_narrow
    while abs(y) <= m:
        f = int(math.trunc(float(y)))
        y = T(y - T(f))
        a, c = f * a + c, a
        b, d = f * b + d, b
        if not max(abs(a), abs(b)) <= m:
            return c, d
        if b != 0 and T(T(a) / T(b)) == x:
            break
        if y == 0:
            break
        y = T(T(1) / y)
    return a, b


def round_to(fr: Fraction, T=np.float32):
    """Correctly rounded (nearest, ties to even) value of an exact rational in type T."""
    T = np.dtype(T).type
    v = T(float(fr))
    if not np.isfinite(v):
        return v
    best = v
    for cand in (np.nextafter(v, T(-np.inf)), np.nextafter(v, T(np.inf))):
        dc = abs(Fraction(float(cand)) - fr)
        db = abs(Fraction(float(best)) - fr)
        if dc < db:
            best = cand
        elif dc == db:
            # ties to even mantissa
            bits_c = np.array(cand, dtype=T).view(np.uint32 if T is np.float32 else np.uint64)
            if int(bits_c) % 2 == 0:
                best = cand
    return T(best)


def _isbetween(a, x, b):
    return (a <= x <= b) or (b <= x <= a)


class JlRange:
    """A Julia float/int range: ``first``, ``step`` (may be None for linspace), ``len``."""

    def __init__(self, values: np.ndarray):
        self._v = values

    @property
    def first(self):
        return self._v[0]

    @property
    def last(self):
        return self._v[-1]

    def __len__(self):
        return int(self._v.shape[0])

    def collect(self):
        return self._v.copy()

    def __array__(self, dtype=None, copy=None):
        return self._v if dtype is None else self._v.astype(dtype)

    def __repr__(self):
        return f"JlRange({self._v!r})"


def colon(start, step, stop, T=np.float32) -> JlRange:
    """``start:step:stop`` for IEEE floats (Base (:)(start::T, step::T, stop::T))."""
    T, _, maxint = _cfg(T)
    start, step, stop = T(start), T(step), T(stop)
    if step == 0:
        raise ValueError("range step cannot be zero")
    step_n, step_d = rat(step, T)
    if step_d != 0 and T(step_n / step_d) == step:
        start_n, start_d = rat(start, T)
        stop_n, stop_d = rat(stop, T)
        if start_d != 0 and stop_d != 0 and T(start_n / start_d) == start and T(stop_n / stop_d) == stop:
            den = start_d * step_d // math.gcd(start_d, step_d)
            if den != 0 and abs(T(start * T(den))) <= maxint and abs(T(step * T(den))) <= maxint and \
                    den % start_d == 0 and den % step_d == 0:
                s_n = int(round(float(T(start * T(den)))))
                st_n = int(round(float(T(step * T(den)))))
                num = den * stop_n - stop_d * s_n + st_n * stop_d
                dd = st_n * stop_d
                q = abs(num) // abs(dd)
                if (num < 0) != (dd < 0) and q != 0:
                    q = -q
                n = max(0, int(q))
                if _isbetween(start, T(start + T(n - 1) * step), T(stop + step / T(2))) and \
                        not _isbetween(start, T(start + T(n) * step), stop):
                    vals = np.array([round_to(Fraction(s_n + k * st_n, den), T) for k in range(n)], dtype=T)
                    return JlRange(vals)
    lf = T((stop - start) / step)
    if lf < 0:
        n = 0
    elif lf == 0:
        n = 1
    else:
        n = int(np.rint(lf)) + 1
        stop2 = T(start + T(n - 1) * step)
        n -= int(start < stop < stop2) + int(start > stop > stop2)
    fs, fst = Fraction(float(start)), Fraction(float(step))
    vals = np.array([round_to(fs + k * fst, T) for k in range(n)], dtype=T)
    return JlRange(vals)


def linrange(start, stop, length: int, T=np.float32) -> JlRange:
    """``range(start, stop, length = n)`` for IEEE floats: correctly rounded points of the
    exact rational sequence start + k (stop - start) / (n - 1)."""
    T, _, _ = _cfg(T)
    start, stop = T(start), T(stop)
    n = int(length)
    if n < 0:
        raise ValueError("length must be non-negative")
    if n == 0:
        return JlRange(np.zeros(0, T))
    if n == 1:
        if start != stop:
            raise ValueError("range(start, stop, length=1) needs start == stop")
        return JlRange(np.array([start], T))
    a, b = Fraction(float(start)), Fraction(float(stop))
    vals = np.array([round_to(a + (b - a) * k / (n - 1), T) for k in range(n)], dtype=T)
    vals[0], vals[-1] = start, stop
    return JlRange(vals)


def julia_range_length(start, step, stop, T=np.float32) -> int:
    return len(colon(start, step, stop, T))
