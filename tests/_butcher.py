"""Rooted trees and Runge-Kutta order conditions (Butcher): sum_i b_i Phi_i(t) = 1/gamma(t)."""
from functools import lru_cache

import numpy as np


@lru_cache(maxsize=None)
def trees(order):
    """All rooted trees with `order` vertices; a tree is a sorted tuple of child trees."""
    if order == 1:
        return ((),)
    out = set()
    for parts in _partitions(order - 1):
        for combo in _forests(parts):
            out.add(tuple(sorted(combo)))
    return tuple(sorted(out))


def _partitions(n, maxpart=None):
    if maxpart is None:
        maxpart = n
    if n == 0:
        yield ()
        return
    for k in range(min(n, maxpart), 0, -1):
        for rest in _partitions(n - k, k):
            yield (k,) + rest


def _forests(parts):
    if not parts:
        yield ()
        return
    k = parts[0]
    for t in trees(k):
        for rest in _forests(parts[1:]):
            yield (t,) + rest


def size(t):
    return 1 + sum(size(c) for c in t)


def gamma(t):
    g = size(t)
    for c in t:
        g *= gamma(c)
    return g


def phi(A, t):
    """Stage vector g(t): product over children of A g(child); ones for a leaf."""
    v = np.ones(A.shape[0])
    for c in t:
        v = v * (A @ phi(A, c))
    return v


def order_residuals(A, b, p):
    """max |b . Phi(t) - 1/gamma(t)| over trees of each order 1..p."""
    res = {}
    for q in range(1, p + 1):
        res[q] = max(abs(b @ phi(A, t) - 1.0 / gamma(t)) for t in trees(q))
    return res
