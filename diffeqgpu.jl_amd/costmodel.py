"""A fitted cost model for the start order of persistent grids.

The persistent kernel's tail -- lanes idling while the wave's last trajectories finish -- is
shortest when the longest trajectories start first. ``deg_start_order`` orders by span / h0 (the
standard initial step estimate); this module fits a better predictor of each trajectory's attempt
count from a pilot ensemble (a held-out draw of the same problem, or a small subset solved first),
and ``HIPDevice(start_order="model", cost_model=w)`` starts the solve in descending order of the
prediction (``deg_start_order_model``, computed on the device before every launch).

The model is linear in log2 features, log2(attempts) ~ w . x with x = [1, log2(span / h0),
log2(|u0_c| + 2^-20) for each state c, log2(|p_j| + 2^-20) for each parameter j, log2 span]: the same
features the device computes, fitted by ridge least squares on the pilot's attempt counts. Results
do not depend on the order; only the time does. The reference starts trajectories in index order
(one thread per trajectory, src/ensemblegpukernel/kernels.jl:74-152) and has no counterpart.
"""
from __future__ import annotations

import copy
import dataclasses

import numpy as np

TINY = 2.0 ** -20


def features(key, u0, p, span) -> np.ndarray:
    """(N, 3 + n + np) float64 design matrix of the device's features."""
    key = np.asarray(key, np.float64)
    u0 = np.asarray(u0, np.float64).reshape(-1, key.size)
    p = np.asarray(p, np.float64).reshape(-1, key.size)
    span = np.broadcast_to(np.asarray(span, np.float64), key.shape)
    cols = [np.ones_like(key), np.log2(np.maximum(key, 1e-30))]
    cols += [np.log2(np.abs(u0[c]) + TINY) for c in range(u0.shape[0])]
    cols += [np.log2(np.abs(p[j]) + TINY) for j in range(p.shape[0])]
    cols.append(np.log2(np.maximum(span, 1e-30)))
    return np.stack(cols, axis=1)


def fit_arrays(key, u0, p, span, attempts, ridge: float = 1e-6) -> tuple:
    """Ridge least squares of log2(attempts) on the features (the intercept is not penalised)."""
    X = features(key, u0, p, span)
    y = np.log2(np.maximum(np.asarray(attempts, np.float64), 1.0))
    # features constant over the pilot carry no information (the range is exact for a constant column;
    # the standard deviation of many equal values is a rounding residue, not zero)
    keep = np.ptp(X, axis=0) > 1e-9 * (1.0 + np.abs(X).max(axis=0))
    keep[0] = True
    sd = np.where(keep, X.std(axis=0), 1.0)
    sd[0] = 1.0
    Xs = X[:, keep] / np.where(keep, sd, 1.0)[keep]
    reg = ridge * Xs.shape[0] * np.eye(Xs.shape[1])
    reg[0, 0] = 0.0
    ws = np.linalg.solve(Xs.T @ Xs + reg, Xs.T @ y)
    w = np.zeros(X.shape[1])
    w[keep] = ws / np.where(keep, sd, 1.0)[keep]
    return tuple(float(np.float32(v)) for v in w)


def fit(probs, prob, alg, **solve_kw) -> tuple:
    """Fits the cost model on a pilot ensemble: one solve of ``probs`` (its attempt counts) and the
    span / h0 keys its cost-ordered start computes. Returns the 3 + n + np weights."""
    import torch

    from .lowerlevel import vectorized_asolve

    dev = dataclasses.replace(probs.dev, start_order="cost", persistent=True, cost_model=None)
    pilot = copy.copy(probs)
    pilot.dev = dev
    res = vectorized_asolve(pilot, prob, alg, **solve_kw)
    torch.cuda.synchronize()
    att = (res.naccept.to(torch.int64) + res.nreject.to(torch.int64)).cpu().numpy()
    key = res.order_key.cpu().numpy()
    u0 = probs.u0.cpu().numpy()
    p = probs.p.cpu().numpy() if probs.np > 0 else np.zeros((0, probs.N), np.float32)
    t0 = probs.t0s.cpu().numpy() if probs.t0s is not None else np.float32(probs.tspan[0])
    tf = probs.tfs.cpu().numpy() if probs.tfs is not None else np.float32(probs.tspan[1])
    span = np.abs(np.asarray(tf, np.float64) - np.asarray(t0, np.float64))
    return fit_arrays(key, u0, p, span, att)
