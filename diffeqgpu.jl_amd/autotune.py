"""Launch-configuration autotuner.

Picks, per (right-hand side, method, solver flags, ensemble size), the configuration with the
lowest measured kernel time:
  block size       threads per block (64, 128, 192, 256);
  register cap     DEG_F_WAVES: the solver compiled for at least w waves per SIMD (0 = the
                   compiler's own allocation, 4..7 = capped, possibly spilling);
  refill batch     persistent grids: idle lanes per wave that trigger a refill;
  chunk            persistent grids: trajectory ids claimed per queue atomic;
  blocks per CU    persistent grids: resident blocks per compute unit (0 = as many as fit);
  start order      persistent adaptive grids: index order, the cost order of deg_start_order, or (when the
                   device carries a fitted cost model) the order of deg_start_order_model
                   (which helps where the initial step estimate predicts the work, and can hurt
                   where it points the wrong way).
The LDS depth of the saveat cache needs no tuning: it is sized to the save grid at every launch.

The search is coordinate descent, starting from the defaults: the start order first (it changes
which launch shape is best), then the other axes in the order above, then the start order again; the
whole sweep runs twice, because the best register cap depends on the refill batch and chunk chosen
after it (a second sweep re-times only new candidates).
Each candidate is
timed with HIP events around full relaunches of the solve (start-order kernels included; the best
of ``repeats``) on the caller's problems, so the result reflects the real trajectory mix. Results are cached in
``_jitcache/autotune.json``, keyed by the traced source, n, np, method, flags and the ensemble
size rounded to a power of two. ``autotune_mi355x.json`` beside this file ships the results measured
on MI355X for the benchmark workloads; the local cache takes precedence over it. The reference has
no tuner: KernelAbstractions picks its own workgroup size.
"""
from __future__ import annotations

import copy
import dataclasses
import hashlib
import json
import math
import os

from . import _lib
from .algorithms import GPUODEImplicitAlgorithm, HIPDevice

CACHE_PATH = os.path.join(_lib.CACHE_DIR, "autotune.json")
SHIPPED_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "autotune_mi355x.json")
BLOCKS = (64, 128, 192, 256)
WAVES = (0, 4, 5, 6, 7)
REFILL = (1, 2, 4, 8, 16)
CHUNK = (4, 16, 64)
WAVES_PER_CU = (8, 12, 16)   # blocks-per-CU candidates are these resident waves per CU


def blocks_per_cu_candidates(block: int):
    return (0,) + tuple(max(1, w * 64 // block) for w in WAVES_PER_CU)


def _key(probs, alg, adaptive, nsave_hint, learned: bool = False) -> str:
    implicit = isinstance(alg, GPUODEImplicitAlgorithm)
    src, _ = probs.f.hip_source(probs.n, probs.np, implicit, fd=implicit and not alg.autodiff)
    h = hashlib.sha1(src.encode()).hexdigest()[:16]
    base = dataclasses.replace(probs.dev, block_size=None, refill=0, chunk=0, waves=0, blocks_per_cu=0,
                               start_order=None)
    nb = int(round(math.log2(max(1, probs.N))))
    flags = base.flags(adaptive) | (_lib.F_LEARNED if learned else 0)
    # a device that carries a cost model can also be tuned to start in its order: a separate entry
    model = ":model" if probs.dev.cost_model and adaptive and bool(base.flags(adaptive) & _lib.F_PERSISTENT) else ""
    return f"{h}:{probs.n}:{probs.np}:{alg.method}:{flags}:{nb}:{nsave_hint}{model}"


def _read(path) -> dict:
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _load() -> dict:
    """The local cache."""
    return _read(CACHE_PATH)


def lookup(key: str):
    """A tuned configuration: the local cache first, then the shipped MI355X table."""
    for table in (_load(), _read(SHIPPED_PATH)):
        if key in table:
            return table[key]["config"]
    return None


def _store(table: dict):
    os.makedirs(os.path.dirname(CACHE_PATH), exist_ok=True)
    tmp = CACHE_PATH + f".tmp{os.getpid()}"
    with open(tmp, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
    os.replace(tmp, CACHE_PATH)


def _with_dev(probs, dev):
    q = copy.copy(probs)
    q.dev = dev
    return q


def tune(probs, prob, alg, *, adaptive: bool = True, repeats: int = 2, use_cache: bool = True, refresh: bool = False,
         log=None, **solve_kw) -> HIPDevice:
    """Returns the fastest HIPDevice for this ensemble (also stored in the cache). ``refresh``
    measures again even when a tuned entry exists."""
    import torch

    from .lowerlevel import vectorized_asolve, vectorized_solve

    fn = vectorized_asolve if adaptive else vectorized_solve
    nsave_hint = "saveat" if solve_kw.get("saveat") is not None else ("every" if solve_kw.get("save_everystep") else "end")
    key = _key(probs, alg, adaptive, nsave_hint, learned=solve_kw.get("mlp") is not None)
    if use_cache and not refresh:
        hit = lookup(key)
        if hit is not None:
            return dataclasses.replace(probs.dev, **hit)

    persistent = bool(probs.dev.flags(adaptive) & _lib.F_PERSISTENT)
    trials = {}

    def measure(cfg):
        k = tuple(sorted(cfg.items()))
        if k not in trials:
            dev = dataclasses.replace(probs.dev, **cfg)
            q = _with_dev(probs, dev)
            r = fn(q, prob, alg, **solve_kw)  # the first launch also loads (or JIT-compiles) the variant
            st = torch.cuda.current_stream(q.u0.device)
            best = math.inf
            for _ in range(repeats):
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record(st)
                r.relaunch()
                e1.record(st)
                e1.synchronize()
                best = min(best, e0.elapsed_time(e1))
            trials[k] = best
            if log:
                log(f"[autotune] {cfg} {best:.3f} ms")
        return trials[k]

    default = dict(block_size=256, waves=0, refill=4 if persistent else 0, chunk=16 if persistent else 0,
                   blocks_per_cu=0)
    if persistent and adaptive:
        default["start_order"] = "none"
    cfg = dict(default)
    orders = ("none", "cost", "model") if probs.dev.cost_model else ("none", "cost")
    axes = []
    if persistent and adaptive:
        # the start order changes which launch shape is best: chosen first, and checked again last
        axes += [("start_order", orders)]
    axes += [("block_size", BLOCKS), ("waves", WAVES)]
    if persistent:
        axes += [("refill", REFILL), ("chunk", CHUNK), ("blocks_per_cu", None)]
    if persistent and adaptive:
        axes += [("start_order", orders)]
    # two sweeps over the axes: the best register cap, for one, depends on the refill batch and chunk
    # chosen after it (repeated candidates are timed once)
    for _sweep in range(2):
        for name, values in axes:
            if values is None:
                values = blocks_per_cu_candidates(cfg["block_size"])
            best_v, best_t = cfg[name], measure(cfg)
            for v in values:
                c = dict(cfg, **{name: v})
                t = measure(c)
                if t < best_t:
                    best_v, best_t = v, t
            cfg[name] = best_v
    table = _load() if use_cache else {}
    table[key] = dict(config=cfg, kernel_ms=trials[tuple(sorted(cfg.items()))],
                      default_ms=trials.get(tuple(sorted(default.items()))), trials=len(trials))
    if use_cache:
        _store(table)
    return dataclasses.replace(probs.dev, **cfg)
