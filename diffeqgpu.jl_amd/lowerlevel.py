"""Lower-level API: ``vectorized_solve`` / ``vectorized_asolve``.

Host half of src/ensemblegpukernel/lowerlevel_solve.jl:53-131 (fixed step), :134-199 (fixed-step
SDE) and :253-356 (adaptive; none for SDEs): saveat normalisation, output sizing and the t0
prefill of ``ts``. The device half is one call of ``deg_solve`` in libdeghip.so.

Layout. The device writes structure-of-arrays: ``ts`` is a (nsave, N) tensor and the state
output is stored as (nsave, n, N). ``us`` is returned as the (nsave, N, n) view of that
storage, so ``us[s, i]`` is the n-vector of trajectory i at save point s, the same indexing as
the reference's ``us[s, i]::SVector`` of its (nsave, N) matrix.
"""
from __future__ import annotations

import math
from typing import Optional

import numpy as np

from . import _lib
from .algorithms import GPUODEAlgorithm, GPUODEImplicitAlgorithm, GPUSDEAlgorithm, GPUSIEA, HIPDevice
from .callbacks import emit_callbacks, gpu_callback_set
from .problems import HIPProbVector, ODEProblem, get_backend
from .ranges import JlRange, colon, linrange

MAX_SAVEAT_LENGTH = 100_000  # lowerlevel_solve.jl:284-290


class SolveResult(tuple):
    """``(ts, us)`` with per-trajectory statistics attached.

    Attributes: ``retcode`` (int32, 0 Success / 1 DtLessThanMin / 2 MaxIters / 3 Unstable),
    ``naccept``, ``nreject`` (int32 per trajectory), ``us_soa`` ((nsave, n, N) storage),
    ``kernel_ms`` (float or None), ``saveat`` (normalised device tensor or None)."""

    def __new__(cls, ts, us, **attrs):
        obj = super().__new__(cls, (ts, us))
        for k, v in attrs.items():
            setattr(obj, k, v)
        return obj

    @property
    def ts(self):
        return self[0]

    @property
    def us(self):
        return self[1]


def _is_range(x) -> bool:
    return isinstance(x, (range, JlRange))


def _to_T(x):
    return np.asarray(x, dtype=np.float64).astype(np.float32)


def normalize_saveat_solve(saveat, tspan) -> Optional[np.ndarray]:
    """lowerlevel_solve.jl:89-102 (fixed-step entry point)."""
    if saveat is None:
        return None
    t0, tf = np.float32(tspan[0]), np.float32(tspan[1])
    if _is_range(saveat):
        vals = np.asarray(list(saveat) if isinstance(saveat, range) else saveat.collect(), dtype=np.float64)
        if len(vals) == 0:
            return np.zeros(0, np.float32)
        return linrange(np.float32(vals[0]), np.float32(vals[-1]), len(vals)).collect()
    if np.ndim(saveat) >= 1:
        return _to_T(saveat).reshape(-1)
    s = np.float32(saveat)
    if s == np.float32(0.0):
        return np.array([t0, tf], np.float32)
    num = int(math.ceil(float(np.float32(abs(np.float32(tf - t0)) / abs(s))))) + 1
    return linrange(t0, tf, num).collect()


def normalize_saveat_asolve(saveat, tspan) -> Optional[np.ndarray]:
    """lowerlevel_solve.jl:271-306 (adaptive entry point)."""
    if saveat is None:
        return None
    t0, tf = np.float32(tspan[0]), np.float32(tspan[1])
    if np.ndim(saveat) == 0 and not _is_range(saveat):
        s = np.float32(saveat)
        if s == np.float32(0.0):
            return np.array([t0, tf], np.float32)
        num = int(math.ceil(float(np.float32(abs(np.float32(tf - t0)) / abs(s))))) + 1
        if num > MAX_SAVEAT_LENGTH:
            raise ValueError(f"saveat would create too many save points ({num}). "
                             "Consider using a larger saveat value.")
        return linrange(t0, tf, num).collect()
    if isinstance(saveat, range):
        return _to_T(list(saveat))
    if isinstance(saveat, JlRange):
        return saveat.collect().astype(np.float32)
    return _to_T(saveat).reshape(-1)


def _rep_tspan(prob) -> tuple:
    if isinstance(prob, ODEProblem) or hasattr(prob, "tspan"):
        return (np.float32(prob.tspan[0]), np.float32(prob.tspan[1]))
    raise TypeError("prob must be an ODEProblem or SDEProblem")


def check_mass(f, alg):
    """A mass matrix is solved by the implicit methods only (the explicit kernels have no W)."""
    if getattr(f, "mass_matrix", None) is not None and not isinstance(alg, GPUODEImplicitAlgorithm):
        raise ValueError(f"a mass matrix needs an implicit method (GPURosenbrock23, GPURodas4, GPURodas5P, "
                         f"GPUKvaerno3, GPUKvaerno5), not {type(alg).__name__}")


def _callback_set(callback):
    """The GPU CallbackSet for the kernel (SciML callbacks are converted as solve.jl:386-397
    does), or None when there are no callbacks."""
    if callback is None:
        return None
    cs = gpu_callback_set(callback)
    return cs if len(cs) else None


def _tstops(tstops) -> Optional[np.ndarray]:
    """``tstops`` as the Float32 vector the kernel walks in order, or None when empty."""
    if tstops is None:
        return None
    a = np.asarray(list(tstops) if isinstance(tstops, (range, JlRange)) else tstops, dtype=np.float64)
    a = a.astype(np.float32).reshape(-1)
    return a if a.size else None


def _launch(probs: HIPProbVector, alg, adaptive: bool, *, dt, abstol, reltol, saveat_np,
            nsave, save_everystep, t0_fill, timing=False, stream=None, mlp=None, steplog=None, block_size=None,
            grid_blocks=0, outputs=None, order=None, mlp_every=1, callback=None, tstops_np=None, traj_base=0):
    """Launch one solve. ``outputs`` (optional) supplies preallocated device tensors
    {"ts", "us_soa", "retcode", "naccept", "nreject", "saveat"} so that the call allocates and
    copies nothing (required inside a graph capture). ``order`` (optional, int32 device tensor)
    is the start order of the persistent queue, e.g. from ``order_by_span``. ``callback`` is a
    CallbackSet of GPU callbacks or None; ``tstops_np`` a Float32 vector or None. For an SDE
    batch, ``traj_base`` is the ensemble index of its first trajectory (its noise stream)."""
    import torch

    sde = isinstance(alg, GPUSDEAlgorithm)
    if not (sde or isinstance(alg, GPUODEAlgorithm)):
        raise TypeError(f"unsupported algorithm {alg!r}")
    if sde != probs.is_sde:
        raise TypeError("SDE algorithms solve SDE problems and ODE algorithms ODE problems")
    dev: HIPDevice = probs.dev
    if sde:
        if adaptive:
            raise NotImplementedError("Adaptive time-stepping is not supported yet with GPUEM.")
        if mlp is not None or steplog is not None or order is not None:
            raise ValueError("the learned proposer, step logs and start orders are ODE-kernel options")
        if isinstance(alg, GPUSIEA) and probs.noise_m is not None:
            raise ValueError("The algorithm is not compatible with the chosen noise type. Please see the "
                             "documentation on the solver methods")
        if int(traj_base) < 0:
            raise ValueError("traj_base must be >= 0")
        flags = dev.sde_flags()
        src, _flops = probs.f.hip_source(probs.n, probs.np, probs.noise_m)
    else:
        flags = dev.flags(adaptive)
        if mlp is not None:
            flags |= _lib.F_LEARNED
        if steplog is not None:
            flags |= _lib.F_STEPLOG
        if saveat_np is not None and len(saveat_np) > _lib.SAVEAT_LDS_CAP:
            flags |= _lib.F_SAVEAT_GLOBAL
        if tstops_np is not None:
            flags |= _lib.F_TSTOPS
        if callback is not None:
            flags |= _lib.F_CALLBACKS
        if flags & (_lib.F_TSTOPS | _lib.F_CALLBACKS):
            if flags & (_lib.F_LEARNED | _lib.F_STEPLOG | _lib.F_COMPACT | _lib.F_BATCH_SAVES):
                raise ValueError("tstops and callbacks do not combine with the learned proposer, step logs, "
                                 "tail compaction or batched saves")
        with_jac = isinstance(alg, GPUODEImplicitAlgorithm)
        check_mass(probs.f, alg)
        src, _flops = probs.f.hip_source(probs.n, probs.np, with_jac, fd=with_jac and not alg.autodiff)
        if callback is not None:
            src = src + emit_callbacks(callback, probs.n, probs.np)
    ctx = _lib.context(dev.device)
    kern = ctx.kernel(src, probs.n, probs.np, alg.method, flags)
    device = probs.u0.device
    N, n = probs.N, probs.n
    if outputs is None:
        outputs = {}
    ts = outputs.get("ts")
    if ts is None:
        ts = torch.empty((nsave, N), dtype=torch.float32, device=device)
    us_soa = outputs.get("us_soa")
    if us_soa is None:
        us_soa = torch.empty((nsave, n, N), dtype=torch.float32, device=device)
    retcode, nacc, nrej = (outputs.get(k) for k in ("retcode", "naccept", "nreject"))
    if retcode is None:
        retcode = torch.empty(N, dtype=torch.int32, device=device)
    if nacc is None:
        nacc = torch.empty(N, dtype=torch.int32, device=device)
    if nrej is None:
        nrej = torch.empty(N, dtype=torch.int32, device=device)
    if tuple(ts.shape) != (nsave, N) or tuple(us_soa.shape) != (nsave, n, N):
        raise ValueError("preallocated ts/us have the wrong shape")
    for t in (ts, us_soa, retcode, nacc, nrej):
        if not t.is_contiguous() or t.device != device:
            raise ValueError("outputs must be contiguous tensors on the problems' device")
    if retcode.numel() != N or nacc.numel() != N or nrej.numel() != N or retcode.dtype != torch.int32:
        raise ValueError("retcode/naccept/nreject must be int32 tensors of length N")
    sv = None
    if saveat_np is not None:
        sv = outputs.get("saveat")
        if sv is None:
            sv = torch.as_tensor(np.ascontiguousarray(saveat_np, dtype=np.float32)).to(device)
        elif sv.numel() != len(saveat_np):
            raise ValueError("preallocated saveat has the wrong length")
    a = _lib.SolveArgs()
    a.n = n
    a.n_traj = N
    a.stride = N
    a.id_offset = 0
    a.u0 = probs.u0.data_ptr()
    a.p = probs.p.data_ptr() if probs.np > 0 else None
    a.t0s = probs.t0s.data_ptr() if probs.t0s is not None else None
    a.tfs = probs.tfs.data_ptr() if probs.tfs is not None else None
    a.t0 = float(t0_fill)
    a.tf = float(probs.tspan[1])
    a.dt = float(np.float32(dt))
    a.abstol = float(np.float32(abstol))
    a.reltol = float(np.float32(reltol))
    a.saveat = sv.data_ptr() if sv is not None else None
    a.nsaveat = int(sv.numel()) if sv is not None else 0
    a.nsave = int(nsave)
    a.save_everystep = int(bool(save_everystep))
    a.block_size = int(block_size or dev.block_size or 0)
    if not grid_blocks and dev.blocks_per_cu:
        grid_blocks = int(dev.blocks_per_cu) * ctx.num_cus
    a.grid_blocks = int(grid_blocks)
    a.ts = ts.data_ptr()
    a.us = us_soa.data_ptr()
    a.retcode = retcode.data_ptr()
    a.naccept = nacc.data_ptr()
    a.nreject = nrej.data_ptr()
    a.max_attempts = int(dev.max_attempts)
    a.refill_batch = int(dev.refill)
    a.chunk = int(dev.chunk)
    st = stream if stream is not None else torch.cuda.current_stream(device)
    a.stream = st.cuda_stream
    ms = None
    if timing:
        import ctypes

        box = ctypes.c_float(0.0)
        a.elapsed_ms = ctypes.addressof(box)
    if mlp is not None:
        a.mlp = mlp.data_ptr()
        a.mlp_every = int(mlp_every)
    okey = owork = omodel = oinp = None
    if order is None and not sde and dev.cost_order(adaptive) and N > 0:
        # cost-ordered start, recomputed on the device before every launch of this solve
        okey, order, owork = (outputs.get(k) for k in ("order_key", "order_auto", "order_work"))
        if dev.start_order == "model":
            if len(dev.cost_model) != 3 + n + probs.np:
                raise ValueError("cost_model needs 3 + n + np weights for this problem")
            omodel = outputs.get("cost_model")
            if omodel is None:
                omodel = torch.as_tensor(np.asarray(dev.cost_model, np.float32)).to(device)
        if okey is None:
            okey = torch.empty(N, dtype=torch.float32, device=device)
        if order is None:
            order = torch.empty(N, dtype=torch.int32, device=device)
        if owork is None:
            owork = torch.empty(_lib.ORDER_WORK_INTS, dtype=torch.int32, device=device)
        if okey.numel() != N or order.numel() != N or owork.numel() != _lib.ORDER_WORK_INTS:
            raise ValueError("preallocated start-order buffers have the wrong size")
        if dev.ordered_inputs:
            # the ordering pass also lays out u0 and p in the order the queues hand the trajectories out
            ninp = (n + probs.np) * 8 * ((N + 7) // 8)
            oinp = outputs.get("inputs_ordered")
            if oinp is None:
                oinp = torch.empty(ninp, dtype=torch.float32, device=device)
            if oinp.numel() != ninp:
                raise ValueError("preallocated inputs_ordered has the wrong size")
            a.inputs_ordered = oinp.data_ptr()
    if order is not None:
        if order.dtype != torch.int32 or order.numel() != N or order.device != device:
            raise ValueError("order must be an int32 tensor of N trajectory indices on the problems' device")
        a.order = order.data_ptr()
    if steplog is not None:
        buf, counter = steplog
        a.steplog = buf.data_ptr()
        a.steplog_count = counter.data_ptr()
        a.steplog_cap = int(buf.shape[0])
    tst = None
    if tstops_np is not None:
        tst = outputs.get("tstops")
        if tst is None:
            tst = torch.as_tensor(np.ascontiguousarray(tstops_np, dtype=np.float32)).to(device)
        elif tst.numel() != len(tstops_np):
            raise ValueError("preallocated tstops has the wrong length")
        a.tstops = tst.data_ptr()
        a.ntstops = int(tst.numel())
    if sde:
        a.seeds = probs.seeds.data_ptr() if probs.seeds is not None else None
        a.seed = int(probs.seed) & 0xFFFFFFFFFFFFFFFF
        a.traj_base = int(traj_base)

    def order_start():
        if omodel is not None:
            kern.start_order_model(a, omodel.data_ptr(), omodel.numel(), okey.data_ptr(), order.data_ptr(),
                                   owork.data_ptr())
        elif okey is not None:
            kern.start_order(a, okey.data_ptr(), order.data_ptr(), owork.data_ptr())

    order_start()
    kern.solve(a)
    if timing:
        ms = float(box.value)
        a.elapsed_ms = None  # the box does not outlive this call

    def relaunch():
        # same arguments, same output buffers: recomputes the start order and every output
        order_start()
        kern.solve(a)

    # the result keeps the argument block and every tensor it points to alive
    return SolveResult(ts, us_soa.permute(0, 2, 1), retcode=retcode, naccept=nacc, nreject=nrej, us_soa=us_soa,
                       kernel_ms=ms, saveat=sv, kernel=kern, flags=flags, relaunch=relaunch, args=a,
                       probs=probs, tstops=tst, order_key=okey, order_auto=order if okey is not None else None,
                       order_work=owork, cost_model=omodel, inputs_ordered=oinp)


def vectorized_solve(probs, prob, alg, *, dt, saveat=None, save_everystep=True, debug=False, callback=None,
                     tstops=None, timing=False, **kwargs):
    """Fixed-step ensemble solve (lowerlevel_solve.jl:53-131; SDEs :134-199). Returns ``(ts, us)``.

    ``callback``: GPU callbacks (a callback or CallbackSet; SciML callbacks are converted).
    ``tstops``: stopping times shared by all trajectories; each adds a row to the every-step
    output unless it lies on the t0:dt:tf grid (lowerlevel_solve.jl:73-77).
    SDE batches (GPUEM, GPUSIEA) take neither; ``traj_base`` (keyword) numbers their noise
    streams from the ensemble index of the first trajectory."""
    get_backend(probs)
    t0, tf = _rep_tspan(prob)
    dt32 = np.float32(dt)
    if not dt32 > 0:
        raise ValueError("dt must be positive for a fixed-step solve")
    if isinstance(alg, GPUSDEAlgorithm):
        return _vectorized_solve_sde(probs, alg, t0, tf, dt32, saveat, save_everystep, callback, tstops, timing,
                                     **kwargs)
    cbs = _callback_set(callback)
    tst = _tstops(tstops)
    sv = normalize_saveat_solve(saveat, (t0, tf))
    if sv is None:
        if save_everystep:
            grid = colon(t0, dt32, tf)
            nsave = len(grid)
            if tst is not None:
                on_grid = int(np.isin(grid.collect().astype(np.float32), tst).sum())
                nsave += len(tst) - on_grid
        else:
            nsave = 2
    else:
        nsave = len(sv)
    return _launch(probs, alg, False, dt=dt32, abstol=0.0, reltol=0.0, saveat_np=sv, nsave=nsave,
                   save_everystep=save_everystep and sv is None, t0_fill=t0, timing=timing, callback=cbs,
                   tstops_np=tst, **kwargs)


def _vectorized_solve_sde(probs, alg, t0, tf, dt32, saveat, save_everystep, callback, tstops, timing, **kwargs):
    """lowerlevel_solve.jl:134-199: output sizing and saveat normalisation of the SDE kernels."""
    if callback is not None or tstops is not None:
        raise NotImplementedError("callbacks and tstops are not provided for SDE solves")
    sv = normalize_saveat_solve(saveat, (t0, tf))
    if sv is None:
        nsave = len(colon(t0, dt32, tf)) if save_everystep else 2
    else:
        if len(sv) == 0:
            raise ValueError("saveat is empty")
        nsave = len(sv)
    return _launch(probs, alg, False, dt=dt32, abstol=0.0, reltol=0.0, saveat_np=sv, nsave=nsave,
                   save_everystep=save_everystep and sv is None, t0_fill=t0, timing=timing, **kwargs)


def vectorized_asolve(probs, prob, alg, *, dt=0.1, saveat=None, save_everystep=False, abstol=1.0e-6,
                      reltol=1.0e-3, debug=False, callback=None, tstops=None, timing=False, **kwargs):
    """Adaptive ensemble solve (lowerlevel_solve.jl:253-346). Returns ``(ts, us)``.
    ``callback`` and ``tstops`` as for vectorized_solve (tstops add no output rows here).
    SDE methods have no adaptive solve (lowerlevel_solve.jl:349-356)."""
    if isinstance(alg, GPUSDEAlgorithm) or getattr(probs, "is_sde", False):
        raise NotImplementedError("Adaptive time-stepping is not supported yet with GPUEM.")
    get_backend(probs)
    cbs = _callback_set(callback)
    tst = _tstops(tstops)
    t0, tf = _rep_tspan(prob)
    sv = normalize_saveat_asolve(saveat, (t0, tf))
    dt32 = np.float32(dt)
    if sv is None:
        if save_everystep:
            nsave = int(math.ceil(float(np.float32(np.float32(tf - t0) / dt32)))) + 1
        else:
            nsave = 2
    else:
        nsave = len(sv)
    # The adaptive integrator never saves every step (nonstiff/types.jl:340-402 passes
    # save_everystep = false); the kernel only skips the endpoint write in that case.
    return _launch(probs, alg, True, dt=dt32, abstol=abstol, reltol=reltol, saveat_np=sv, nsave=nsave,
                   save_everystep=save_everystep and sv is None, t0_fill=t0, timing=timing, callback=cbs,
                   tstops_np=tst, **kwargs)


def order_by_span(probs, stream=None):
    """Longest-span-first start order of the trajectories (int32 device tensor), computed on the
    device (deg_order_by_span). Pass it as ``order=`` to the solve: long trajectories start first,
    which shortens the tail in which lanes idle; results do not depend on the order."""
    import torch

    dev = probs.u0.device
    order = torch.empty(probs.N, dtype=torch.int32, device=dev)
    work = torch.empty(_lib.ORDER_WORK_INTS, dtype=torch.int32, device=dev)
    st = stream if stream is not None else torch.cuda.current_stream(dev)
    _lib.order_by_span(probs.t0s.data_ptr() if probs.t0s is not None else 0,
                       probs.tfs.data_ptr() if probs.tfs is not None else 0, float(probs.tspan[0]),
                       float(probs.tspan[1]), probs.N, order.data_ptr(), work.data_ptr(), st.cuda_stream)
    order.work = work  # keep the workspace alive with the order until the stream has used it
    return order
