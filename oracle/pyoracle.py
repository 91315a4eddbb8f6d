"""ctypes binding of the oracle library (oracle/_build/libdeg_oracle.so).

TEST INFRASTRUCTURE: used only by tests/, the smoke check and the benchmark's CPU-baseline
leg, as the checker of the HIP solver. The C source is oracle/deg_oracle.c.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "libdeg_oracle.so")

SYSTEM_IDS = {"lorenz": 0, "decay": 1, "growth": 2, "quad": 3, "robertson": 4, "vdp": 5, "tdep": 6, "rot": 7,
              "ball": 8, "rober_dae": 9, "robertson_eye": 10, "dae2": 11, "lorenz_dual": 12}
SYSTEM_DIMS = {"lorenz": (3, 3), "decay": (1, 1), "growth": (1, 1), "quad": (1, 1), "robertson": (3, 3),
               "vdp": (2, 1), "tdep": (1, 1), "rot": (2, 1), "ball": (2, 1), "rober_dae": (3, 3),
               "robertson_eye": (3, 3), "dae2": (2, 0), "lorenz_dual": (21, 21)}
METHOD_IDS = {"tsit5": 0, "vern7": 1, "rb23": 2, "vern9": 3, "rodas4": 4, "rodas5p": 5, "kvaerno3": 6,
              "kvaerno5": 7}
KAHAN, DETPOW, FDJAC, AUGMENTED, PIVLU = 1, 2, 4, 8, 16

_fp = ctypes.POINTER(ctypes.c_float)
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)
_lp = ctypes.POINTER(ctypes.c_int64)
_lib = None

# callback descriptors (deg_oracle.c: ocb)
DISCRETE, CONTINUOUS = 0, 1
COND_T_EQ, COND_U0, COND_U0_MINUS, COND_U0_LT = 0, 1, 2, 3
AFF_ADD, AFF_SCALE, AFF_TERMINATE, AFF_KEEP_FSAL = 1, 2, 4, 8
EPS32 = float(np.finfo(np.float32).eps)


class OCB(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("cond", ctypes.c_int32), ("affect", ctypes.c_int32),
                ("affect_neg", ctypes.c_int32), ("rootfind", ctypes.c_int32), ("cval", ctypes.c_float),
                ("aval", ctypes.c_float * 4), ("abstol", ctypes.c_float), ("nudge", ctypes.c_float),
                ("dtrelax", ctypes.c_float)]


def discrete(cond, cval=0.0, affect=0, aval=()):
    """A discrete callback descriptor: condition id (COND_T_EQ / COND_U0_LT) with its constant, and
    affect bits (AFF_*) with their constants."""
    return dict(kind=DISCRETE, cond=cond, cval=cval, affect=affect, affect_neg=-1, aval=aval)


def continuous(cond, cval=0.0, affect=0, affect_neg="same", aval=(), rootfind=1, abstol=10 * EPS32,
               nudge=0.01, dtrelax=1.0):
    """A continuous callback descriptor (GPUContinuousCallback defaults). affect / affect_neg: AFF_*
    bits or None for `nothing`; affect_neg "same" copies affect."""
    an = affect if affect_neg == "same" else affect_neg
    return dict(kind=CONTINUOUS, cond=cond, cval=cval, affect=-1 if affect is None else affect,
                affect_neg=-1 if an is None else an, aval=aval, rootfind=rootfind, abstol=abstol, nudge=nudge,
                dtrelax=dtrelax)


def _ocb_array(cbs):
    arr = (OCB * max(1, len(cbs)))()
    for q, c in enumerate(cbs):
        o = arr[q]
        o.kind, o.cond, o.affect, o.affect_neg = c["kind"], c["cond"], c["affect"], c.get("affect_neg", -1)
        o.rootfind = c.get("rootfind", 0)
        o.cval = float(np.float32(c.get("cval", 0.0)))
        av = list(c.get("aval", ())) + [0.0] * 4
        for j in range(4):
            o.aval[j] = float(np.float32(av[j]))
        o.abstol = float(np.float32(c.get("abstol", 0.0)))
        o.nudge = float(np.float32(c.get("nudge", 0.01)))
        o.dtrelax = float(np.float32(c.get("dtrelax", 1.0)))
    return arr


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        L.deg_oracle_solve.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int64, _fp, _fp, _fp, _fp,
                                       ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                       _fp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int64,
                                       _fp, _fp, _ip, _ip, _ip, ctypes.c_int]
        L.deg_oracle_solve.restype = ctypes.c_int
        L.deg_oracle_solve_cb.argtypes = L.deg_oracle_solve.argtypes + [_fp, ctypes.c_int, ctypes.c_void_p,
                                                                         ctypes.c_int]
        L.deg_oracle_solve_cb.restype = ctypes.c_int
        L.deg_oracle_vern9.argtypes = [ctypes.c_int, ctypes.c_int64, _fp, _fp, _fp, _fp, ctypes.c_double,
                                       ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp,
                                       ctypes.c_int, _dp, _ip, _lp, ctypes.c_int]
        L.deg_oracle_vern9.restype = ctypes.c_int
        L.deg_oracle_jac.argtypes = [ctypes.c_int, ctypes.c_int, _fp, _fp, ctypes.c_float, _fp, _fp]
        L.deg_oracle_jac.restype = ctypes.c_int
        L.deg_oracle_det_pow.argtypes = [ctypes.c_float, ctypes.c_float]
        L.deg_oracle_det_pow.restype = ctypes.c_float
        L.deg_oracle_dual_eest.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _fp, _fp, _fp, ctypes.c_float,
                                           ctypes.c_float]
        L.deg_oracle_dual_eest.restype = ctypes.c_float
        L.deg_oracle_tableau.argtypes = [ctypes.c_int, _dp, _dp, _dp, _dp]
        L.deg_oracle_tableau.restype = ctypes.c_int
        L.deg_oracle_tsit5_btheta.argtypes = [ctypes.c_float, _fp]
        L.deg_oracle_tsit5_r.argtypes = [ctypes.c_int, ctypes.c_int]
        L.deg_oracle_tsit5_r.restype = ctypes.c_double
        L.deg_oracle_vern9_dense.argtypes = [_dp, _dp, _dp]
        L.deg_oracle_vern9_btheta.argtypes = [ctypes.c_float, _fp]
        _u32p = ctypes.POINTER(ctypes.c_uint32)
        _u64p = ctypes.POINTER(ctypes.c_uint64)
        L.deg_oracle_philox.argtypes = [_u32p, _u32p, _u32p]
        L.deg_oracle_bm_m2log.argtypes = [ctypes.c_uint32]
        L.deg_oracle_bm_m2log.restype = ctypes.c_float
        L.deg_oracle_bm_sincos.argtypes = [ctypes.c_uint32, _fp, _fp]
        L.deg_oracle_normals.argtypes = [ctypes.c_uint64, ctypes.c_int64, ctypes.c_int64, ctypes.c_uint32,
                                         ctypes.c_int, _fp]
        L.deg_oracle_sde_dims.argtypes = [ctypes.c_int, _ip, _ip, _ip]
        L.deg_oracle_sde_dims.restype = ctypes.c_int
        L.deg_oracle_sde_solve.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int64, _fp, _fp, _fp, _fp,
                                           ctypes.c_float, ctypes.c_float, ctypes.c_float, _u64p, ctypes.c_uint64,
                                           ctypes.c_int64, _fp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _fp, _fp,
                                           _ip, _ip, _dp, ctypes.c_int]
        L.deg_oracle_sde_solve.restype = ctypes.c_int
        _lib = L
    return _lib


def _p(a, t):
    return None if a is None else a.ctypes.data_as(t)


def _f32(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float32)


def jac(system: str, u, p, t, fd=False):
    """(J, dT) at one state: the system's analytic Jacobian and time gradient, or with fd the
    forward differences an ``autodiff=False`` implicit method uses."""
    n, np_ = SYSTEM_DIMS[system]
    u = np.ascontiguousarray(u, np.float32).reshape(n)
    p = np.ascontiguousarray(p if np_ else [0.0], np.float32).reshape(-1)
    J = np.zeros(n * n, np.float32)
    dT = np.zeros(n, np.float32)
    r = lib().deg_oracle_jac(SYSTEM_IDS[system], int(bool(fd)), _p(u, _fp), _p(p, _fp), ctypes.c_float(t),
                             _p(J, _fp), _p(dT, _fp))
    if r != 0:
        raise ValueError(f"deg_oracle_jac failed ({r})")
    return J.reshape(n, n), dT


def solve(system: str, method: str, adaptive: bool, u0, p, tspan, dt, *, abstol=1e-6, reltol=1e-3, saveat=None,
          nsave=None, save_everystep=False, t0s=None, tfs=None, kahan=False, detpow=False, max_attempts=0,
          nthreads=None, tstops=None, callbacks=(), fdjac=False, augmented=False, pivlu=False):
    """Float32 restatement over an ensemble. u0 (n, N), p (np, N) float32 SoA. Returns dict with
    ts (nsave, N), us (nsave, n, N), retcode, naccept, nreject. Rows the reference leaves
    uninitialised are NaN here. tstops: stopping times; callbacks: descriptors from discrete() /
    continuous(). augmented: a dual-number system is solved as a plain system of all its entries
    (full augmented Jacobian and W, plain norm) -- a check of the dual W factor, not a reference
    behaviour. fdjac: the implicit methods take J and the time gradient by forward differences
    of f (autodiff = false) instead of the system's analytic ones. pivlu: W is solved by LU with
    partial pivoting also for n <= 3, where the reference (and the default) use Cramer's rule -- a
    check of the Cramer solve, not a reference behaviour."""
    n, np_ = SYSTEM_DIMS[system]
    u0 = _f32(u0)
    p = _f32(p) if np_ > 0 else np.zeros((0, u0.shape[1]), np.float32)
    N = u0.shape[1]
    sv = _f32(saveat)
    if nsave is None:
        nsave = len(sv) if sv is not None else 2
    ts = np.full((nsave, N), np.float32(tspan[0]), np.float32)
    us = np.full((nsave, n, N), np.nan, np.float32)
    rc = np.zeros(N, np.int32)
    na = np.zeros(N, np.int32)
    nr = np.zeros(N, np.int32)
    flags = (KAHAN if kahan else 0) | (DETPOW if detpow else 0) | (FDJAC if fdjac else 0) | (AUGMENTED if augmented else 0)
    flags |= PIVLU if pivlu else 0
    t0s, tfs = _f32(t0s), _f32(tfs)
    tst = None if tstops is None else np.ascontiguousarray(np.asarray(tstops, np.float64).astype(np.float32).reshape(-1))
    cbs = list(callbacks or ())
    arr = _ocb_array(cbs)
    r = lib().deg_oracle_solve_cb(SYSTEM_IDS[system], METHOD_IDS[method], int(adaptive), N, _p(u0, _fp),
                                  _p(p, _fp), _p(t0s, _fp), _p(tfs, _fp), float(np.float32(tspan[0])),
                                  float(np.float32(tspan[1])), float(np.float32(dt)), float(np.float32(abstol)),
                                  float(np.float32(reltol)), _p(sv, _fp), 0 if sv is None else len(sv), int(nsave),
                                  int(bool(save_everystep)), flags, int(max_attempts), _p(ts, _fp), _p(us, _fp),
                                  _p(rc, _ip), _p(na, _ip), _p(nr, _ip), int(nthreads or os.cpu_count() or 1),
                                  _p(tst, _fp), 0 if tst is None else len(tst), ctypes.addressof(arr), len(cbs))
    if r != 0:
        raise RuntimeError(f"oracle solve failed ({r})")
    return dict(ts=ts, us=us, retcode=rc, naccept=na, nreject=nr)


def vern9(system: str, u0, p, tspan, *, abstol=1e-12, reltol=1e-12, saveat=None, t0s=None, tfs=None, dt0=1e-3,
          nthreads=None):
    """FP64 Vern9 reference solution at the saveat times (or the end time). Returns (rows, n, N)."""
    n, np_ = SYSTEM_DIMS[system]
    u0 = _f32(u0)
    p = _f32(p) if np_ > 0 else np.zeros((0, u0.shape[1]), np.float32)
    N = u0.shape[1]
    sv = None if saveat is None else np.ascontiguousarray(saveat, dtype=np.float64)
    rows = 1 if sv is None else len(sv)
    out = np.zeros((rows, n, N), np.float64)
    st = np.zeros(N, np.int32)
    ns = np.zeros(N, np.int64)
    t0s, tfs = _f32(t0s), _f32(tfs)
    r = lib().deg_oracle_vern9(SYSTEM_IDS[system], N, _p(u0, _fp), _p(p, _fp), _p(t0s, _fp), _p(tfs, _fp),
                               float(tspan[0]), float(tspan[1]), abstol, reltol, dt0, _p(sv, _dp),
                               0 if sv is None else len(sv), _p(out, _dp), _p(st, _ip), _p(ns, _lp),
                               int(nthreads or os.cpu_count() or 1))
    if r != 0:
        raise RuntimeError(f"oracle vern9 failed ({r})")
    return out, st, ns


def det_pow(x: float, y: float) -> float:
    return float(lib().deg_oracle_det_pow(float(x), float(y)))


def dual_eest(nv: int, P: int, n: int, e, uprev, u, abstol: float, reltol: float) -> float:
    """The oracle's dual error norm (EEst) of one state."""
    arrs = [np.ascontiguousarray(x, dtype=np.float32) for x in (e, uprev, u)]
    return float(lib().deg_oracle_dual_eest(nv, P, n, *(_p(a, _fp) for a in arrs), float(abstol), float(reltol)))


def tableau(which: str):
    idx = {"tsit5": 0, "vern7": 1, "vern9": 2}[which]
    c = np.zeros(18)
    A = np.zeros(18 * 18)
    b = np.zeros(18)
    bt = np.zeros(18)
    s = lib().deg_oracle_tableau(idx, _p(c, _dp), _p(A, _dp), _p(b, _dp), _p(bt, _dp))
    A = A.reshape(18, 18)
    # 1-based stage arrays -> 0-based
    return c[1:s + 1], A[1:s + 1, 1:s + 1], b[1:s + 1], bt[1:s + 1]


def tsit5_btheta(theta: float) -> np.ndarray:
    out = np.zeros(7, np.float32)
    lib().deg_oracle_tsit5_btheta(float(np.float32(theta)), _p(out, _fp))
    return out


def tsit5_r():
    R = np.zeros((8, 5))
    for i in range(8):
        for k in range(5):
            R[i, k] = lib().deg_oracle_tsit5_r(i, k)
    return R


def vern9_dense():
    """Vern9 dense output tables: nodes C (27), coefficients A (27x27, rows 17..26), polynomial
    coefficients R (27x10), stage indices 1-based."""
    C = np.zeros(27)
    A = np.zeros(27 * 27)
    R = np.zeros(27 * 10)
    lib().deg_oracle_vern9_dense(_p(C, _dp), _p(A, _dp), _p(R, _dp))
    return C, A.reshape(27, 27), R.reshape(27, 10)


def vern9_btheta(theta: float) -> np.ndarray:
    """The Float32 weights b_i(theta) (index 1..26, 0 unused) the oracle's Vern9 dense output uses."""
    out = np.zeros(27, np.float32)
    lib().deg_oracle_vern9_btheta(float(np.float32(theta)), _p(out, _fp))
    return out


# ---- SDE kernels (deg_oracle_sde.c) ----
SDE_SYSTEM_IDS = {"gbm": 0, "gbm1": 1, "lorenz": 2, "nondiag": 3, "ou": 4, "bm": 5}
SDE_NP = {"gbm": 2, "gbm1": 0, "lorenz": 3, "nondiag": 0, "ou": 3, "bm": 1}
SDE_METHOD_IDS = {"em": 8, "siea": 9}


def philox(ctr, key) -> np.ndarray:
    """Philox4x32-10 of a 4-word counter and a 2-word key (uint32 words)."""
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    out = np.zeros(4, np.uint32)
    t = ctypes.POINTER(ctypes.c_uint32)
    lib().deg_oracle_philox(c.ctypes.data_as(t), k.ctypes.data_as(t), out.ctypes.data_as(t))
    return out


def bm_m2log(w: int) -> float:
    """-2 ln U of the Box-Muller transform for the 32-bit word w (U = ((w >> 8) + 1) 2^-24)."""
    return float(lib().deg_oracle_bm_m2log(int(w)))


def bm_sincos(w: int):
    """(cos, sin) of 2 pi (w >> 8) 2^-24 as the Box-Muller transform evaluates them."""
    c = ctypes.c_float()
    s = ctypes.c_float()
    lib().deg_oracle_bm_sincos(int(w), ctypes.byref(c), ctypes.byref(s))
    return float(c.value), float(s.value)


def normals(seed: int, id0: int, n: int, step: int, m: int) -> np.ndarray:
    """(m, n) standard normals of step `step` for streams id0 .. id0 + n - 1."""
    out = np.zeros((m, n), np.float32)
    lib().deg_oracle_normals(int(seed) & 0xFFFFFFFFFFFFFFFF, int(id0), int(n), int(step), int(m), _p(out, _fp))
    return out


def sde_dims(system: str):
    n, m, d = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
    lib().deg_oracle_sde_dims(SDE_SYSTEM_IDS[system], ctypes.byref(n), ctypes.byref(m), ctypes.byref(d))
    return n.value, m.value, bool(d.value)


def sde_solve(system: str, method: str, u0, p, tspan, dt, *, seed=0, seeds=None, traj_base=0, saveat=None,
              nsave=None, save_everystep=True, t0s=None, tfs=None, want_wsum=False, nthreads=None):
    """Float32 restatement of em_kernel / siea_kernel over an ensemble. u0 (n, N), p (np, N) float32
    SoA. Returns dict with ts (nsave, N), us (nsave, n, N), retcode, naccept and (want_wsum) wsum
    (m, N): the sum of the Wiener increments of each noise component (float64)."""
    n, m, _diag = sde_dims(system)
    u0 = _f32(u0)
    N = u0.shape[1]
    np_ = SDE_NP[system]
    p = _f32(p) if np_ > 0 else np.zeros((1, N), np.float32)
    sv = _f32(saveat)
    if nsave is None:
        if sv is not None:
            nsave = len(sv)
        elif save_everystep:
            raise ValueError("nsave is required for every-step output")
        else:
            nsave = 2
    ts = np.full((nsave, N), np.float32(tspan[0]), np.float32)
    us = np.full((nsave, n, N), np.nan, np.float32)
    rc = np.zeros(N, np.int32)
    na = np.zeros(N, np.int32)
    ws = np.zeros((m, N), np.float64) if want_wsum else None
    sd = None if seeds is None else np.ascontiguousarray(np.asarray(seeds, dtype=np.uint64))
    t0s, tfs = _f32(t0s), _f32(tfs)
    r = lib().deg_oracle_sde_solve(SDE_SYSTEM_IDS[system], SDE_METHOD_IDS[method], N, _p(u0, _fp), _p(p, _fp),
                                   _p(t0s, _fp), _p(tfs, _fp), float(np.float32(tspan[0])),
                                   float(np.float32(tspan[1])), float(np.float32(dt)),
                                   None if sd is None else sd.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                                   int(seed) & 0xFFFFFFFFFFFFFFFF, int(traj_base), _p(sv, _fp),
                                   0 if sv is None else len(sv), int(nsave), int(bool(save_everystep)),
                                   _p(ts, _fp), _p(us, _fp), _p(rc, _ip), _p(na, _ip), _p(ws, _dp),
                                   int(nthreads or os.cpu_count() or 1))
    if r != 0:
        raise RuntimeError(f"oracle SDE solve failed ({r})")
    out = dict(ts=ts, us=us, retcode=rc, naccept=na)
    if want_wsum:
        out["wsum"] = ws
    return out
