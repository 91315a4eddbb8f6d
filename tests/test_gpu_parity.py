"""GPU parity: the HIP solver (through libdeghip.so) against the CPU oracle on the same seeded
inputs.

Bars (each stated where it is asserted):
  * fixed step, no contraction: bit-identical ts and us, identical step counts;
  * adaptive with the deterministic controller pow: bit-identical ts, us, return codes and
    accept/reject counts, persistent and non-persistent grids alike;
  * adaptive with the device powf (the library call AMDGPU.jl makes for Float32 ^): the
    FP64-oracle error within 2x of the oracle's own Float32 error, same return codes;
  * the default fast path (FMA contraction, persistent): FP64-oracle endpoint error no worse
    than the oracle's own Float32 error beyond a small factor;
  * the reference's own KATs (public_interface.jl, conversions.jl) through the high-level API.
"""
import numpy as np
import pytest

import pyoracle

pytestmark = pytest.mark.gpu
F32 = np.float32
METHODS = {"tsit5": "GPUTsit5", "vern7": "GPUVern7", "rb23": "GPURosenbrock23", "vern9": "GPUVern9",
           "rodas4": "GPURodas4", "rodas5p": "GPURodas5P", "kvaerno3": "GPUKvaerno3", "kvaerno5": "GPUKvaerno5"}


@pytest.fixture(scope="module")
def torch_cuda():
    import torch

    if not torch.cuda.is_available():
        pytest.fail("no HIP device visible: the GPU tests need an MI355X")
    return torch


def inputs(system, N, seed):
    """Seeded SoA inputs for a system."""
    from diffeqgpu_amd import rng, workloads

    U = rng.uniform_table(seed, 4, N)
    if system == "lorenz":
        return workloads.lorenz_random(N, seed)
    if system == "vdp":
        mu = (0.1 + 4.9 * U[0]).astype(F32)[None, :]
        u0 = (-2.0 + 4.0 * U[1:3]).astype(F32)
        return u0, mu
    if system == "robertson":
        p = (np.array([0.04, 3e7, 1e4], np.float64)[:, None] * (0.5 + U[:3])).astype(F32)
        u0 = np.zeros((3, N), F32)
        u0[0] = 1.0
        return u0, p
    if system in ("decay", "quad", "growth", "tdep"):
        return (1.0 + U[0:1]).astype(F32), (0.5 + U[1:2]).astype(F32)
    if system == "rot":
        return (U[0:2] - 0.5).astype(F32), (0.5 + U[2:3]).astype(F32)
    raise KeyError(system)


def gpu_solve(deg, torch, system, method, adaptive, u0, p, tspan, dt, dev, t0s=None, tfs=None, **kw):
    from diffeqgpu_amd import systems

    f, n, np_, _ = systems.SYSTEMS[system]
    probs = deg.HIPProbVector.from_arrays(f, u0, p, tspan, dev, t0s=t0s, tfs=tfs)
    prob = deg.ODEProblem(f, u0[:, 0], tspan, p[:, 0])
    alg = getattr(deg, METHODS[method])()
    fn = deg.vectorized_asolve if adaptive else deg.vectorized_solve
    print(f"[gpu] {system} {method} adaptive={adaptive} N={u0.shape[1]} flags={dev} kw={sorted(kw)}", flush=True)
    res = fn(probs, prob, alg, dt=dt, **kw)
    torch.cuda.synchronize()
    return dict(ts=res.ts.cpu().numpy(), us=res.us_soa.cpu().numpy(), retcode=res.retcode.cpu().numpy(),
                naccept=res.naccept.cpu().numpy(), nreject=res.nreject.cpu().numpy(), res=res)


def assert_bitwise(g, o):
    np.testing.assert_array_equal(g["ts"], o["ts"])
    written = ~np.isnan(o["us"])
    assert np.array_equal(g["us"][written], o["us"][written]), \
        f"{np.sum(g['us'][written] != o['us'][written])} of {written.sum()} state entries differ"
    np.testing.assert_array_equal(g["retcode"], o["retcode"])
    np.testing.assert_array_equal(g["naccept"], o["naccept"])
    np.testing.assert_array_equal(g["nreject"], o["nreject"])


# ---------------------------------------------------------------------------------------
# fixed step, bit for bit
# ---------------------------------------------------------------------------------------
FIXED_CASES = [
    ("lorenz", "tsit5", (0.0, 10.0), 0.1), ("lorenz", "vern7", (0.0, 10.0), 0.1), ("lorenz", "rb23", (0.0, 2.0), 0.01),
    ("vdp", "tsit5", (0.0, 5.0), 0.05), ("vdp", "vern7", (0.0, 5.0), 0.05), ("vdp", "rb23", (0.0, 5.0), 0.01),
    ("tdep", "tsit5", (0.0, 2.0), 0.015625), ("rot", "vern7", (0.0, 8.0), 0.5), ("robertson", "rb23", (0.0, 1.0), 0.001),
    ("lorenz", "vern9", (0.0, 10.0), 0.1), ("vdp", "vern9", (0.0, 5.0), 0.05), ("tdep", "vern9", (0.0, 2.0), 0.0625),
    ("lorenz", "rodas4", (0.0, 2.0), 0.01), ("vdp", "rodas4", (0.0, 5.0), 0.01), ("tdep", "rodas4", (0.0, 2.0), 0.0625),
    ("robertson", "rodas5p", (0.0, 1.0), 0.001), ("vdp", "rodas5p", (0.0, 5.0), 0.01), ("tdep", "rodas5p", (0.0, 2.0), 0.0625),
    ("vdp", "kvaerno3", (0.0, 5.0), 0.01), ("tdep", "kvaerno3", (0.0, 2.0), 0.0625), ("lorenz", "kvaerno3", (0.0, 1.0), 0.01),
    ("robertson", "kvaerno5", (0.0, 1.0), 0.001), ("vdp", "kvaerno5", (0.0, 5.0), 0.01), ("tdep", "kvaerno5", (0.0, 2.0), 0.0625),
]


@pytest.mark.parametrize("system,method,tspan,dt", FIXED_CASES)
@pytest.mark.parametrize("mode", ["endpoint", "everystep", "saveat"])
def test_fixed_step_bitwise(deg, torch_cuda, system, method, tspan, dt, mode):
    N = 300
    u0, p = inputs(system, N, seed=11)
    dev = deg.HIPDevice(contract="off")
    kw = {}
    okw = {}
    if mode == "endpoint":
        kw["save_everystep"] = False
    elif mode == "everystep":
        kw["save_everystep"] = True
        nsave = len(deg.colon(tspan[0], dt, tspan[1]))
        okw = dict(save_everystep=True, nsave=nsave)
    else:
        sv = np.linspace(tspan[0], tspan[1], 7).astype(F32)
        sv[3] = F32(tspan[1] * 0.37)
        sv.sort()
        kw["saveat"] = sv
        okw = dict(saveat=sv)
    g = gpu_solve(deg, torch_cuda, system, method, False, u0, p, tspan, dt, dev, **kw)
    o = pyoracle.solve(system, method, False, u0, p, tspan, dt, **okw)
    assert_bitwise(g, o)


# ---------------------------------------------------------------------------------------
# adaptive, deterministic controller pow: bit for bit
# ---------------------------------------------------------------------------------------
ADAPTIVE_CASES = [
    ("lorenz", "tsit5", (0.0, 10.0), 0.1, 1e-6, 1e-6), ("lorenz", "vern7", (0.0, 10.0), 0.1, 1e-8, 1e-8),
    ("lorenz", "rb23", (0.0, 3.0), 0.01, 1e-5, 1e-5), ("vdp", "tsit5", (0.0, 10.0), 0.01, 1e-6, 1e-6),
    ("vdp", "vern7", (0.0, 10.0), 0.01, 1e-7, 1e-7), ("vdp", "rb23", (0.0, 10.0), 0.01, 1e-6, 1e-4),
    ("robertson", "rb23", (0.0, 1e5), 1e-6, 1e-8, 1e-4), ("quad", "rb23", (0.0, 10.0), 0.01, 1e-6, 1e-3),
    ("tdep", "tsit5", (0.0, 3.0), 0.1, 1e-7, 1e-7), ("lorenz", "vern9", (0.0, 10.0), 0.1, 1e-8, 1e-8),
    ("vdp", "vern9", (0.0, 10.0), 0.01, 1e-7, 1e-7), ("robertson", "rodas4", (0.0, 1e5), 1e-6, 1e-8, 1e-4),
    ("vdp", "rodas4", (0.0, 10.0), 0.01, 1e-6, 1e-4), ("robertson", "rodas5p", (0.0, 1e5), 1e-6, 1e-8, 1e-4),
    ("lorenz", "rodas5p", (0.0, 3.0), 0.01, 1e-5, 1e-5), ("robertson", "kvaerno3", (0.0, 1e5), 1e-6, 1e-8, 1e-4),
    ("vdp", "kvaerno3", (0.0, 10.0), 0.01, 1e-6, 1e-4), ("robertson", "kvaerno5", (0.0, 1e5), 1e-6, 1e-8, 1e-4),
    ("vdp", "kvaerno5", (0.0, 10.0), 0.01, 1e-6, 1e-5),
]


@pytest.mark.parametrize("system,method,tspan,dt,abstol,reltol", ADAPTIVE_CASES)
@pytest.mark.parametrize("persistent", [False, True])
@pytest.mark.parametrize("mode", ["endpoint", "saveat"])
def test_adaptive_detpow_bitwise(deg, torch_cuda, system, method, tspan, dt, abstol, reltol, persistent, mode):
    N = 777
    u0, p = inputs(system, N, seed=23)
    dev = deg.HIPDevice(contract="off", powmode="det", persistent=persistent)
    kw = dict(abstol=abstol, reltol=reltol)
    okw = dict(abstol=abstol, reltol=reltol, detpow=True)
    if mode == "saveat":
        if system == "robertson":
            sv = np.array([10.0 ** (-4 + 9 * k / 19) for k in range(20)], F32)
        else:
            sv = np.linspace(tspan[0], tspan[1], 11).astype(F32)
        kw["saveat"] = sv
        okw["saveat"] = sv
    g = gpu_solve(deg, torch_cuda, system, method, True, u0, p, tspan, dt, dev, **kw)
    o = pyoracle.solve(system, method, True, u0, p, tspan, dt, **okw)
    assert_bitwise(g, o)
    assert np.all(o["retcode"] == 0)


def test_adaptive_uneven_tspans_kahan_persistent(deg, torch_cuda):
    """Per-trajectory end times (the uneven-batch workload shape) with compensated t."""
    from diffeqgpu_amd import workloads

    cfg = workloads.config("lorenz_uneven", N=2000, seed=5)
    for kahan in (False, True):
        dev = deg.HIPDevice(contract="off", powmode="det", persistent=True, kahan=kahan)
        g = gpu_solve(deg, torch_cuda, "lorenz", "vern7", True, cfg["u0"], cfg["p"], (0.0, 50.0), 0.01, dev,
                      tfs=cfg["tfs"], abstol=1e-6, reltol=1e-6)
        o = pyoracle.solve("lorenz", "vern7", True, cfg["u0"], cfg["p"], (0.0, 50.0), 0.01, tfs=cfg["tfs"],
                           abstol=1e-6, reltol=1e-6, detpow=True, kahan=kahan)
        assert_bitwise(g, o)
        np.testing.assert_array_equal(g["ts"][1], cfg["tfs"])


def test_persistent_equals_plain_grid(deg, torch_cuda):
    """The work queue only changes which lane runs a trajectory, never its result."""
    u0, p = inputs("lorenz", 5000, seed=3)
    outs = []
    for persistent in (False, True):
        dev = deg.HIPDevice(persistent=persistent)
        g = gpu_solve(deg, torch_cuda, "lorenz", "tsit5", True, u0, p, (0.0, 10.0), 0.1, dev, abstol=1e-6,
                      reltol=1e-6, saveat=1.0)
        outs.append(g)
    for k in ("ts", "us", "naccept", "nreject", "retcode"):
        np.testing.assert_array_equal(outs[0][k], outs[1][k])


# ---------------------------------------------------------------------------------------
# adaptive with the device powf
# ---------------------------------------------------------------------------------------
@pytest.mark.parametrize("system,method,tspan,tol", [("vdp", "tsit5", (0.0, 10.0), 1e-6), ("decay", "vern7", (0.0, 5.0), 1e-8),
                                                     ("rot", "tsit5", (0.0, 20.0), 1e-7), ("quad", "rb23", (0.0, 10.0), 1e-5),
                                                     ("lorenz", "vern7", (0.0, 1.0), 1e-8), ("decay", "vern9", (0.0, 5.0), 1e-8)])
def test_adaptive_powf_fp64_error(deg, torch_cuda, system, method, tspan, tol):
    """Exact operation order with the device powf instead of the oracle's powf: the step
    sequences may drift apart by rounding of the controller, so the bar is the FP64-oracle
    error: the GPU error is within 2x of the oracle's own Float32 error at the median and the
    90th percentile (plus 1e-7 absolute), with identical return codes and end times."""
    u0, p = inputs(system, 1000, seed=31)
    dev = deg.HIPDevice(contract="off", powmode="ocml", persistent=True)
    g = gpu_solve(deg, torch_cuda, system, method, True, u0, p, tspan, 0.01, dev, abstol=tol, reltol=tol)
    o = pyoracle.solve(system, method, True, u0, p, tspan, 0.01, abstol=tol, reltol=tol)
    ref, st, _ = pyoracle.vern9(system, u0, p, tspan)
    assert np.all(st == 0)
    np.testing.assert_array_equal(g["retcode"], o["retcode"])
    np.testing.assert_array_equal(g["ts"], o["ts"])
    eg = np.abs(g["us"][1] - ref[0]).max(axis=0)
    eo = np.abs(o["us"][1] - ref[0]).max(axis=0)
    for q in (50, 90):
        assert np.percentile(eg, q) <= 2 * np.percentile(eo, q) + 1e-7, (q, np.percentile(eg, q), np.percentile(eo, q))
    # step counts stay close on average (near Float32 resolution the controller is noisy)
    assert abs(g["naccept"].mean() - o["naccept"].mean()) <= 0.1 * o["naccept"].mean()


# ---------------------------------------------------------------------------------------
# default fast path against the FP64 oracle
# ---------------------------------------------------------------------------------------
def test_fast_path_fp64_error_lorenz_adaptive(deg, torch_cuda):
    """The benchmark's adaptive Lorenz workload (saveat 0:1:10, tol 1e-6) on 10,000 trajectories:
    the error against FP64 Vern9 at 1e-12, the largest over the 11 save rows and 3 states of each
    trajectory (so the dense output at interior save points counts), is within 1.25x of the oracle's
    own Float32 error on the same trajectories at the median and the 90th percentile. The default path
    differs from the oracle only by FMA contraction and the log2/exp2 controller, whose rounding moves
    step sequences, not error levels; a wrong dense-output coefficient or a lost order would miss by
    far more. Every trajectory succeeds and the saved times are the grid."""
    from diffeqgpu_amd import workloads

    cfg = workloads.config("lorenz_adaptive", N=10_000, seed=202)
    sv = np.arange(11, dtype=F32)
    ref, st, _ = pyoracle.vern9("lorenz", cfg["u0"], cfg["p"], (0.0, 10.0), saveat=sv.astype(np.float64))
    assert np.all(st == 0)
    g = gpu_solve(deg, torch_cuda, "lorenz", "tsit5", True, cfg["u0"], cfg["p"], (0.0, 10.0), 0.1, deg.HIPDevice(),
                  abstol=1e-6, reltol=1e-6, saveat=1.0)
    o = pyoracle.solve("lorenz", "tsit5", True, cfg["u0"], cfg["p"], (0.0, 10.0), 0.1, abstol=1e-6, reltol=1e-6,
                       saveat=sv)
    assert np.all(g["retcode"] == 0)
    np.testing.assert_array_equal(g["ts"], o["ts"])
    eg = np.abs(g["us"] - ref).max(axis=(0, 1))
    eo = np.abs(o["us"] - ref).max(axis=(0, 1))
    for q in (50, 90):
        assert np.percentile(eg, q) <= 1.25 * np.percentile(eo, q), (q, np.percentile(eg, q), np.percentile(eo, q))


@pytest.mark.parametrize("method", ["rb23", "rodas4", "rodas5p", "kvaerno3", "kvaerno5"])
def test_fast_path_implicit_robertson_fixture(deg, torch_cuda, method):
    """The implicit methods on the default path (contraction, hardware controller pow, the W factor
    keeping pivot reciprocals instead of dividing) on the stiff workload's 1,000-trajectory FP64 Radau
    fixture: every trajectory succeeds, y1 + y2 + y3 = 1 to 1e-4, and the error against the fixture
    is within 2x of the oracle's own Float32 error on the same trajectories at the median and 4x at
    the 99th percentile (the Kvaerno Newton solves stop at 100 eps(Float32), where contraction moves
    the iteration counts), and below 1e-3 everywhere. A wrong factorisation errs by 1e-2 or more."""
    import os

    from diffeqgpu_amd import workloads

    cfg = workloads.config("robertson", N=1000)
    gold = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "robertson_radau.npz"))
    sv = np.asarray(cfg["saveat"], F32)
    g = gpu_solve(deg, torch_cuda, "robertson", method, True, cfg["u0"], cfg["p"], cfg["tspan"], cfg["dt"],
                  deg.HIPDevice(), abstol=1e-8, reltol=1e-4, saveat=sv)
    o = pyoracle.solve("robertson", method, True, cfg["u0"], cfg["p"], cfg["tspan"], cfg["dt"], abstol=1e-8,
                       reltol=1e-4, saveat=sv)
    assert np.all(g["retcode"] == 0)
    us = g["us"].astype(np.float64)
    assert np.abs(us.sum(axis=1) - 1.0).max() < 1e-4
    eg = np.abs(us - gold["u"]).max(axis=(0, 1))
    eo = np.abs(o["us"].astype(np.float64) - gold["u"]).max(axis=(0, 1))
    assert np.median(eg) <= 2 * np.median(eo) + 1e-7, (np.median(eg), np.median(eo))
    assert np.percentile(eg, 99) <= 4 * np.percentile(eo, 99) + 1e-6, (np.percentile(eg, 99), np.percentile(eo, 99))
    assert eg.max() < 1e-3 if method != "kvaerno5" else np.sum(eg >= 1e-3) <= 5, np.sort(eg)[-6:]


# ---------------------------------------------------------------------------------------
# the fixed-step Lorenz workload at full size: bitwise and within 1e-5 relative
# ---------------------------------------------------------------------------------------
def test_lorenz_fixed_full_size(deg, torch_cuda):
    from diffeqgpu_amd import workloads

    cfg = workloads.config("lorenz_fixed")
    g = gpu_solve(deg, torch_cuda, "lorenz", "tsit5", False, cfg["u0"], cfg["p"], (0.0, 10.0), 0.1,
                  deg.HIPDevice(contract="off"), save_everystep=False)
    o = pyoracle.solve("lorenz", "tsit5", False, cfg["u0"], cfg["p"], (0.0, 10.0), 0.1)
    assert_bitwise(g, o)
    rel = np.abs(g["us"][1] - o["us"][1]) / np.maximum(np.abs(o["us"][1]), 1e-30)
    assert rel.max() <= 1e-5
    assert np.all(g["ts"][1] == F32(10.000002)) and np.all(g["naccept"] == 100)


# ---------------------------------------------------------------------------------------
# edge cases
# ---------------------------------------------------------------------------------------
def test_edge_cases(deg, torch_cuda):
    dev = deg.HIPDevice(contract="off", powmode="det", persistent=True)
    # one trajectory, and a count that is not a multiple of the block size
    for N in (1, 65, 257):
        u0, p = inputs("lorenz", N, seed=N)
        g = gpu_solve(deg, torch_cuda, "lorenz", "tsit5", True, u0, p, (0.0, 1.0), 0.1, dev, abstol=1e-6, reltol=1e-6)
        o = pyoracle.solve("lorenz", "tsit5", True, u0, p, (0.0, 1.0), 0.1, abstol=1e-6, reltol=1e-6, detpow=True)
        assert_bitwise(g, o)
    # empty span: no step, the endpoint row holds the initial state
    u0, p = inputs("lorenz", 64, seed=1)
    g = gpu_solve(deg, torch_cuda, "lorenz", "tsit5", True, u0, p, (0.0, 0.0), 0.1, dev)
    np.testing.assert_array_equal(g["us"][1], u0)
    assert np.all(g["naccept"] == 0) and np.all(g["ts"] == 0)
    # tolerance below Float32 resolution: lanes end with DtLessThanMin, or oscillate between
    # EEst == 0 accepts and rejects (the reference loops there) until the attempt cap
    capd = deg.HIPDevice(contract="off", powmode="det", persistent=True, max_attempts=20000)
    g = gpu_solve(deg, torch_cuda, "decay", "tsit5", True, *inputs("decay", 64, 2), (0.0, 1.0), 0.1, capd,
                  abstol=1e-30, reltol=1e-30)
    o = pyoracle.solve("decay", "tsit5", True, *inputs("decay", 64, 2), (0.0, 1.0), 0.1, abstol=1e-30,
                       reltol=1e-30, detpow=True, max_attempts=20000)
    assert set(np.unique(g["retcode"])) <= {1, 2} and np.sum(g["retcode"] == 1) > 32
    assert_bitwise(g, o)
    # attempt cap
    capped = deg.HIPDevice(contract="off", powmode="det", persistent=True, max_attempts=7)
    g = gpu_solve(deg, torch_cuda, "lorenz", "tsit5", True, u0, p, (0.0, 10.0), 0.1, capped, abstol=1e-6, reltol=1e-6)
    assert np.all(g["retcode"] == 2) and np.all(g["naccept"] + g["nreject"] == 7)
    # non-finite parameters end as Unstable
    pn = p.copy()
    pn[1, :5] = np.nan
    g = gpu_solve(deg, torch_cuda, "lorenz", "tsit5", True, u0, pn, (0.0, 1.0), 0.1, dev)
    assert np.all(g["retcode"][:5] == 3) and np.all(g["retcode"][5:] == 0)
    # saveat points outside the span: before t0 extrapolates like the reference, after tf stays t0
    sv = np.array([-0.5, 0.25, 0.5, 2.0], F32)
    g = gpu_solve(deg, torch_cuda, "lorenz", "tsit5", True, u0, p, (0.0, 1.0), 0.1, dev, saveat=sv)
    o = pyoracle.solve("lorenz", "tsit5", True, u0, p, (0.0, 1.0), 0.1, saveat=sv, detpow=True)
    assert_bitwise(g, o)
    assert np.all(g["ts"][3] == 0.0)


# ---------------------------------------------------------------------------------------
# the reference's KATs through the high-level API
# ---------------------------------------------------------------------------------------
def test_public_interface_kats(deg, torch_cuda):
    """test/public_interface.jl:9-41, :71-80."""
    from diffeqgpu_amd import systems

    ode_prob = deg.ODEProblem(systems.growth, [1.0], (0.0, 0.2), [1.0])
    probs = deg.adapt(deg.HIPDevice(), [deg.make_prob_compatible(ode_prob) for _ in range(2)])
    ts, us = deg.vectorized_solve(probs, ode_prob, deg.GPUTsit5(), dt=0.1, save_everystep=False)
    assert tuple(ts.shape) == (2, 2) and tuple(us.shape[:2]) == (2, 2)
    assert np.all(us[0].cpu().numpy() == 1.0)
    ats, aus = deg.vectorized_asolve(probs, ode_prob, deg.GPUTsit5(), dt=0.1, saveat=0.1, save_everystep=False)
    assert ats.shape[1] == 2 and aus.shape[1] == 2
    assert np.all(aus[0, 0].cpu().numpy() == 1.0)
    sol = deg.solve(deg.EnsembleProblem(ode_prob), deg.GPUTsit5(), deg.EnsembleGPUKernel(deg.HIPDevice()),
                    trajectories=2, adaptive=False, dt=0.1)
    assert len(sol.u) == 2
    assert all(s.u[0][0] == 1.0 for s in sol.u)


def test_conversions_kats(deg, torch_cuda):
    """test/gpu_kernel_de/conversions.jl:23-59: the saved time grids, 10,000 trajectories."""
    from diffeqgpu_amd import systems
    from diffeqgpu_amd.ranges import colon

    base = np.array([10.0, 28.0, F32(8.0) / F32(3.0)], F32)
    prob = deg.ODEProblem(systems.SYSTEMS["lorenz"][0], [1.0, 0.0, 0.0], (0.0, 10.0), base)

    def prob_func(pr, ctx):
        return deg.remake(pr, p=ctx.rng.random(3).astype(F32) * base)

    monte = deg.EnsembleProblem(prob, prob_func=prob_func, safetycopy=False)
    alg, ens = deg.GPUTsit5(), deg.EnsembleGPUKernel(deg.HIPDevice())

    def t_of(saveat):
        return deg.solve(monte, alg, ens, trajectories=10_000, seed=1, saveat=saveat).u[0].t

    np.testing.assert_array_equal(t_of(range(1, 11)), np.arange(1, 11, dtype=F32))
    np.testing.assert_array_equal(t_of(colon(1.0, 0.1, 10.0, np.float64)), colon(1.0, 0.1, 10.0).collect())
    np.testing.assert_array_equal(t_of(1.0), np.arange(11, dtype=F32))
    np.testing.assert_array_equal(t_of([1.0, 5.0, 10.0]), [1.0, 5.0, 10.0])


def test_lorenz_regression_gpu(deg, torch_cuda):
    """gpu_ode_regression.jl:19-46 on the GPU for Tsit5 and Vern7: fixed dt = 0.01 endpoint within
    5e-3 and adaptive 1e-7 within 1e-3 of the FP64 Vern9 solution."""
    from diffeqgpu_amd import systems

    u0 = np.array([[1.0], [0.0], [0.0]], F32)
    p = np.array([[10.0], [28.0], [F32(8.0) / F32(3.0)]], F32)
    ref, _, _ = pyoracle.vern9("lorenz", u0, p, (0.0, 10.0))
    prob = deg.ODEProblem(systems.SYSTEMS["lorenz"][0], u0[:, 0], (0.0, 10.0), p[:, 0])
    monte = deg.EnsembleProblem(prob, safetycopy=False)
    for alg in (deg.GPUTsit5(), deg.GPUVern7()):
        sol = deg.solve(monte, alg, deg.EnsembleGPUKernel(deg.HIPDevice()), trajectories=10, adaptive=False, dt=0.01)
        assert np.linalg.norm(sol.u[0].u[-1] - ref[0, :, 0]) < 5e-3
        asol = deg.solve(monte, alg, deg.EnsembleGPUKernel(deg.HIPDevice()), trajectories=10, adaptive=True, dt=0.01,
                         abstol=1e-7, reltol=1e-7)
        assert np.linalg.norm(asol.u[0].u[-1] - ref[0, :, 0]) < 1e-3
        assert asol.u[0].retcode == "Success"
