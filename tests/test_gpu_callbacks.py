"""GPU: tstops and callbacks (discrete, continuous with ITP root finding, terminate!) through the
C ABI, against the CPU oracle's restatement of integrators/integrator_utils.jl:52-479.

Bars:
  * with contraction off and the deterministic controller pow, every output (ts, us, retcode,
    naccept, nreject) equals the oracle bit for bit, over ensembles whose event times differ per
    trajectory, for GPUTsit5 / GPUVern7 / GPURosenbrock23, fixed and adaptive, every-step,
    saveat and endpoint output, persistent and one-shot grids;
  * the reference's own callback tests (test/gpu_kernel_de/gpu_ode_discrete_callbacks.jl,
    gpu_ode_continuous_callbacks.jl) through solve(EnsembleProblem, ...; callback, tstops,
    merge_callbacks = true), at their tolerances against the closed forms those problems have.
"""
import math

import numpy as np
import pytest

import pyoracle as P

pytestmark = pytest.mark.gpu
F32 = np.float32


@pytest.fixture(scope="module")
def torch_cuda():
    import torch

    if not torch.cuda.is_available():
        pytest.fail("no HIP device visible: the GPU tests need an MI355X")
    return torch


def _cases(deg):
    """Traced callbacks for the GPU and the oracle's descriptors of the same functions."""
    S = deg.SVector

    def add10(i):
        i.u += S(10.0)

    def bounce(i):
        i.u = i.u + S(0.0, -2.0) * i.u

    def stop(i):
        deg.terminate(i)

    def add10_keep(i):
        i.u += S(10.0)
        deg.u_modified(i, False)

    def lift(i):
        i.u += S(1000.0, 0.0)

    G = dict(
        jump24=(deg.GPUDiscreteCallback(lambda u, t, i: t == 2.4, add10), P.discrete(P.COND_T_EQ, 2.4, P.AFF_ADD, (10.0,))),
        jump40=(deg.GPUDiscreteCallback(lambda u, t, i: t == 4.0, add10), P.discrete(P.COND_T_EQ, 4.0, P.AFF_ADD, (10.0,))),
        stop24=(deg.GPUDiscreteCallback(lambda u, t, i: t == 2.4, stop), P.discrete(P.COND_T_EQ, 2.4, P.AFF_TERMINATE)),
        keep24=(deg.GPUDiscreteCallback(lambda u, t, i: t == 2.4, add10_keep),
                P.discrete(P.COND_T_EQ, 2.4, P.AFF_ADD | P.AFF_KEEP_FSAL, (10.0,))),
        low=(deg.GPUDiscreteCallback(lambda u, t, i: u[0] < 3.0, add10), P.discrete(P.COND_U0_LT, 3.0, P.AFF_ADD, (10.0,))),
        bounce=(deg.GPUContinuousCallback(lambda u, t, i: u[0], bounce),
                P.continuous(P.COND_U0, affect=P.AFF_SCALE, aval=(0.0, -2.0))),
        bounce_sciml=(deg.ContinuousCallback(lambda u, t, i: u[0], bounce, save_positions=(False, False)),
                      P.continuous(P.COND_U0, affect=P.AFF_SCALE, aval=(0.0, -2.0), abstol=100 * P.EPS32)),
        cross5_right=(deg.GPUContinuousCallback(lambda u, t, i: u[0] - 5.0, add10, rootfind=deg.RightRootFind),
                      P.continuous(P.COND_U0_MINUS, cval=5.0, affect=P.AFF_ADD, aval=(10.0,), rootfind=2)),
        cross5_none=(deg.GPUContinuousCallback(lambda u, t, i: u[0] - 5.0, add10, rootfind=deg.NoRootFind),
                     P.continuous(P.COND_U0_MINUS, cval=5.0, affect=P.AFF_ADD, aval=(10.0,), rootfind=0)),
        bounce_none=(deg.GPUContinuousCallback(lambda u, t, i: u[0], bounce, rootfind=deg.NoRootFind),
                     P.continuous(P.COND_U0, affect=P.AFF_SCALE, aval=(0.0, -2.0), rootfind=0)),
        up_only=(deg.GPUContinuousCallback(lambda u, t, i: u[0], bounce, None),
                 P.continuous(P.COND_U0, affect=P.AFF_SCALE, affect_neg=None, aval=(0.0, -2.0))),
        height20=(deg.GPUContinuousCallback(lambda u, t, i: u[0] - 20.0, lift),
                  P.continuous(P.COND_U0_MINUS, cval=20.0, affect=P.AFF_ADD, aval=(1000.0, 0.0))),
        cross5=(deg.GPUContinuousCallback(lambda u, t, i: u[0] - 5.0, add10),
                P.continuous(P.COND_U0_MINUS, cval=5.0, affect=P.AFF_ADD, aval=(10.0,))),
        ball_stop7=(deg.GPUDiscreteCallback(lambda u, t, i: t == 7.0, stop), P.discrete(P.COND_T_EQ, 7.0, P.AFF_TERMINATE)),
    )
    return G


def _ensemble(system, N, seed):
    r = np.random.default_rng(seed)
    if system == "decay":
        return r.uniform(5.0, 15.0, (1, N)).astype(F32), r.uniform(0.5, 1.5, (1, N)).astype(F32), (0.0, 10.0)
    u0 = np.stack([r.uniform(20.0, 60.0, N), np.zeros(N)]).astype(F32)
    return u0, r.uniform(5.0, 15.0, (1, N)).astype(F32), (0.0, 15.0)


# attempt cap of every run here (GPU and oracle alike): a configuration whose events repeat
# forever in the reference (e.g. a bounce found just below the floor that re-triggers) ends with
# MaxIters after a bounded amount of work
MAX_ATTEMPTS = 20000


def _run(deg, torch, system, method, adaptive, names, *, dt, N=192, seed=0, saveat=None, save_everystep=None,
         tstops=None, persistent=False, abstol=1e-6, reltol=1e-3):
    from diffeqgpu_amd import systems

    G = _cases(deg)
    cbs = deg.CallbackSet(*[G[k][0] for k in names]) if names else None
    ocbs = [G[k][1] for k in names]
    f, n, np_, _ = systems.SYSTEMS[system]
    u0, p, tspan = _ensemble(system, N, seed)
    dev = deg.HIPDevice(contract="off", powmode="det", persistent=persistent, max_attempts=MAX_ATTEMPTS)
    probs = deg.HIPProbVector.from_arrays(f, u0, p, tspan, dev)
    prob = deg.ODEProblem(f, u0[:, 0], tspan, p[:, 0])
    alg = {"tsit5": deg.GPUTsit5, "vern7": deg.GPUVern7, "rb23": deg.GPURosenbrock23, "vern9": deg.GPUVern9,
           "rodas4": deg.GPURodas4, "rodas5p": deg.GPURodas5P, "kvaerno3": deg.GPUKvaerno3,
           "kvaerno5": deg.GPUKvaerno5}[method]()
    kw = dict(dt=dt, callback=cbs, tstops=tstops)
    if saveat is not None:
        kw["saveat"] = saveat
    if save_everystep is not None:
        kw["save_everystep"] = save_everystep
    print(f"[gpu] callbacks {system} {method} adaptive={adaptive} {names} dt={dt} saveat={saveat} "
          f"se={save_everystep} tstops={tstops} persistent={persistent}", flush=True)
    if adaptive:
        r = deg.vectorized_asolve(probs, prob, alg, abstol=abstol, reltol=reltol, **kw)
    else:
        r = deg.vectorized_solve(probs, prob, alg, **kw)
    torch.cuda.synchronize()
    nsave = r.ts.shape[0]
    sv = r.saveat.cpu().numpy() if r.saveat is not None else None
    se = bool(r.args.save_everystep)
    o = P.solve(system, method, adaptive, u0, p, tspan, dt, abstol=abstol, reltol=reltol, saveat=sv, nsave=nsave,
                save_everystep=se, detpow=True, tstops=tstops, callbacks=ocbs, max_attempts=MAX_ATTEMPTS)
    got = dict(ts=r.ts.cpu().numpy(), us=r.us_soa.cpu().numpy(), retcode=r.retcode.cpu().numpy(),
               naccept=r.naccept.cpu().numpy(), nreject=r.nreject.cpu().numpy())
    return got, o


def _assert_same(got, o, rows_written=None):
    for key in ("retcode", "naccept", "nreject", "ts"):
        np.testing.assert_array_equal(got[key], o[key], err_msg=key)
    # the oracle leaves never-written rows NaN (the reference leaves them uninitialised)
    mask = ~np.isnan(o["us"])
    np.testing.assert_array_equal(got["us"][mask], o["us"][mask], err_msg="us")


@pytest.mark.parametrize("method", ["tsit5", "vern7", "rb23", "vern9", "rodas4", "rodas5p", "kvaerno3", "kvaerno5"])
@pytest.mark.parametrize("persistent", [False, True])
def test_discrete_fixed_step(deg, torch_cuda, method, persistent):
    for names, tst, dt, kw in ((["jump24"], [2.4], 1.0, {}),
                               (["jump24", "jump40"], [2.4, 4.0], 1.0, {}),
                               (["jump24", "jump40"], [2.4, 4.0], 1.0, dict(saveat=[0.0, 6.0])),
                               (["jump24", "jump40"], [2.4, 4.0], 1.0, dict(save_everystep=False)),
                               (["jump24"], [4.0], 0.01, {}),
                               (["low"], None, 0.1, {})):
        got, o = _run(deg, torch_cuda, "decay", method, False, names, dt=dt, tstops=tst, persistent=persistent, **kw)
        _assert_same(got, o)


@pytest.mark.parametrize("method", ["tsit5", "vern7", "rb23", "vern9", "rodas4", "rodas5p", "kvaerno3", "kvaerno5"])
@pytest.mark.parametrize("persistent", [False, True])
def test_discrete_adaptive(deg, torch_cuda, method, persistent):
    for names, tst, kw in ((["jump24"], [4.0], {}),
                           (["jump24", "jump40"], [2.4, 4.0], {}),
                           (["jump24", "jump40"], [2.4, 4.0], dict(saveat=[0.0, 6.0], abstol=1e-7, reltol=1e-7)),
                           (["keep24", "jump40"], [2.4, 4.0], dict(saveat=np.linspace(0, 10, 21))),
                           (["low"], [1.0, 3.0, 5.0], dict(saveat=np.linspace(0, 10, 11)))):
        got, o = _run(deg, torch_cuda, "decay", method, True, names, dt=1.0, tstops=tst, persistent=persistent, **kw)
        _assert_same(got, o)


@pytest.mark.parametrize("method", ["tsit5", "vern7", "rb23", "rodas5p"])
def test_terminate(deg, torch_cuda, method):
    for adaptive in (False, True):
        got, o = _run(deg, torch_cuda, "decay", method, adaptive, ["stop24"], dt=1.0, tstops=[2.4])
        assert np.all(got["retcode"] == 4)
        _assert_same(got, o)


@pytest.mark.parametrize("method", ["tsit5", "vern7", "rb23", "rodas4", "rodas5p", "kvaerno3", "kvaerno5"])
@pytest.mark.parametrize("persistent", [False, True])
def test_continuous_ball(deg, torch_cuda, method, persistent):
    for adaptive, dt, names, kw in ((False, 0.1, ["bounce"], {}),
                                    (False, 1.0, ["bounce"], dict(saveat=[0.0, 9.1])),
                                    (False, 0.1, ["bounce"], dict(save_everystep=False)),
                                    (True, 0.1, ["bounce"], {}),
                                    (True, 0.1, ["bounce", "bounce"], {}),
                                    (True, 0.1, ["bounce_sciml"], dict(saveat=[0.0, 9.1], abstol=1e-6, reltol=1e-6)),
                                    (False, 0.1, ["bounce_none"], dict(saveat=np.linspace(0, 15, 31))),
                                    (True, 0.1, ["up_only"], {}),
                                    (True, 0.1, ["bounce", "height20"], dict(saveat=np.linspace(0, 15, 16))),
                                    (True, 0.1, ["bounce", "ball_stop7"], dict(tstops=[7.0]))):
        got, o = _run(deg, torch_cuda, "ball", method, adaptive, names, dt=dt, persistent=persistent, **kw)
        _assert_same(got, o)


@pytest.mark.parametrize("method", ["tsit5", "vern7"])
def test_repeated_crossings_itp(deg, torch_cuda, method):
    """u - 5 crossed again after every +10 affect; the callback saves record each event."""
    for adaptive, dt, kw in ((False, 0.25, {}), (True, 0.5, dict(saveat=np.linspace(0, 10, 41)))):
        for name in ("cross5", "cross5_right", "cross5_none"):
            got, o = _run(deg, torch_cuda, "decay", method, adaptive, [name], dt=dt, **kw)
            _assert_same(got, o)


def test_tstops_only(deg, torch_cuda):
    for method in ("tsit5", "vern7", "rb23", "vern9", "rodas4", "rodas5p", "kvaerno3", "kvaerno5"):
        for adaptive in (False, True):
            got, o = _run(deg, torch_cuda, "decay", method, adaptive, [], dt=0.3, tstops=[0.5, 1.25, 2.0, 7.77],
                          saveat=np.linspace(0, 10, 9) if adaptive else None)
            _assert_same(got, o)


# ---------------------------------------------------------------------------------------
# the reference's callback tests through the high-level solve
# ---------------------------------------------------------------------------------------
def _decay_exact(t, jumps):
    out = []
    for x in np.asarray(t, np.float64):
        u, t0 = 10.0, 0.0
        for tj in jumps:
            if x <= tj + 1e-7:
                break
            u = u * math.exp(-(tj - t0)) + 10.0
            t0 = tj
        out.append(u * math.exp(-(x - t0)))
    return np.array(out)


@pytest.mark.parametrize("alg_name", ["GPUTsit5", "GPUVern7", "GPUVern9"])
def test_reference_discrete_callbacks(deg, torch_cuda, alg_name):
    """test/gpu_kernel_de/gpu_ode_discrete_callbacks.jl, u' = -u, u(0) = 10 on (0, 10): the
    bounds it asserts against Vern9, here against the closed form."""
    S = deg.SVector
    f = lambda u, p, t: S(-u[0])  # noqa: E731
    prob = deg.ODEProblem(f, [10.0], (0.0, 10.0))
    mp = deg.EnsembleProblem(prob, safetycopy=False)
    alg = getattr(deg, alg_name)()
    ens = deg.EnsembleGPUKernel(deg.HIPDevice())

    def affect(i):
        i.u += S(10.0)

    cb = deg.DiscreteCallback(lambda u, t, i: t == 2.4, affect, save_positions=(False, False))
    cb2 = deg.DiscreteCallback(lambda u, t, i: t == 4.0, affect, save_positions=(False, False))
    cbs = deg.CallbackSet(cb, cb2)

    def run(**kw):
        print(f"[gpu] reference discrete {alg_name} {sorted(kw)}", flush=True)
        return deg.solve(mp, alg, ens, trajectories=2, merge_callbacks=True, **kw).u[0]

    def bench(s, jumps):
        # the exact solution saved at the same times as the GPU solution: what the reference's
        # Vern9 solution with the same dt / save settings stands for (sol(t) interpolates linearly
        # between saved points in both)
        return deg.ODESolution(s.t, _decay_exact(s.t, jumps)[:, None], "Success")

    s = run(adaptive=False, dt=1.0, callback=cb, tstops=[2.4])
    b = bench(s, [2.4])
    assert np.linalg.norm(b(2.4) - s(2.4)) < 2e-3 and s.t[3] == F32(2.4)
    assert np.linalg.norm(b.u - s.u) < 5e-3
    s = run(adaptive=False, dt=0.01, callback=cb, tstops=[4.0])
    b = bench(s, [])
    assert np.linalg.norm(b(4.0) - s(4.0)) < 2e-6
    assert np.linalg.norm(b(s.t[-1]) - s(s.t[-1])) < 3e-5
    s = run(adaptive=False, dt=1.0, callback=cbs, tstops=[2.4, 4.0])
    b = bench(s, [2.4, 4.0])
    assert np.linalg.norm(b(2.4) - s(2.4)) < 2e-3
    assert np.linalg.norm(b(4.0) - s(4.0)) < 3e-3
    assert np.linalg.norm(b.u - s.u) < 7e-3
    s = run(adaptive=False, dt=1.0, callback=cbs, tstops=[2.4, 4.0], saveat=[0.0, 6.0])
    assert np.linalg.norm(bench(s, [2.4, 4.0]).u - s.u) < 3e-3
    s = run(adaptive=False, dt=1.0, callback=cbs, tstops=[2.4, 4.0], save_everystep=False)
    assert np.linalg.norm(bench(s, [2.4, 4.0]).u - s.u) < 2e-4
    s = run(adaptive=True, dt=1.0, callback=cb, tstops=[4.0])
    assert np.linalg.norm(bench(s, [])(4.0) - s(4.0)) < 5e-5
    s = run(adaptive=True, dt=1.0, callback=cbs, tstops=[2.4, 4.0])
    assert np.linalg.norm(bench(s, [2.4, 4.0]).u - s.u) < 3e-3
    s = run(adaptive=True, dt=1.0, callback=cbs, tstops=[2.4, 4.0], saveat=[0.0, 6.0], reltol=1e-7, abstol=1e-7)
    assert np.linalg.norm(bench(s, [2.4, 4.0])(6.0) - s(6.0)) < 2e-2
    # without merge_callbacks the keyword callback is not used (ensemblegpuarray/callbacks.jl:63)
    print(f"[gpu] reference discrete {alg_name} no merge", flush=True)
    s = deg.solve(mp, alg, ens, trajectories=2, adaptive=False, dt=1.0, callback=cb, tstops=[2.4]).u[0]
    assert np.linalg.norm(bench(s, []).u - s.u) < 5e-3
    # terminate
    s = run(adaptive=False, dt=1.0, callback=deg.DiscreteCallback(lambda u, t, i: t == 2.4, deg.terminate,
                                                                   save_positions=(False, False)), tstops=[2.4])
    assert s.retcode == "Terminated" and s.t[-1] == F32(2.4)


@pytest.mark.parametrize("alg_name", ["GPUTsit5", "GPUVern7"])
def test_reference_continuous_callbacks(deg, torch_cuda, alg_name):
    """test/gpu_kernel_de/gpu_ode_continuous_callbacks.jl: the ball from 45 m under g = 10 with
    a bounce at u1 = 0, on (0, 15)."""
    S = deg.SVector
    f = lambda u, p, t: S(u[1], -p[0])  # noqa: E731
    prob = deg.ODEProblem(f, [45.0, 0.0], (0.0, 15.0), [10.0])
    mp = deg.EnsembleProblem(prob, safetycopy=False)
    alg = getattr(deg, alg_name)()
    ens = deg.EnsembleGPUKernel(deg.HIPDevice())

    def affect(i):
        i.u = i.u + S(0.0, -2.0) * i.u

    cb = deg.ContinuousCallback(lambda u, t, i: u[0], affect, save_positions=(False, False))

    def at_end(s):
        # u1 back on the floor at t = 15, |u2| = 30 (the bounce is at the end point itself)
        return abs(s.u[-1, 0]) + abs(abs(s.u[-1, 1]) - 30.0)

    def run(**kw):
        print(f"[gpu] reference continuous {alg_name} {sorted(kw)}", flush=True)
        return deg.solve(mp, alg, ens, trajectories=2, merge_callbacks=True, **kw).u[0]

    for c in (cb, deg.CallbackSet(cb, cb)):
        s = run(adaptive=False, dt=0.1, callback=c)
        assert at_end(s) < 2e-3 * 10 and s.retcode == "Success"
        s = run(adaptive=True, dt=0.1, callback=c)
        assert at_end(s) < 1e-2 * 10
    s = run(adaptive=False, dt=1.0, callback=cb, saveat=[0.0, 9.1])
    assert abs(s.u[-1, 0] - (30.0 * 0.1 - 5.0 * 0.01)) < 2e-3 and abs(s.u[-1, 1] - 29.0) < 2e-3
    s = run(adaptive=True, dt=0.1, callback=deg.CallbackSet(cb, cb), saveat=[0.0, 9.1], reltol=1e-6, abstol=1e-6)
    assert abs(s.u[-1, 0] - 2.95) < 2e-3 and abs(s.u[-1, 1] - 29.0) < 2e-3
    s = run(adaptive=False, dt=0.1, callback=cb, save_everystep=False)
    assert at_end(s) < 7e-4 * 10
