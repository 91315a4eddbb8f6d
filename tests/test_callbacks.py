"""Callbacks and tstops on the CPU: pins of the oracle's restatement of
integrators/integrator_utils.jl:52-479 and perform_step tstops handling against closed forms,
and the host side (callback types, conversion, merging, tracing, output sizing, JIT compile).

Closed forms used as pins:
  * u' = -u, u(0) = 10 with u += 10 at t = 2.4 (and 4.0): piecewise exponentials; the row saved
    at an event time holds the value before the affect (savevalues! runs before affect!);
  * the falling ball u1'' = -10 from 45 (test/gpu_kernel_de/gpu_ode_continuous_callbacks.jl):
    u1 = 45 - 5 s^2, u2 = -10 s with s the time since the nearest apex (bounces at 3, 9, 15);
  * the ITP event time of u - 5 on u = 10 e^-t is ln 2; LeftRootFind lands on the side where
    the condition keeps its sign before the event, RightRootFind on the other, NoRootFind at
    the step end;
  * affect! acts on upcrossings (condition rising through 0) and affect_neg! on downcrossings:
    without affect_neg! the falling ball passes the floor (45 - 5 t^2 throughout);
  * with two continuous callbacks the affect applied is the first one's, at the earliest event
    (integrator_utils.jl:296-310): the ball then bounces off height 20 at t = sqrt(5).
"""
import math

import numpy as np
import pytest

import pyoracle as P

F32 = np.float32


def _decay_exact(t, jumps):
    """u' = -u from 10 with +10 at each jump time; at a jump time the value before the jump."""
    t = np.asarray(t, dtype=np.float64)
    out = np.empty_like(t)
    for i, x in enumerate(t):
        u, t0 = 10.0, 0.0
        for tj in jumps:
            if x <= tj + 1e-7:
                break
            u = u * math.exp(-(tj - t0)) + 10.0
            t0 = tj
        out[i] = u * math.exp(-(x - t0))
    return out


def _ball_exact(t):
    t = np.asarray(t, dtype=np.float64)
    s = t - 6.0 * np.round(t / 6.0)
    return 45.0 - 5.0 * s * s, -10.0 * s


DECAY_CB = [P.discrete(P.COND_T_EQ, 2.4, P.AFF_ADD, (10.0,))]
BALL_CB = [P.continuous(P.COND_U0, affect=P.AFF_SCALE, aval=(0.0, -2.0))]


@pytest.mark.parametrize("method,tol", [("tsit5", 3e-3), ("vern7", 3e-5), ("rb23", 0.3)])
def test_discrete_jump_fixed_step_with_tstop(method, tol):
    """The reference's first discrete case: dt = 1, tstops = [2.4], t == 2.4 -> u += 10."""
    u0 = np.array([[10.0, 10.0]], F32)
    p = np.ones((1, 2), F32)
    o = P.solve("decay", method, False, u0, p, (0.0, 10.0), 1.0, nsave=12, save_everystep=True, tstops=[2.4],
                callbacks=DECAY_CB, detpow=True)
    t = o["ts"][:, 0]
    want_t = np.array([0, 1, 2, 2.4, 3.4, 4.4, 5.4, 6.4, 7.4, 8.4, 9.4, 10.0], F32)
    np.testing.assert_allclose(t, want_t, atol=2e-6)
    assert t[3] == F32(2.4)  # the step ends exactly on the stopping time
    assert np.all(o["retcode"] == 0)
    np.testing.assert_allclose(o["us"][:, 0, 0], _decay_exact(t, [2.4]), atol=tol)
    assert abs(o["us"][3, 0, 0] - 10 * math.exp(-2.4)) < tol  # saved before the affect


@pytest.mark.parametrize("adaptive", [False, True])
@pytest.mark.parametrize("method", ["tsit5", "vern7"])
def test_discrete_two_jumps_saveat(method, adaptive):
    cbs = [P.discrete(P.COND_T_EQ, 2.4, P.AFF_ADD, (10.0,)), P.discrete(P.COND_T_EQ, 4.0, P.AFF_ADD, (10.0,))]
    u0 = np.array([[10.0]], F32)
    sv = np.array([0.0, 2.4, 3.0, 6.0], F32)
    o = P.solve("decay", method, adaptive, u0, np.ones((1, 1), F32), (0.0, 10.0), 0.1 if adaptive else 0.05,
                saveat=sv, tstops=[2.4, 4.0], callbacks=cbs, detpow=True, abstol=1e-7, reltol=1e-7)
    assert o["retcode"][0] == 0
    np.testing.assert_array_equal(o["ts"][:, 0], sv)
    np.testing.assert_allclose(o["us"][:, 0, 0], _decay_exact(sv, [2.4, 4.0]), rtol=2e-5, atol=2e-5)


def test_tstops_without_callbacks_land_exactly():
    """Every stopping time becomes a step end (fixed dt 0.3 is off the grid of tstops)."""
    tst = np.array([0.5, 1.25, 2.0], F32)
    o = P.solve("decay", "tsit5", False, np.array([[1.0]], F32), np.ones((1, 1), F32), (0.0, 3.0), 0.3,
                nsave=14, save_everystep=True, tstops=tst)
    t = o["ts"][:, 0]
    for x in tst:
        assert x in t
    np.testing.assert_allclose(o["us"][:, 0, 0], np.exp(-t.astype(np.float64)), rtol=2e-5)


@pytest.mark.parametrize("method", ["tsit5", "vern7", "rb23"])
@pytest.mark.parametrize("adaptive,dt", [(False, 0.1), (True, 0.1)])
def test_bouncing_ball_closed_form(method, adaptive, dt):
    u0 = np.array([[45.0], [0.0]], F32)
    sv = np.array([0.0, 1.5, 4.5, 9.1, 12.0, 14.5], F32)
    o = P.solve("ball", method, adaptive, u0, np.full((1, 1), 10.0, F32), (0.0, 15.0), dt, saveat=sv,
                callbacks=BALL_CB, detpow=True, abstol=1e-6, reltol=1e-6)
    assert o["retcode"][0] == 0
    e1, e2 = _ball_exact(sv)
    np.testing.assert_allclose(o["us"][:, 0, 0], e1, atol=2e-3)
    np.testing.assert_allclose(o["us"][:, 1, 0], e2, atol=2e-3)


def test_affect_neg_handles_downcrossings():
    u0 = np.array([[45.0], [0.0]], F32)
    sv = np.array([0.0, 2.0, 4.0, 5.0], F32)
    p = np.full((1, 1), 10.0, F32)
    only_up = [P.continuous(P.COND_U0, affect=P.AFF_SCALE, affect_neg=None, aval=(0.0, -2.0))]
    o = P.solve("ball", "tsit5", True, u0, p, (0.0, 5.0), 0.1, saveat=sv, callbacks=only_up, detpow=True)
    np.testing.assert_allclose(o["us"][:, 0, 0], 45.0 - 5.0 * sv.astype(np.float64) ** 2, atol=1e-3)
    only_down = [P.continuous(P.COND_U0, affect=None, affect_neg=P.AFF_SCALE, aval=(0.0, -2.0))]
    o = P.solve("ball", "tsit5", True, u0, p, (0.0, 5.0), 0.1, saveat=sv, callbacks=only_down, detpow=True)
    np.testing.assert_allclose(o["us"][:, 0, 0], _ball_exact(sv)[0], atol=1e-3)


@pytest.mark.parametrize("rootfind", [1, 2, 0])
def test_event_time_from_itp(rootfind):
    """u = 10 e^-t crosses 5 at ln 2 inside the step 0.5 -> 0.75; the callback's save records
    the event time and the state before the affect (u += 10)."""
    cb = [P.continuous(P.COND_U0_MINUS, cval=5.0, affect=P.AFF_ADD, aval=(10.0,), rootfind=rootfind)]
    o = P.solve("decay", "tsit5", False, np.array([[10.0]], F32), np.ones((1, 1), F32), (0.0, 1.5), 0.25,
                nsave=7, save_everystep=True, callbacks=cb, detpow=True)
    t, u = o["ts"][:, 0].astype(np.float64), o["us"][:, 0, 0].astype(np.float64)
    te, ue = t[3], u[3]
    if rootfind == 0:
        assert te == 0.75  # NoRootFind: the step end
    else:
        assert abs(te - math.log(2.0)) < 2e-6
        assert abs(ue - 5.0) < 2e-5
        # the bracket side: left keeps the sign the condition had before the event
        assert (ue - 5.0 >= 0.0) if rootfind == 1 else (ue - 5.0 <= 0.0)
    # after the event: 15 e^-(t - te)
    np.testing.assert_allclose(u[4:], (ue + 10.0) * np.exp(-(t[4:] - te)), rtol=5e-5)


def test_first_continuous_callbacks_affect_applies():
    """CallbackSet(ground, height 20): the earlier event (height 20 at sqrt 5) applies the first
    callback's affect (bounce)."""
    cbs = [P.continuous(P.COND_U0, affect=P.AFF_SCALE, aval=(0.0, -2.0)),
           P.continuous(P.COND_U0_MINUS, cval=20.0, affect=P.AFF_ADD, aval=(1000.0, 0.0))]
    sv = np.array([1.0, 3.0, 4.0], F32)
    o = P.solve("ball", "tsit5", True, np.array([[45.0], [0.0]], F32), np.full((1, 1), 10.0, F32), (0.0, 4.0), 0.1,
                saveat=sv, callbacks=cbs, detpow=True, abstol=1e-7, reltol=1e-7)
    t1 = math.sqrt(5.0)
    v = 10.0 * t1
    s = sv.astype(np.float64) - t1
    want = np.where(s < 0, 45.0 - 5.0 * sv.astype(np.float64) ** 2, 20.0 + v * s - 5.0 * s * s)
    np.testing.assert_allclose(o["us"][:, 0, 0], want, atol=2e-3)


@pytest.mark.parametrize("save_everystep", [True, False])
def test_terminate(save_everystep):
    cb = [P.discrete(P.COND_T_EQ, 2.4, P.AFF_TERMINATE)]
    o = P.solve("decay", "tsit5", False, np.array([[10.0]], F32), np.ones((1, 1), F32), (0.0, 10.0), 1.0,
                nsave=12 if save_everystep else 2, save_everystep=save_everystep, tstops=[2.4], callbacks=cb)
    assert o["retcode"][0] == 4
    t = o["ts"][:, 0]
    if save_everystep:
        np.testing.assert_array_equal(t[:4], np.array([0, 1, 2, 2.4], F32))
        assert np.all(t[4:] == 0.0)  # unsaved rows keep the t0 prefill
    else:
        assert t[1] == F32(2.4)
    assert o["naccept"][0] == 3


# ---------------------------------------------------------------------------------------
# host side
# ---------------------------------------------------------------------------------------
def test_callback_types_and_conversion(deg):
    C = deg.callbacks
    with pytest.raises(ValueError, match="save_positions"):
        deg.GPUDiscreteCallback(lambda u, t, i: t == 1.0, lambda i: None, save_positions=(True, False))
    cc = deg.ContinuousCallback(lambda u, t, i: u[0], lambda i: None, save_positions=(False, False))
    g = C.to_gpu(cc)
    assert isinstance(g, deg.GPUContinuousCallback)
    assert g.abstol == 100 * C.EPS32 and g.affect_neg is cc.affect and g.rootfind is deg.LeftRootFind
    assert deg.GPUContinuousCallback(lambda u, t, i: u[0], None).abstol == 10 * C.EPS32
    with pytest.raises(ValueError, match="save_positions"):
        C.to_gpu(deg.DiscreteCallback(lambda u, t, i: t == 1.0, lambda i: None))
    d = deg.GPUDiscreteCallback(lambda u, t, i: t == 1.0, lambda i: None)
    cs = deg.CallbackSet(cc, deg.CallbackSet(d, None), [g])
    assert len(cs.continuous_callbacks) == 2 and cs.discrete_callbacks == [d]
    # merge_callbacks: the keyword callback only counts when merging (callbacks.jl:61-74)
    assert C.merged_callback(None, d, False) is None
    m = C.merged_callback(g, d, True)
    assert m.continuous_callbacks == [g] and m.discrete_callbacks == [d]
    with pytest.raises(TypeError):
        deg.CallbackSet(object())


def test_condition_and_affect_tracing(deg):
    C = deg.callbacks

    def aff(integ):
        integ.u = integ.u + deg.SVector(0.0, -2.0) * integ.u
        integ.p = [integ.p[0] * 2.0]
        deg.terminate(integ)

    cs = deg.CallbackSet(deg.GPUDiscreteCallback(lambda u, t, i: (t == 2.4) | ~(u[0] < 1.0), aff),
                         deg.GPUContinuousCallback(lambda u, t, i: u[0] - i.p[0], aff, None,
                                                   rootfind=deg.RightRootFind))
    src = C.emit_callbacks(cs, 2, 1)
    assert "#define DEG_NDCB 1" in src and "#define DEG_NCCB 1" in src
    assert "(t == (float)2.4)" in src and "(u[0] < 1.0f)" in src and "||" in src
    assert "u[1] + v" in src  # u + (a .* u), left operand first
    assert "return 1;" in src  # terminate flag
    assert "return 9;" in src  # mode: affect only (1) | RightRootFind (2 << 2)
    with pytest.raises(deg.tracer.TraceError):
        C.emit_callbacks(deg.CallbackSet(deg.GPUDiscreteCallback(lambda u, t, i: t == 1.0 and u[0] > 0,
                                                                 lambda i: None)), 2, 1)
    with pytest.raises(deg.tracer.TraceError):
        def bad(integ):
            integ.t = 1.0
        C.emit_callbacks(deg.CallbackSet(deg.GPUDiscreteCallback(lambda u, t, i: t == 1.0, bad)), 2, 1)


@pytest.mark.parametrize("method", [0, 1, 2])
@pytest.mark.parametrize("adaptive", [False, True])
def test_callback_kernels_compile(deg, method, adaptive):
    """Every method compiles with discrete + continuous callbacks and tstops (hiprtc, no GPU)."""
    from diffeqgpu_amd import _lib, systems

    C = deg.callbacks
    f, n, np_, _ = systems.SYSTEMS["ball"]
    cs = deg.CallbackSet(deg.GPUContinuousCallback(lambda u, t, i: u[0], lambda i: None),
                         deg.GPUDiscreteCallback(lambda u, t, i: t == 2.0, lambda i: deg.terminate(i)))
    src = f.hip_source(n, np_, method == 2)[0] + C.emit_callbacks(cs, n, np_)
    flags = _lib.F_TSTOPS | _lib.F_CALLBACKS | _lib.F_PARITY | (_lib.F_ADAPTIVE | _lib.F_DETPOW if adaptive else 0)
    L = _lib.lib()
    assert L.deg_jit_precompile(src.encode(), n, np_, method, flags, None) == 0, L.deg_last_error()
    # the flag must match the source
    assert L.deg_jit_precompile(src.encode(), n, np_, method, flags & ~_lib.F_CALLBACKS, None) != 0
    assert L.deg_jit_precompile(src.encode(), n, np_, method, flags | _lib.F_COMPACT | _lib.F_PERSISTENT
                                | _lib.F_ADAPTIVE, None) != 0


def test_fixed_step_rows_with_tstops(deg):
    """lowerlevel_solve.jl:73-77: each stopping time off the t0:dt:tf grid adds a row."""
    from diffeqgpu_amd import lowerlevel as LL

    t0, dt, tf = F32(0.0), F32(1.0), F32(10.0)
    grid = deg.colon(t0, dt, tf)
    tst = LL._tstops([2.4, 4.0])
    assert len(grid) == 11
    assert len(tst) - int(np.isin(grid.collect().astype(F32), tst).sum()) == 1
    assert LL._tstops([]) is None and LL._tstops(None) is None


def test_solution_linear_interpolation(deg):
    sol = deg.ODESolution(np.array([0.0, 1.0, 1.0, 2.0], F32), np.array([[0.0], [2.0], [5.0], [7.0]], F32), "Success")
    assert sol(0.5)[0] == 1.0
    assert sol(1.0)[0] == 2.0  # the first entry saved at that time
    assert sol(1.5)[0] == 6.0
    with pytest.raises(ValueError):
        sol(2.5)
