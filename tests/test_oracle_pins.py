"""Pins for the CPU oracle (oracle/deg_oracle.c) before it is trusted as the GPU checker.

The reference holds no golden vectors (its tests recompute an OrdinaryDiffEq oracle at run
time), so the oracle is pinned by what the mathematics fixes and by the reference's own
assertions re-expressed against an independent FP64 solution:
  * Butcher order conditions of every tableau (all rooted trees up to the method order),
    the embedded pairs and the Tsit5 dense output;
  * closed forms of the reference test problems (u' = -p u, u' = -p u^2, u' = u) and of
    u' = p t u and a rotation, with the empirical convergence order under step halving;
  * the reference's regression tolerances (test/gpu_kernel_de/gpu_ode_regression.jl,
    stiff_ode/gpu_ode_regression.jl, finite_diff.jl) against FP64 Vern9 / closed forms;
  * the FP64 Vern9 oracle against scipy's DOP853 and Radau;
  * the fixed-step time grid (100 steps of 0.1f0 end at 10.000002f0) and the controller
    invariants (every accepted step has EEst <= 1, counted via accept/reject counters).
"""
import math

import numpy as np
import pytest
from scipy.integrate import solve_ivp

import _butcher as B
import pyoracle

F32 = np.float32


# ---------------------------------------------------------------------------------------
# tableaus
# ---------------------------------------------------------------------------------------
@pytest.mark.parametrize("name,order", [("tsit5", 5), ("vern7", 7), ("vern9", 9)])
def test_tableau_row_sums_and_order(name, order):
    c, A, b, bt = pyoracle.tableau(name)
    np.testing.assert_allclose(A.sum(axis=1), c, atol=2e-14)
    if name == "tsit5":
        # FSAL: the last row of A is b
        np.testing.assert_allclose(A[-1, :-1], b[:-1], atol=0)
        assert b[-1] == 0.0
    res = B.order_residuals(A, b, order)
    assert max(res.values()) < 1e-12, res
    # the method is not of order p+1 (a planted wrong coefficient would usually break the above)
    extra = max(abs(b @ B.phi(A, t) - 1.0 / B.gamma(t)) for t in B.trees(order + 1))
    assert extra > 1e-6


@pytest.mark.parametrize("name,order", [("tsit5", 4), ("vern7", 6), ("vern9", 8)])
def test_embedded_pair(name, order):
    c, A, b, bt = pyoracle.tableau(name)
    best = None
    for sgn in (+1, -1):
        bhat = b + sgn * bt
        r = B.order_residuals(A, bhat, order)
        best = r if best is None or max(r.values()) < max(best.values()) else best
    assert max(best.values()) < 1e-12, best
    assert np.abs(bt).max() > 1e-4


def test_tableau_perturbation_is_detected():
    """A single transposed or sign-flipped coefficient breaks the order conditions."""
    c, A, b, bt = pyoracle.tableau("vern7")
    for (i, j, k, l) in [(2, 0, 2, 1), (5, 3, 5, 4)]:
        A2 = A.copy()
        A2[i, j], A2[k, l] = A[k, l], A[i, j]
        assert max(B.order_residuals(A2, b, 7).values()) > 1e-6
    b2 = b.copy()
    b2[3] = -b2[3]
    assert max(B.order_residuals(A, b2, 7).values()) > 1e-6


def test_tsit5_dense_output():
    c, A, b, bt = pyoracle.tableau("tsit5")
    R = pyoracle.tsit5_r()

    def bth(th):
        out = np.zeros(7)
        out[0] = th * (R[1, 1] + th * (R[1, 2] + th * (R[1, 3] + th * R[1, 4])))
        for i in range(2, 8):
            out[i - 1] = th * th * (R[i, 2] + th * (R[i, 3] + th * R[i, 4]))
        return out

    np.testing.assert_allclose(bth(1.0), b, atol=1e-13)
    assert np.all(bth(0.0) == 0.0)
    # continuous extension of order 4: sum_i b_i(th) Phi_i(t) = th^|t| / gamma(t)
    for th in (0.3, 0.7):
        w = bth(th)
        for q in range(1, 5):
            for t in B.trees(q):
                assert abs(w @ B.phi(A, t) - th ** q / B.gamma(t)) < 1e-12
    # the oracle evaluates the same polynomials in Float32 (fused Horner)
    for th in (0.0, 0.25, 0.5, 1.0):
        np.testing.assert_allclose(pyoracle.tsit5_btheta(th), bth(th), rtol=0, atol=3e-6)
    assert np.all(pyoracle.tsit5_btheta(0.0) == 0.0)


def test_vern9_dense_output():
    """The Vern9 dense output (ten extra stages, degree-9 weights) is a continuous extension of
    order 9: b(1) = b with zero weight on the extra stages, b(0) = 0, the extra stages' nodes are
    their row sums, and sum_i b_i(th) Phi_i(t) = th^|t| / gamma(t) for every tree up to order 9
    at interior th over the 26-stage extended tableau."""
    c, A, b, bt = pyoracle.tableau("vern9")
    C, AX, R = pyoracle.vern9_dense()
    Ae = np.zeros((26, 26))
    Ae[:16, :16] = A
    Ae[16:, :] = AX[17:27, 1:27]
    np.testing.assert_allclose(Ae[16:].sum(axis=1), C[17:27], atol=1e-14)
    stages = [1] + list(range(8, 16)) + list(range(17, 27))

    def bth(th):
        w = np.zeros(26)
        for i in stages:
            w[i - 1] = sum(R[i, k] * th ** k for k in range(10))
        return w

    np.testing.assert_allclose(bth(1.0)[:16], b, atol=1e-11)
    assert np.abs(bth(1.0)[16:]).max() < 1e-11 and np.all(bth(0.0) == 0.0)
    for th in (0.3, 0.7):
        w = bth(th)
        for q in range(1, 10):
            for t in B.trees(q):
                assert abs(w @ B.phi(Ae, t) - th ** q / B.gamma(t)) < 1e-12, (th, q)
    # a coefficient off in its last digits breaks the order-9 conditions
    R2 = R.copy()
    R2[12, 5] *= 1 + 1e-9
    w = np.zeros(26)
    for i in stages:
        w[i - 1] = sum(R2[i, k] * 0.5 ** k for k in range(10))
    assert max(abs(w @ B.phi(Ae, t) - 0.5 ** 5 / B.gamma(t)) for t in B.trees(5)) > 1e-9
    # the oracle evaluates the same polynomials in Float32 (fused Horner): within the Horner
    # rounding bound 2 * 9 * eps32 * sum_k |r_ik| th^k
    eps = float(np.finfo(np.float32).eps)
    for th in (0.0, 0.25, 0.5, 1.0):
        got = pyoracle.vern9_btheta(th)[1:]
        bound = np.array([18 * eps * sum(abs(R[i, k]) * th ** k for k in range(10)) for i in range(1, 27)])
        assert np.all(np.abs(got - bth(th)) <= bound + 1e-12), th
    # a saveat point at a step end (th = 1) reproduces the step's solution up to the Float32
    # cancellation of the degree-9 weights (coefficients up to 6e3: about 1e-3 relative here)
    sv = np.array([0.0, 0.05, 0.1], F32)
    r = pyoracle.solve("lorenz", "vern9", False, LZ_U0, LZ_P, (0.0, 0.1), 0.05, saveat=sv)
    r2 = pyoracle.solve("lorenz", "vern9", False, LZ_U0, LZ_P, (0.0, 0.1), 0.05, nsave=3, save_everystep=True)
    assert np.all(np.isfinite(r2["us"]))
    np.testing.assert_allclose(r["us"][1:], r2["us"][1:], rtol=2e-3)


# ---------------------------------------------------------------------------------------
# Rodas4 / Rodas5P
# ---------------------------------------------------------------------------------------
def _rodas_tab(tag):
    import os
    import re

    src = open(os.path.join(os.path.dirname(pyoracle.__file__), "tableaus.h")).read()
    return {m.group(1): float(m.group(2)) for m in re.finditer(rf"#define {tag}_(\w+) (\S+)", src)}


def _rodas_step64(T, stages, f, jac, tgrad, u0, t, dt):
    """One Rosenbrock step in FP64 in the generic form: W = J - I/(dt gamma); stage i solves
    W K_i = -(f(U_i, t + c_i dt) + d_i dt f_t + sum_j C_ij / dt K_j) with U_i = uprev + sum_j a_ij K_j
    while a_i exists, then U_i = U_(i-1) + K_(i-1) at t + dt; the solution is U_s + K_s."""
    n = len(u0)
    W = jac(u0, t) - np.eye(n) / (dt * T["gamma"])
    dT = tgrad(u0, t)
    K = []
    U = u0
    for i in range(1, stages + 1):
        if i > 1:
            if f"a{i}1" in T:
                U = u0 + sum(T[f"a{i}{j}"] * K[j - 1] for j in range(1, i))
            else:
                U = U + K[-1]
        ti = t + T.get(f"c{i}", 0.0 if i == 1 else 1.0) * dt
        rhs = f(U, ti) + T.get(f"d{i}", 0.0) * dt * dT + sum(T.get(f"C{i}{j}", 0.0) / dt * K[j - 1] for j in range(1, i))
        K.append(np.linalg.solve(W, -rhs))
    return U + K[-1]


@pytest.mark.parametrize("tag,stages,order", [("ORO4", 6, 4), ("ORO5", 8, 5)])
def test_rodas_tableaus_convergence(tag, stages, order):
    """The Rodas coefficients (a, C, gamma, c, d) give the method's order in FP64 on a rotation
    and on the non-autonomous u' = p t u (exercises c_i and d_i); a perturbed C fails it."""
    T = _rodas_tab(tag)
    problems = [
        (lambda u, t: np.array([-u[1], u[0]]), lambda u, t: np.array([[0.0, -1.0], [1.0, 0.0]]),
         lambda u, t: np.zeros(2), np.array([1.0, 0.0]), 8.0, np.array([math.cos(8.0), math.sin(8.0)])),
        (lambda u, t: 0.5 * t * u, lambda u, t: np.array([[0.5 * t]]), lambda u, t: 0.5 * u, np.array([1.0]), 2.0,
         np.array([math.exp(0.25 * 4.0)])),
    ]

    def err(Tab, prob, dt):
        f, jac, tg, u, tf, ex = prob
        t = 0.0
        for _ in range(int(round(tf / dt))):
            u = _rodas_step64(Tab, stages, f, jac, tg, u, t, dt)
            t += dt
        return np.abs(u - ex).max()

    for prob in problems:
        e1, e2 = err(T, prob, 0.125), err(T, prob, 0.0625)
        assert math.log2(e1 / e2) > order - 0.5, (e1, e2)
    T2 = dict(T)
    T2["C43"] *= 1.001
    e1, e2 = err(T2, problems[1], 0.125), err(T2, problems[1], 0.0625)
    assert math.log2(e1 / e2) < order - 0.5


@pytest.mark.parametrize("method,tag,stages", [("rodas4", "ORO4", 6), ("rodas5p", "ORO5", 8)])
def test_rodas_oracle_step_vs_fp64(method, tag, stages):
    """The Float32 oracle's Rodas step agrees with the FP64 generic form to Float32 rounding
    (u' = -u^2 from 1 and u' = p t u from 1 at t0 = 0.5)."""
    T = _rodas_tab(tag)
    for system, u0, p, t0, f, jac, tg in (
            ("quad", 1.0, 1.0, 0.0, lambda u, t: -u * u, lambda u, t: np.array([[-2.0 * u[0]]]), lambda u, t: np.zeros(1)),
            ("tdep", 1.0, 0.5, 0.5, lambda u, t: 0.5 * t * u, lambda u, t: np.array([[0.5 * t]]), lambda u, t: 0.5 * u)):
        for dt in (0.25, 0.05):
            want = _rodas_step64(T, stages, f, jac, tg, np.array([u0]), t0, dt)
            r = pyoracle.solve(system, method, False, np.array([[u0]], F32), np.array([[p]], F32), (t0, t0 + dt), dt)
            assert abs(float(r["us"][1, 0, 0]) - want[0]) < 2e-6, (system, dt)


@pytest.mark.parametrize("method", ["kvaerno3", "kvaerno5"])
def test_kvaerno_adaptive_closed_forms(method):
    """Adaptive ESDIRK with Newton-solved stages and Hermite dense output: the rotation (exact
    circle) and u' = -u^2 from 1 (1 / (1 + t)) at interior save points. (The reference runs the
    second stage's nonlinear solve at t + gamma dt, not at its node 2 gamma: on an autonomous
    right-hand side that is invisible, so these problems pin the method itself.)"""
    sv = np.linspace(0.0, 8.0, 17).astype(F32)
    r = pyoracle.solve("rot", method, True, np.array([[1.0], [0.0]], F32), np.array([[1.0]], F32), (0.0, 8.0), 0.01,
                       abstol=1e-8, reltol=1e-8, saveat=sv, detpow=True)
    assert r["retcode"][0] == 0
    t = sv.astype(np.float64)
    np.testing.assert_allclose(r["us"][:, 0, 0], np.cos(t), atol=2e-5)
    np.testing.assert_allclose(r["us"][:, 1, 0], np.sin(t), atol=2e-5)
    r = pyoracle.solve("quad", method, True, np.array([[1.0]], F32), np.array([[1.0]], F32), (0.0, 8.0), 0.01,
                       abstol=1e-8, reltol=1e-8, saveat=sv, detpow=True)
    np.testing.assert_allclose(r["us"][:, 0, 0], 1.0 / (1.0 + t), atol=2e-5)


def test_rodas_dense_output_endpoints():
    """(1-th) y0 + th (y1 + (1-th)(...)) is y0 at th = 0 and y1 at th = 1 exactly in real
    arithmetic: saveat at a step end reproduces the every-step value."""
    for method in ("rodas4", "rodas5p", "kvaerno3", "kvaerno5"):
        sv = np.array([0.0, 0.05, 0.1], F32)
        r = pyoracle.solve("lorenz", method, False, LZ_U0, LZ_P, (0.0, 0.1), 0.05, saveat=sv)
        r2 = pyoracle.solve("lorenz", method, False, LZ_U0, LZ_P, (0.0, 0.1), 0.05, nsave=3, save_everystep=True)
        np.testing.assert_allclose(r["us"][1:], r2["us"][1:], rtol=1e-6, atol=1e-6)


# ---------------------------------------------------------------------------------------
# deterministic pow of the controller
# ---------------------------------------------------------------------------------------
def test_det_pow_accuracy():
    xs = [1e-38, 3.3e-20, 1e-10, 1e-4, 0.012, 0.5, 0.9, 1.0, 1.7, 2.0, 37.0, 1e5, 1e20, 3e38]
    ys = [0.02, 0.0571428, 0.08, 0.1, 0.14, 0.2, 0.35]
    for x in xs:
        for y in ys:
            got = pyoracle.det_pow(x, y)
            want = float(F32(x)) ** float(F32(y))
            assert abs(got - want) <= 3e-6 * want, (x, y, got, want)
    assert pyoracle.det_pow(0.0, 0.14) == 0.0
    assert math.isinf(pyoracle.det_pow(float("inf"), 0.14))
    assert math.isnan(pyoracle.det_pow(float("nan"), 0.14))
    # subnormal input
    x = float(F32(1e-42))
    want = x ** float(F32(0.2))
    assert abs(pyoracle.det_pow(x, 0.2) - want) < 3e-6 * want


# ---------------------------------------------------------------------------------------
# time grid facts
# ---------------------------------------------------------------------------------------
def lorenz_ref_inputs(N, seed=1):
    rng = np.random.default_rng(seed)
    p = (np.array([10.0, 28.0, F32(8.0) / F32(3.0)], F32)[:, None] * rng.random((3, N), dtype=np.float32))
    u0 = np.zeros((3, N), F32)
    u0[0] = 1
    return u0, p.astype(F32)


def test_fixed_step_time_grid():
    u0, p = lorenz_ref_inputs(16)
    r = pyoracle.solve("lorenz", "tsit5", False, u0, p, (0.0, 10.0), 0.1)
    assert np.all(r["naccept"] == 100)
    assert np.all(r["ts"][1] == F32(10.000002))
    assert np.all(r["ts"][0] == 0.0)
    np.testing.assert_array_equal(r["us"][0], u0)


# ---------------------------------------------------------------------------------------
# closed forms and convergence order
# ---------------------------------------------------------------------------------------
def _endpoint(system, method, u0, p, tspan, dt, adaptive=False, **kw):
    u0 = np.asarray(u0, F32).reshape(-1, 1)
    p = np.asarray(p, F32).reshape(-1, 1)
    r = pyoracle.solve(system, method, adaptive, u0, p, tspan, dt, **kw)
    return r["us"][1, :, 0].astype(np.float64), r


@pytest.mark.parametrize("method,order,system,dts", [("tsit5", 5, "rot", (0.5, 0.25)), ("vern7", 7, "quad", (1.0, 0.5)),
                                                    ("rb23", 2, "rot", (0.0625, 0.03125)), ("vern9", 9, "rot", (2.0, 1.0)),
                                                    ("kvaerno3", 3, "rot", (0.25, 0.125)),
                                                    ("kvaerno5", 5, "rot", (0.5, 0.25))])
def test_convergence_order(method, order, system, dts):
    """Empirical order under step halving on closed-form problems over (0, 8) with binary-exact
    steps: rotation u1' = -u2, u2' = u1 and u' = -u^2, u(0) = 1 (1 / (1 + t))."""
    T = 8.0
    if system == "rot":
        u0, p, exact = [1.0, 0.0], [1.0], np.array([math.cos(T), math.sin(T)])
    else:
        u0, p, exact = [1.0], [1.0], np.array([1.0 / (1.0 + T)])
    errs = []
    for dt in dts:
        u, r = _endpoint(system, method, u0, p, (0.0, T), dt)
        assert r["ts"][1, 0] == T and r["naccept"][0] == round(T / dt)
        errs.append(np.abs(u - exact).max())
    rate = math.log2(errs[0] / errs[1])
    assert abs(rate - order) < 1.0, (errs, rate)


@pytest.mark.parametrize("method", ["tsit5", "vern7", "rb23", "vern9"])
def test_nonautonomous_closed_form(method):
    # u' = p t u, u(0) = 1 -> exp(p t^2 / 2); exercises the stage times t + c_i dt
    p0, T = 0.5, 2.0
    u, r = _endpoint("tdep", method, [1.0], [p0], (0.0, T), 0.015625)
    assert r["ts"][1, 0] == T
    exact = math.exp(p0 * T * T / 2)
    tol = {"tsit5": 2e-6, "vern7": 2e-6, "rb23": 2e-4, "vern9": 2e-6}[method]
    assert abs(u[0] - exact) / exact < tol


def test_decay_closed_forms_stiff_regression():
    # stiff_ode/gpu_ode_regression.jl:5-20,58-59: u' = -p u, u0 = 10, p = 1, tspan (0, 10)
    exact = 10.0 * math.exp(-10.0)
    u, _ = _endpoint("decay", "rb23", [10.0], [1.0], (0.0, 10.0), 0.01)
    assert abs(u[0] - exact) < 5e-3
    u, r = _endpoint("decay", "rb23", [10.0], [1.0], (0.0, 10.0), 0.01, adaptive=True, abstol=1e-7, reltol=1e-7)
    assert abs(u[0] - exact) < 6e-3
    assert r["retcode"][0] == 0


def test_quad_closed_form_finite_diff_problem():
    # finite_diff.jl:6-32: u' = -p u^2, u0 = 10, p = 1 -> 10 / (1 + 10 t); tolerance 2e-4 vs Rodas5P
    exact = 10.0 / (1.0 + 100.0)
    u, r = _endpoint("quad", "rb23", [10.0], [1.0], (0.0, 10.0), 0.01, adaptive=True, abstol=1e-6, reltol=1e-3)
    assert abs(u[0] - exact) < 2e-4 * 10, (u, exact)
    for method in ("tsit5", "vern7"):
        u, r = _endpoint("quad", method, [10.0], [1.0], (0.0, 10.0), 0.01, adaptive=True, abstol=1e-8, reltol=1e-8)
        assert abs(u[0] - exact) < 2e-4


def test_growth_public_interface_problem():
    # public_interface.jl:9-41: u' = u, u0 = 1, tspan (0, 0.2)
    u, r = _endpoint("growth", "tsit5", [1.0], [1.0], (0.0, 0.2), 0.1)
    assert abs(u[0] - math.exp(0.2)) < 1e-6
    assert r["ts"][1, 0] == F32(0.2) or r["ts"][1, 0] >= F32(0.2)
    np.testing.assert_array_equal(r["us"][0, :, 0], [1.0])


# ---------------------------------------------------------------------------------------
# FP64 Vern9 oracle vs scipy
# ---------------------------------------------------------------------------------------
def test_vern9_matches_dop853():
    u0, p = lorenz_ref_inputs(6, seed=3)
    out, st, ns = pyoracle.vern9("lorenz", u0, p, (0.0, 10.0), saveat=[2.0, 5.0, 10.0])
    assert np.all(st == 0)
    for i in range(6):
        pp = p[:, i].astype(np.float64)

        def f(t, u):
            return [pp[0] * (u[1] - u[0]), u[0] * (pp[1] - u[2]) - u[1], u[0] * u[1] - pp[2] * u[2]]

        s = solve_ivp(f, (0, 10), [1.0, 0.0, 0.0], method="DOP853", rtol=1e-13, atol=1e-13, t_eval=[2.0, 5.0, 10.0])
        scale = max(1.0, np.abs(s.y).max())
        assert np.abs(out[:, :, i] - s.y.T).max() < 1e-8 * scale


def test_robertson_rb23_vs_radau_and_mass():
    """Rosenbrock23 with the FP64 W factorisation on the stiff Robertson kinetics (the 1e6
    stiff workload's problem and tolerances: reltol 1e-4, abstol 1e-8, 20 log-spaced save
    times to 1e5) against scipy Radau at rtol 1e-12, and the linear invariant y1 + y2 + y3 = 1
    that every Rosenbrock step preserves up to rounding."""
    sv = (10.0 ** (-4 + 9 * np.arange(20) / 19)).astype(F32)
    sv[-1] = F32(1e5)
    ks = np.array([[0.04, 0.05, 0.03], [3e7, 2e7, 4.5e7], [1e4, 1.3e4, 0.8e4]], F32)
    u0 = np.zeros((3, 3), F32)
    u0[0] = 1
    r = pyoracle.solve("robertson", "rb23", True, u0, ks, (0.0, 1e5), 1e-6, abstol=1e-8, reltol=1e-4, saveat=sv)
    assert np.all(r["retcode"] == 0)
    for i in range(3):
        k = ks[:, i].astype(np.float64)

        def f(t, y):
            return [-k[0] * y[0] + k[2] * y[1] * y[2], k[0] * y[0] - k[1] * y[1] ** 2 - k[2] * y[1] * y[2],
                    k[1] * y[1] ** 2]

        s = solve_ivp(f, (0, 1e5), [1.0, 0.0, 0.0], method="Radau", rtol=1e-12, atol=1e-20,
                      t_eval=sv.astype(np.float64))
        got = r["us"][:, :, i].astype(np.float64)
        err = np.abs(got - s.y.T).max(axis=0)
        assert err[0] < 1e-4 and err[2] < 1e-4 and err[1] < 1e-7, err
        assert np.abs(got.sum(axis=1) - 1.0).max() < 1e-5


def test_vern9_vdp_matches_radau():
    mu = np.array([[0.5, 5.0]], F32)
    u0 = np.array([[1.5, -1.0], [0.0, 0.5]], F32)
    out, st, ns = pyoracle.vern9("vdp", u0, mu, (0.0, 10.0))
    for i in range(2):
        m = float(mu[0, i])
        s = solve_ivp(lambda t, u: [u[1], m * (1 - u[0] ** 2) * u[1] - u[0]], (0, 10), u0[:, i].astype(float),
                      method="Radau", rtol=1e-12, atol=1e-12)
        assert np.abs(out[0, :, i] - s.y[:, -1]).max() < 1e-7


# ---------------------------------------------------------------------------------------
# the reference's Lorenz regression tolerances (gpu_ode_regression.jl:14-124)
# ---------------------------------------------------------------------------------------
LZ_U0 = np.array([[1.0], [0.0], [0.0]], F32)
LZ_P = np.array([[10.0], [28.0], [F32(8.0) / F32(3.0)]], F32)


@pytest.fixture(scope="module")
def lorenz_vern9():
    out, st, _ = pyoracle.vern9("lorenz", LZ_U0, LZ_P, (0.0, 10.0), saveat=[2.0, 4.0, 10.0])
    dense, _, _ = pyoracle.vern9("lorenz", LZ_U0, LZ_P, (0.0, 10.0), saveat=np.arange(101) * 0.1)
    return out[:, :, 0], dense[:, :, 0]


@pytest.mark.parametrize("method", ["tsit5", "vern7"])
def test_lorenz_regression_tolerances(method, lorenz_vern9):
    ref, dense = lorenz_vern9
    # fixed dt = 0.01, save_everystep: norm(bench.u[end] - sol.u[end]) < 5e-3
    r = pyoracle.solve("lorenz", method, False, LZ_U0, LZ_P, (0.0, 10.0), 0.01, nsave=1001, save_everystep=True)
    assert np.linalg.norm(r["us"][-1, :, 0] - ref[2]) < 5e-3
    # adaptive 1e-7, endpoint: < 1e-3
    r = pyoracle.solve("lorenz", method, True, LZ_U0, LZ_P, (0.0, 10.0), 0.01, abstol=1e-7, reltol=1e-7)
    assert np.linalg.norm(r["us"][1, :, 0] - ref[2]) < 1e-3
    assert r["retcode"][0] == 0
    if method == "tsit5":
        # saveat [2, 4]: fixed and adaptive within 5e-4 of the reference solution
        sv = np.array([2.0, 4.0], F32)
        r = pyoracle.solve("lorenz", method, False, LZ_U0, LZ_P, (0.0, 10.0), 0.01, saveat=sv)
        assert np.linalg.norm(r["us"][:, :, 0] - ref[:2]) < 5e-4
        r = pyoracle.solve("lorenz", method, True, LZ_U0, LZ_P, (0.0, 10.0), 0.01, abstol=1e-7, reltol=1e-7,
                           saveat=sv)
        assert np.linalg.norm(r["us"][:, :, 0] - ref[:2]) < 5e-4
        np.testing.assert_array_equal(r["ts"][:, 0], sv)
        # dense saveat 0:0.1:10: the reference asserts fixed < 5e-3 and adaptive < 1e-2 against a
        # Float32 Vern9 run. Against the exact (FP64) solution, Float32 roundoff amplified by the
        # Lorenz Lyapunov exponent (~0.9) dominates after t ~ 5, so the tolerances apply on [0, 5]
        # and the whole grid gets a loose bound.
        sv = np.array([F32(k) / F32(10) for k in range(101)], F32)
        r = pyoracle.solve("lorenz", method, False, LZ_U0, LZ_P, (0.0, 10.0), 0.01, saveat=sv)
        np.testing.assert_array_equal(r["ts"][:, 0], sv)
        assert np.linalg.norm(r["us"][:51, :, 0] - dense[:51]) < 5e-3
        assert np.linalg.norm(r["us"][:, :, 0] - dense) < 5e-2
        r = pyoracle.solve("lorenz", method, True, LZ_U0, LZ_P, (0.0, 10.0), 0.01, abstol=1e-7, reltol=1e-7,
                           saveat=sv)
        assert np.linalg.norm(r["us"][:51, :, 0] - dense[:51]) < 1e-2
        assert np.linalg.norm(r["us"][:, :, 0] - dense) < 1e-1


def test_dense_output_endpoints():
    # saveat at the very first and last step ends: th = 0 gives uprev exactly
    sv = np.array([0.0, 0.1, 0.2], F32)
    r = pyoracle.solve("lorenz", "tsit5", False, LZ_U0, LZ_P, (0.0, 0.2), 0.1, saveat=sv)
    r2 = pyoracle.solve("lorenz", "tsit5", False, LZ_U0, LZ_P, (0.0, 0.2), 0.1, nsave=3, save_everystep=True)
    np.testing.assert_array_equal(r["us"][0], LZ_U0)
    # th = (0.1 - 0.1)/0.1 = 0 on the step that starts at 0.1 is not reached: save at 0.1 happens
    # at the end of step 1 with th = 1 -> equals the step's u up to rounding
    np.testing.assert_allclose(r["us"][1], r2["us"][1], rtol=2e-6)


# ---------------------------------------------------------------------------------------
# controller invariants
# ---------------------------------------------------------------------------------------
def test_controller_counts_and_dtmin():
    u0, p = lorenz_ref_inputs(64, seed=5)
    r = pyoracle.solve("lorenz", "tsit5", True, u0, p, (0.0, 10.0), 0.1, abstol=1e-6, reltol=1e-6)
    assert np.all(r["retcode"] == 0)
    assert np.all(r["naccept"] > 10) and np.all(r["nreject"] >= 0)
    assert np.all(r["ts"][1] == F32(10.0))
    # tolerance far below Float32 resolution drives dt under 1e-14: DtLessThanMin
    r = pyoracle.solve("decay", "tsit5", True, np.array([[1.0]], F32), np.array([[1.0]], F32), (0.0, 1.0), 0.1,
                       abstol=1e-30, reltol=1e-30)
    assert r["retcode"][0] == 1
    # attempt cap
    r = pyoracle.solve("lorenz", "tsit5", True, u0[:, :2], p[:, :2], (0.0, 10.0), 0.1, abstol=1e-6, reltol=1e-6,
                       max_attempts=5)
    assert np.all(r["retcode"] == 2) and np.all(r["naccept"] + r["nreject"] == 5)


def test_kahan_time_accumulation():
    """1e5 fixed steps of Float32(1e-4) (stopped by the attempt cap): plain Float32 summation
    of t drifts, the compensated sum stays within an ulp of the exact total."""
    dt = F32(1e-4)
    n = 100000
    one = np.array([[1.0]], F32)
    r0 = pyoracle.solve("growth", "tsit5", False, one, one, (0.0, 100.0), dt, nsave=2, max_attempts=n)
    rk = pyoracle.solve("growth", "tsit5", False, one, one, (0.0, 100.0), dt, nsave=2, kahan=True, max_attempts=n)
    assert r0["naccept"][0] == n and rk["naccept"][0] == n
    assert r0["retcode"][0] == 2 and rk["retcode"][0] == 2
    exact = float(dt) * n
    err_plain = abs(float(r0["ts"][1, 0]) - exact)
    err_kahan = abs(float(rk["ts"][1, 0]) - exact)
    assert err_kahan <= float(np.spacing(F32(exact)))
    assert err_plain > 100 * err_kahan + 1e-5


def test_no_progress_rule_ends_every_lane():
    """Lanes whose remaining span falls below the resolution of t (the reference repeats such a
    step forever) snap to tf: the uneven Lorenz workload finishes every trajectory at its tf."""
    import __graft_entry__ as g

    g.load_package()
    from diffeqgpu_amd import workloads

    cfg = workloads.config("lorenz_uneven", N=4000, seed=5)
    o = pyoracle.solve("lorenz", "vern7", True, cfg["u0"], cfg["p"], (0.0, 50.0), 0.01, tfs=cfg["tfs"],
                       abstol=1e-6, reltol=1e-6, detpow=True, max_attempts=100000)
    assert np.all(o["retcode"] == 0)
    np.testing.assert_array_equal(o["ts"][1], cfg["tfs"])
    assert (o["naccept"] + o["nreject"]).max() < 10000
    # a fixed step below the resolution of t stops with DtLessThanMin instead of looping
    r = pyoracle.solve("growth", "tsit5", False, np.array([[1.0]], F32), np.array([[1.0]], F32), (1e6, 1e6 + 1.0),
                       1e-3)
    assert r["retcode"][0] == 1 and r["naccept"][0] == 0


# ---------------------------------------------------------------------------------------
# the workload fixtures (tests/golden/) against the Float32 restatement at the workloads'
# own tolerances: pins the fixtures and the oracle together at benchmark scale
# ---------------------------------------------------------------------------------------
def _workload(name, N=1000):
    from diffeqgpu_amd import workloads

    return workloads.config(name, N=N)


def _golden(name):
    import os

    return np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", name))


@pytest.mark.parametrize("method", ["rb23", "rodas4", "rodas5p", "kvaerno3", "kvaerno5"])
def test_robertson_fixture_vs_oracle(method):
    """Every method at the stiff workload's tolerances against the FP64 Radau fixture. Kvaerno5:
    in 3 of the 1,000 trajectories the last large steps (dt ~ 1e3 near t = 1e5) take Newton
    stage solves that stop at 30 iterations with Float32 residuals above the absolute tolerance
    100 eps (the reference's nlsolve has no failure exit, nlsolve/utils.jl:16-25), and the end
    value is off by up to 0.2; all others are within the bar."""
    cfg, gold = _workload("robertson"), _golden("robertson_radau.npz")
    r = pyoracle.solve("robertson", method, True, cfg["u0"], cfg["p"], cfg["tspan"], cfg["dt"], abstol=1e-8,
                       reltol=1e-4, saveat=np.asarray(cfg["saveat"], F32))
    assert np.all(r["retcode"] == 0)
    err = np.abs(r["us"].astype(np.float64) - gold["u"]).max(axis=(0, 1))
    if method == "kvaerno5":
        assert np.sum(err >= 1e-4) <= 3 and np.percentile(err, 99) < 5e-5, np.sort(err)[-4:]
    else:
        assert err.max() < 1e-4, err.max()


@pytest.mark.parametrize("method", ["rb23", "rodas4", "rodas5p", "kvaerno3", "kvaerno5"])
def test_cramer_solve_matches_pivoted_lu(method):
    """For n <= 3 the W solve is Cramer's rule (the reference's formulas, linalg/linsolve.jl:8-57) in
    FP64; the oracle's pivoted LU (flag pivlu), which the convergence orders and the Radau fixture pin
    for every n, solves the same FP64 systems. The two agree to FP64 rounding, which almost never moves
    a Float32 result: the solutions must be bit-identical in at least 99 % of the values and within
    1e-5 relative everywhere, fixed and adaptive, for n = 3 (Robertson), 2 (Van der Pol) and 1 (decay).
    A transposed cofactor (W[0][1] given W[1][0]'s formula) fails the check for every method: on
    Robertson 17-99 % of the values stay equal and the rest differ by up to 5e-4 (checked by planting it)."""
    r = np.random.default_rng(0)
    N = 64
    cases = [
        ("robertson", np.tile(np.array([[1.0], [0.0], [0.0]], F32), (1, N)),
         (np.array([[0.04], [3e7], [1e4]]) * np.exp(0.3 * r.normal(size=(3, N)))).astype(F32), 1e-8, 1e-4),
        ("vdp", r.uniform(-2, 2, (2, N)).astype(F32), np.exp(r.uniform(np.log(0.1), np.log(30), (1, N))).astype(F32),
         1e-6, 1e-6),
        ("decay", r.uniform(0.5, 2, (1, N)).astype(F32), r.uniform(0.1, 5, (1, N)).astype(F32), 1e-6, 1e-6),
    ]
    for system, u0, p, atol, rtol in cases:
        for adaptive, tspan, dt in ((False, (0.0, 1.0), 1e-3), (True, (0.0, 5.0), 1e-3)):
            kw = dict(abstol=atol, reltol=rtol, saveat=np.linspace(0, tspan[1], 6).astype(F32))
            a = pyoracle.solve(system, method, adaptive, u0, p, tspan, dt, **kw)
            b = pyoracle.solve(system, method, adaptive, u0, p, tspan, dt, pivlu=True, **kw)
            assert np.all(a["retcode"] == 0) and np.all(b["retcode"] == 0)
            ua, ub = a["us"].astype(np.float64), b["us"].astype(np.float64)
            assert np.mean(ua == ub) >= 0.99, (system, adaptive, np.mean(ua == ub))
            rel = np.abs(ua - ub) / np.maximum(np.abs(ub), 1e-6)
            assert rel.max() <= 1e-5, (system, adaptive, rel.max())


def test_vdp_fixture_vs_oracle():
    cfg, gold = _workload("vdp"), _golden("vdp_vern9.npz")
    r = pyoracle.solve("vdp", "tsit5", True, cfg["u0"], cfg["p"], cfg["tspan"], cfg["dt"], abstol=1e-6, reltol=1e-6)
    assert np.all(r["retcode"] == 0)
    err = np.abs(r["us"][-1].astype(np.float64) - gold["u"][0]).max(axis=0)
    assert np.median(err) < 5e-5 and np.percentile(err, 99) < 2e-3, (np.median(err), np.percentile(err, 99))


def test_lorenz_uneven_fixture_vs_oracle():
    # the non-chaotic part (rho < 10: every trajectory settles on a fixed point) must agree
    # closely; chaotic trajectories only in distribution
    cfg, gold = _workload("lorenz_uneven"), _golden("lorenz_uneven_vern9.npz")
    r = pyoracle.solve("lorenz", "vern7", True, cfg["u0"], cfg["p"], cfg["tspan"], cfg["dt"], abstol=1e-8,
                       reltol=1e-8, tfs=cfg["tfs"], kahan=True)
    assert np.all(r["retcode"] == 0)
    err = np.abs(r["us"][-1].astype(np.float64) - gold["u"][0]).max(axis=0)
    calm = cfg["p"][1] < 10
    assert calm.sum() > 100 and err[calm].max() < 1e-4, err[calm].max()
    assert np.median(err) < 0.05
