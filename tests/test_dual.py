"""Dual-number (forward-mode) inputs: the host API, the JIT, and the oracle's dual Lorenz pinned by
what fixes it.

  values          fixed step: the value part of a dual solve is bit-identical to the plain solve
                  (each dual operation computes its value exactly as the plain operation does)
  partials        d u(t) / d (u0, p) at t = 1 against central differences of the FP64 Vern9 flow
                  (rtol 1e-12) at the solve's end time, fixed and adaptive, for Tsit5, Vern7 and Vern9; a sign or index slip
                  in a partial rule is off by O(1)
  implicit        the five implicit methods (autodiff = false): fixed-step values bit-identical to the
                  plain solve; partials against the flow derivative as above; the dual W solve against
                  the same method on the augmented system with its full Jacobian (which pins the
                  partials of W: leaving them out misses by 10x the bar)
  dual norm       with zero partials the dual error norm is the plain one divided by sqrt(1 + P);
                  with abstol = 0 the plain norm scales as 1 / reltol, so an adaptive dual solve
                  with zero seeds at reltol r takes the steps of the plain solve at r sqrt(1 + P)
                  (up to the rounding of that product); the plain norm (values alone) at r takes
                  more than 10 % more steps
GPU parity with the oracle is tested in test_gpu_dual.py.
"""
import numpy as np
import pytest

import pyoracle as P

F32 = np.float32
NP = 6  # partials: d/du0 (3) and d/dp (3)


def _seeded(N, rng=None):
    """Augmented (21, N) u0 and p of the dual Lorenz with the identity seeds."""
    rng = rng or np.random.default_rng(5)
    u0v = np.tile(np.array([[1.0], [0.0], [0.0]], F32), (1, N))
    pv = (np.array([[10.0], [28.0], [8.0 / 3.0]], F32) * rng.uniform(0.3, 1.0, (1, N)).astype(F32)).astype(F32)
    seeds = np.eye(NP, dtype=F32)
    u0 = np.concatenate([u0v, np.repeat(seeds[:3].reshape(-1, 1), N, axis=1)]).astype(F32)
    p = np.concatenate([pv, np.repeat(seeds[3:].reshape(-1, 1), N, axis=1)]).astype(F32)
    return u0, p, u0v, pv


def test_dual_api(deg):
    D = deg.Dual
    u0 = [D(1.0, (1, 0, 0, 0, 0, 0)), D(0.0, (0, 1, 0, 0, 0, 0)), D(0.0, (0, 0, 1, 0, 0, 0))]
    p = [D(10.0, (0, 0, 0, 1, 0, 0)), D(28.0, (0, 0, 0, 0, 1, 0)), D(8 / 3, (0, 0, 0, 0, 0, 1))]
    from diffeqgpu_amd import systems

    prob = deg.ODEProblem(systems.lorenz, u0, (0.0, 10.0), p)
    assert prob.is_dual and prob.u0.npartials == 6 and prob.u0[1] == u0[1]
    cp = deg.make_prob_compatible(prob)
    assert cp.dual == (3, 3, 6) and cp.u0.shape == (21,) and cp.p.shape == (21,)
    assert cp.u0[3] == 1.0 and cp.u0[3 + 7] == 1.0 and cp.u0[3 + 14] == 1.0 and cp.u0[3:].sum() == 3.0
    assert cp.p[2] == F32(8 / 3) and cp.p[3 + 6 * 0 + 3] == 1.0 and cp.p[3 + 6 * 2 + 5] == 1.0
    # plain numbers mixed with duals get zero partials; remake keeps the duals
    q = deg.remake(prob, p=[10.0, 28.0, D(2.0, (0, 0, 0, 0, 0, 1))])
    assert q.is_dual and np.all(q.p.partials[:2] == 0) and q.p.partials[2, 5] == 1.0
    with pytest.raises(ValueError):
        deg.ODEProblem(systems.lorenz, [D(1.0, (1,)), D(0.0, (0, 1)), 0.0], (0.0, 1.0), [1.0, 2.0, 3.0])
    # saved rows as (n, 1 + P)
    from diffeqgpu_amd.dual import split_rows

    rows = split_rows(np.arange(2 * 21, dtype=F32).reshape(2, 21), 3, 6)
    assert rows.shape == (2, 3, 7) and rows[1, 2, 0] == 21 + 2 and rows[1, 2, 1] == 21 + 3 + 12
    # callbacks are refused before any device work
    ep = deg.EnsembleProblem(prob)
    cb = deg.GPUDiscreteCallback(lambda u, t, integ: t == 1.0, lambda integ: None)
    with pytest.raises(NotImplementedError, match="dual"):
        deg.solve(ep, deg.GPUTsit5(), deg.EnsembleGPUKernel(), trajectories=2, dt=0.01, adaptive=False,
                  callback=cb, merge_callbacks=True)


def test_dual_source_and_jit(deg):
    from diffeqgpu_amd import _lib, systems

    f, n, np_, _ = systems.SYSTEMS["lorenz_dual"]
    src, flops = f.hip_source(n, np_, False)
    assert "#define DEG_DUAL_NV 3" in src and "#define DEG_DUAL_P 6" in src and flops == 8
    for m in (_lib.DEG_TSIT5, _lib.DEG_VERN7):
        assert len(_lib.precompile(src, n, np_, m, _lib.F_PARITY)) == 16
    # the implicit methods: the dual Jacobian over the 3 value columns of the 21 augmented rows
    fd_src, _ = f.hip_source(n, np_, True, fd=True)
    assert "for (int i = 0; i < 3; ++i)" in fd_src and "J[r * 3 + i]" in fd_src and "deg_tgrad" in fd_src
    ad_src, _ = f.hip_source(n, np_, True, fd=False)
    assert "J[62] = " in ad_src and "J[63] = " not in ad_src and "dT[20] = " in ad_src
    for m in (_lib.DEG_RB23, _lib.DEG_KVAERNO5):
        assert len(_lib.precompile(fd_src, n, np_, m, _lib.F_PARITY)) == 16
    with pytest.raises(_lib.DegError, match="mass"):
        _lib.precompile("#define DEG_MASS 1\n" + fd_src, n, np_, _lib.DEG_RB23, _lib.F_PARITY)
    with pytest.raises(_lib.DegError, match="ODE methods|noise"):
        _lib.precompile(src, n, np_, _lib.DEG_EM, _lib.F_PARITY)


@pytest.mark.parametrize("method", ["tsit5", "vern7", "vern9"])
def test_dual_fixed_values_equal_plain(oracle, method):
    N = 64
    u0, p, u0v, pv = _seeded(N)
    d = P.solve("lorenz_dual", method, False, u0, p, (0.0, 2.0), 0.01, save_everystep=False)
    o = P.solve("lorenz", method, False, u0v, pv, (0.0, 2.0), 0.01, save_everystep=False)
    np.testing.assert_array_equal(d["us"][:, :3], o["us"])
    np.testing.assert_array_equal(d["ts"], o["ts"])


def _fd_jacobian(u0v, pv, T, h=1e-3):
    """d u(T) / d (u0, p) of the exact flow by central differences of FP64 Vern9 (rtol 1e-12). The
    oracle takes Float32 inputs, so the step is h (1 + |x|) and the quotient uses the difference of
    the rounded inputs; the truncation error is O(h^2)."""
    N = u0v.shape[1]
    base = np.concatenate([u0v, pv]).astype(F32)
    J = np.zeros((3, NP, N))
    for j in range(NP):
        up, dn = base.copy(), base.copy()
        step = (h * (1.0 + np.abs(base[j].astype(np.float64))))
        up[j] = (base[j] + step).astype(F32)
        dn[j] = (base[j] - step).astype(F32)
        ru, _, _ = P.vern9("lorenz", up[:3], up[3:], (0.0, T), saveat=np.array([T]))
        rd, _, _ = P.vern9("lorenz", dn[:3], dn[3:], (0.0, T), saveat=np.array([T]))
        J[:, j, :] = (ru[-1] - rd[-1]) / (up[j].astype(np.float64) - dn[j].astype(np.float64))
    return J


@pytest.mark.parametrize("method", ["tsit5", "vern7", "vern9"])
@pytest.mark.parametrize("adaptive", [False, True])
def test_dual_partials_against_flow_derivative(oracle, method, adaptive):
    N = 16
    T = 1.0
    u0, p, u0v, pv = _seeded(N)
    kw = dict(abstol=1e-7, reltol=1e-7) if adaptive else {}
    d = P.solve("lorenz_dual", method, adaptive, u0, p, (0.0, T), 0.001, save_everystep=False, **kw)
    assert np.all(d["retcode"] == 0)
    parts = d["us"][-1, 3:].astype(np.float64).reshape(3, NP, N)
    # a fixed-step solve ends where the Float32 time accumulates to (1000 steps of 0.001f0: 1.00099)
    te = float(d["ts"][-1, 0])
    assert np.all(d["ts"][-1] == d["ts"][-1, 0]) and (adaptive or te > T)
    J = _fd_jacobian(u0v, pv, te, h=1e-5)
    scale = np.abs(J).max(axis=(0, 1))
    err = np.abs(parts - J).max(axis=(0, 1)) / scale
    assert err.max() < 1e-3, err.max()


IMPLICIT = ["rb23", "rodas4", "rodas5p", "kvaerno3", "kvaerno5"]


@pytest.mark.parametrize("method", IMPLICIT)
def test_dual_implicit_values_equal_plain(oracle, method):
    """The implicit methods with autodiff = false (the reference's forward_diff.jl settings): the W
    factor of a dual state is built from the Jacobian's values alone, so the value part of a fixed-step
    dual solve is bit-identical to the plain solve."""
    N = 32
    u0, p, u0v, pv = _seeded(N)
    d = P.solve("lorenz_dual", method, False, u0, p, (0.0, 1.0), 0.005, save_everystep=False, fdjac=True)
    o = P.solve("lorenz", method, False, u0v, pv, (0.0, 1.0), 0.005, save_everystep=False, fdjac=True)
    np.testing.assert_array_equal(d["us"][:, :3], o["us"])
    np.testing.assert_array_equal(d["ts"], o["ts"])


@pytest.mark.parametrize("method", IMPLICIT)
@pytest.mark.parametrize("adaptive", [False, True])
def test_dual_implicit_partials_against_flow_derivative(oracle, method, adaptive):
    N = 16
    T = 1.0
    u0, p, u0v, pv = _seeded(N)
    kw = dict(abstol=1e-7, reltol=1e-7) if adaptive else {}
    d = P.solve("lorenz_dual", method, adaptive, u0, p, (0.0, T), 0.001, save_everystep=False, fdjac=True, **kw)
    assert np.all(d["retcode"] == 0)
    parts = d["us"][-1, 3:].astype(np.float64).reshape(3, NP, N)
    te = float(d["ts"][-1, 0])
    J = _fd_jacobian(u0v, pv, te, h=1e-5)
    err = np.abs(parts - J).max(axis=(0, 1)) / np.abs(J).max(axis=(0, 1))
    assert err.max() < 1e-3, err.max()


# partials of the dual solve against the same method on the augmented system with its full Jacobian;
# dropping the partials of W from the dual solve (x_j = W_v^-1 b_j) misses by 10x these bars or more
W_BARS = {"rb23": 1e-3, "rodas4": 1e-3, "rodas5p": 2e-4, "kvaerno3": 2e-4, "kvaerno5": 1e-4}


@pytest.mark.parametrize("method", IMPLICIT)
def test_dual_implicit_w_partials_against_augmented_solve(oracle, method):
    """The dual W solve x_j = W_v^-1 (b_j - beta J_j x_v) against the method applied to the 21-entry
    augmented system as a plain system (its forward-difference Jacobian holds the partials of J in the
    lower-left block and J again on the block diagonal; one 21 x 21 FP64 LU). Both differentiate the
    discrete step, so at dt = 0.02 they agree to the forward-difference noise of the augmented
    Jacobian's diagonal blocks; leaving out the W partials differs at first order in dt."""
    N = 16
    u0, p, _, _ = _seeded(N)
    d = P.solve("lorenz_dual", method, False, u0, p, (0.0, 0.5), 0.02, save_everystep=False, fdjac=True)
    g = P.solve("lorenz_dual", method, False, u0, p, (0.0, 0.5), 0.02, save_everystep=False, fdjac=True,
                augmented=True)
    pd = d["us"][-1, 3:].reshape(3, NP, N).astype(np.float64)
    pg = g["us"][-1, 3:].reshape(3, NP, N).astype(np.float64)
    rel = (np.abs(pd - pg).max(axis=(0, 1)) / np.abs(pg).max(axis=(0, 1))).max()
    assert rel < W_BARS[method], rel
    # the values: identical for the Rosenbrock methods; Kvaerno's Newton iterations stop on another norm
    # (the values alone against every entry) and so a few iterations apart
    np.testing.assert_allclose(d["us"][-1, :3], g["us"][-1, :3], rtol=1e-4)


def test_dual_norm_zero_seeds(oracle):
    """Zero partials: the dual norm is the plain norm / sqrt(1 + P). With abstol = 0 the scaled error is
    e / (|u| reltol), so the dual solve at reltol r takes the steps of the plain solve at reltol
    r sqrt(1 + P), up to the rounding of that product (step counts within 5 %, 1 % in total)."""
    N = 64
    rng = np.random.default_rng(9)
    u0v = (np.array([[1.0], [0.5], [2.0]], F32) + rng.uniform(0, 1, (3, N)).astype(F32)).astype(F32)
    pv = (np.array([[10.0], [28.0], [8.0 / 3.0]], F32) * rng.uniform(0.3, 0.5, (1, N)).astype(F32)).astype(F32)
    u0 = np.concatenate([u0v, np.zeros((18, N), F32)])
    p = np.concatenate([pv, np.zeros((18, N), F32)])
    r = 1e-6
    d = P.solve("lorenz_dual", "tsit5", True, u0, p, (0.0, 1.0), 0.01, abstol=0.0, reltol=r, save_everystep=False)
    o = P.solve("lorenz", "tsit5", True, u0v, pv, (0.0, 1.0), 0.01, abstol=0.0, reltol=float(F32(r * np.sqrt(7.0))),
                save_everystep=False)
    na_d, na_o = d["naccept"] + d["nreject"], o["naccept"] + o["nreject"]
    assert np.all(d["us"][-1, 3:] == 0.0)
    assert np.abs(na_d.astype(float) / na_o - 1).max() < 0.05, (na_d[:8], na_o[:8])
    assert np.abs(na_d.sum() / na_o.sum() - 1) < 0.01
    np.testing.assert_allclose(d["us"][-1, :3], o["us"][-1], rtol=1e-4, atol=1e-5)
    # and not the plain solve at the unscaled tolerance
    o2 = P.solve("lorenz", "tsit5", True, u0v, pv, (0.0, 1.0), 0.01, abstol=0.0, reltol=r, save_everystep=False)
    assert 1.1 * na_d.sum() < (o2["naccept"] + o2["nreject"]).sum()


def _elementary(u, p, t):
    """Every dual rule: division (dual / dual, real / dual, dual / real), powers (integer, dual base,
    dual exponent, both), sin, cos, exp, log, sqrt, tanh, abs, ifelse, and t."""
    from diffeqgpu_amd import math as m
    from diffeqgpu_amd.tracer import SVector

    a = m.sin(u[0]) * m.exp(p[0] * u[1]) / (1.0 + u[2] ** 2) + 2.0 / (1.0 + p[0] * p[0]) - u[1] / 3.0
    b = m.log(2.0 + u[0] * u[0]) - m.sqrt(1.0 + p[1] * u[1] ** 2) + m.tanh(u[0] - p[0]) + m.cos(t * u[2])
    c = (m.pow(1.5 + u[1] ** 2, p[1]) + m.pow(u[2] * u[2] + 1.0, 0.7) + m.pow(2.0, u[1])
         + m.fabs(u[0] - 0.3) * u[2] + m.ifelse(u[0], ">", 0.5, u[1] * u[2], -u[2]) - u[0] ** 3)
    return SVector(a, b, c)


@pytest.mark.parametrize("point", [0, 1, 2])
def test_dual_rules_against_finite_differences(point):
    """The traced dual evaluation (interpreted in float64) against central differences of the plain
    function along each seed: a wrong derivative factor or sign in any rule fails."""
    from diffeqgpu_amd import tracer
    from test_host import _eval_nodes

    rng = np.random.default_rng(point)
    u = rng.uniform(-1.0, 1.5, 3)
    p = rng.uniform(0.3, 1.2, 2)
    t = 0.4 + 0.3 * point
    n, npp, P = 3, 2, 5
    tr = tracer.trace_dual(_elementary, n, npp, P)
    seeds = np.eye(P)
    ua = np.concatenate([u, seeds[:3].reshape(-1)])
    pa = np.concatenate([p, seeds[3:].reshape(-1)])
    out = np.array(_eval_nodes(tr, tr.outs, ua, pa, t))
    vals, parts = out[:n], out[n:].reshape(n, P)
    np.testing.assert_allclose(vals, tracer.evaluate(_elementary, u, p, t), rtol=1e-12)
    x0 = np.concatenate([u, p])
    h = 1e-6
    for j in range(P):
        xp, xm = x0.copy(), x0.copy()
        xp[j] += h
        xm[j] -= h
        fd = (np.array(tracer.evaluate(_elementary, xp[:3], xp[3:], t))
              - np.array(tracer.evaluate(_elementary, xm[:3], xm[3:], t))) / (2 * h)
        np.testing.assert_allclose(parts[:, j], fd, rtol=1e-6, atol=1e-7)


def _hill(u, p, t):
    """A Hill feedback term and a power with a dual exponent: u' = -u + u^p / (1 + u^p)."""
    from diffeqgpu_amd import math as m
    from diffeqgpu_amd.tracer import SVector

    h = m.pow(u[0], p[0])
    return SVector(-u[0] + h / (1.0 + h))


@pytest.mark.parametrize("uv,pv,seed,want", [
    (0.0, 2.0, (1.0, 0.0), -1.0),        # zero base, exponent with zero partials: d/du = -1 (not NaN)
    (0.0, 2.0, (0.0, 1.0), 0.0),         # zero base, positive dual exponent: the log term is 0
    (-1.5, 3.0, (1.0, 0.0), None),       # negative base, integer-valued exponent with zero partials
])
def test_dual_pow_guards(uv, pv, seed, want):
    """ForwardDiff's guards of the power rule: where x = 0 (exponent > 0) or where the exponent's
    partials are all zero, the log(x) term must not turn the partials into NaN. Checked against
    the derivative by hand and against a central difference of the plain function."""
    from diffeqgpu_amd import tracer
    from test_host import _eval_nodes

    tr = tracer.trace_dual(_hill, 1, 1, 1)
    out = _eval_nodes(tr, tr.outs, [uv, seed[0]], [pv, seed[1]], 0.0)
    assert all(np.isfinite(out)), out
    x0 = np.array([uv, pv])
    h = 1e-6
    d = np.array(seed)
    fd = (tracer.evaluate(_hill, [uv + h * d[0]], [pv + h * d[1]], 0.0)[0]
          - tracer.evaluate(_hill, [uv - h * d[0]], [pv - h * d[1]], 0.0)[0]) / (2 * h)
    np.testing.assert_allclose(out[1], fd, rtol=1e-6, atol=1e-7)
    if want is not None:
        assert out[1] == want
    else:
        # d/du (u^3 / (1 + u^3)) - 1 at u = -1.5: 3u^2 / (1 + u^3)^2 - 1
        assert abs(out[1] - (3 * uv ** 2 / (1 + uv ** 3) ** 2 - 1.0)) < 1e-12
    del x0


def test_dual_abs_and_max_tie_rules(oracle):
    """The dual error norm's abs and max follow DiffRules: abs'(x) = signbit(x) ? -1 : 1 and
    max(x, y) takes y's partials unless x > y. A state with one constant component ties
    |uprev| = |u| on every step: the norm then takes u's partials. Checked with the oracle's norm
    on hand-made states (values and partials chosen so that each rule changes the result)."""
    nv, P_, n = 1, 1, 2
    e = np.array([1e-3, 2e-3], F32)
    # tie |uprev| == |u| = 1 with different partials: the rule takes u's partial (0.5), not uprev's (4)
    uprev = np.array([1.0, 4.0], F32)
    u = np.array([1.0, 0.5], F32)
    got = P.dual_eest(nv, P_, n, e, uprev, u, 0.0, 1.0)

    def norm(md, neg):
        sv = 1.0 * 1.0
        qv = F32(e[0]) / F32(sv)
        sd = F32((-md if neg else md) * 1.0)
        qd = F32(e[1] * (1.0 / sv) + sd * (-(e[0] / (sv * sv))))
        return float(np.sqrt(np.float32(qv * qv + qd * qd) / 2))

    assert abs(got - norm(0.5, False)) < 1e-9 and abs(got - norm(4.0, False)) > 1e-6
    # -0.0 carries the sign bit: its partials are negated (u wins the tie at value 0)
    got0 = P.dual_eest(nv, P_, n, e, np.array([0.0, 0.3], F32), np.array([-0.0, 0.7], F32), 1e-3, 1.0)
    sv = 1e-3
    qv = e[0] / F32(sv)
    sd = F32(-0.7 * 1.0)
    qd = F32(e[1] * F32(1.0 / F32(sv)) + sd * F32(-(e[0] / F32(F32(sv) * F32(sv)))))
    assert abs(got0 - float(np.sqrt(F32(qv * qv + qd * qd) / 2))) < 1e-6 * got0
