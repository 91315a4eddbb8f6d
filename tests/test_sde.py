"""CPU tests of the SDE path (GPUEM, GPUSIEA): the oracle's noise generator and steps against what
the mathematics fixes, and the host API (problem types, output sizing, errors, JIT compile).

Pins of the oracle (oracle/deg_oracle_sde.c):
  Philox4x32-10            the Random123 known-answer vectors
  Box-Muller ln / sincos   float64 math on random and boundary words (relative / absolute error)
  normals                  moments, Kolmogorov-Smirnov, chi-square(2) radius, independence across
                           streams, steps and components; counter layout
  EM, g = 0                explicit Euler in Float32 (bitwise)
  SIEA, g = 0              Heun's method in Float32 (bitwise)
  EM, additive noise       u0 + sum sqrt(dt) sigma w_j in Float32 (bitwise): step and stream indexing
  SIEA, additive noise     reduces to the same Brownian path (to rounding)
  SIEA, GBM                the closed one-step map u R(w), R = 1 + a h + a^2 h^2 / 2 + b (1 + a h) dW
                           + b^2 (dW^2 - h) / 2, evaluated in float64 on the same normals; R's first
                           two moments match exp(a h) and exp((2a + b^2) h) to O(h^3) (weak order 2)
  EM, non-diagonal         u + 1.01 u h + G(u) dW in float64 on the same normals (matrix orientation)
  EM, OU                   discrete mean and variance recursions (Monte Carlo, 5 sigma)
Parity of the HIP kernels with the oracle is tested in test_gpu_sde.py.
"""
import math

import numpy as np
import pytest

import pyoracle as P

KAT = [
    ((0x00000000, 0x00000000, 0x00000000, 0x00000000), (0x00000000, 0x00000000),
     (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff, 0xffffffff, 0xffffffff, 0xffffffff), (0xffffffff, 0xffffffff),
     (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
]


def test_philox_known_answers(oracle):
    for ctr, key, want in KAT:
        assert [int(x) for x in P.philox(ctr, key)] == list(want)


def _words():
    rng = np.random.default_rng(11)
    w = [int(x) for x in rng.integers(0, 2 ** 32, 20000, dtype=np.uint64)]
    # smallest / largest U, powers of two, the sqrt(2) mantissa boundary, quadrant and octant edges
    w += [0, 255, 256, 511, 0xFFFFFFFF, 0xFFFFFF00, 0x7FFFFFFF, 0x80000000, 0x40000000, 0xC0000000]
    w += [((0x3504F3 + d) << 8) & 0xFFFFFFFF for d in (-2, -1, 0, 1, 2)]
    w += [((k << 22) + r) << 8 for k in range(4) for r in (0, 1, 0x1FFFFF, 0x200000, 0x200001, 0x3FFFFF)]
    return w


def test_box_muller_log_accuracy(oracle):
    worst = 0.0
    for w in _words():
        u = ((w >> 8) + 1) * 2.0 ** -24
        ref = -2.0 * math.log(u)
        got = P.bm_m2log(w)
        if ref == 0.0:
            assert got == 0.0
            continue
        worst = max(worst, abs(got - ref) / ref)
    assert worst < 4e-7, worst


def test_box_muller_sincos_accuracy(oracle):
    worst = 0.0
    for w in _words():
        th = 2.0 * math.pi * (w >> 8) * 2.0 ** -24
        c, s = P.bm_sincos(w)
        worst = max(worst, abs(c - math.cos(th)), abs(s - math.sin(th)))
        assert c * c + s * s == pytest.approx(1.0, abs=4e-7)
    assert worst < 2.5e-7, worst


def test_normals_distribution(oracle):
    from scipy import stats

    z = P.normals(123, 0, 500_000, 0, 4).astype(np.float64)
    x = z.ravel()
    N = x.size
    assert abs(x.mean()) < 5 / math.sqrt(N)
    assert abs(x.var() - 1.0) < 5 * math.sqrt(2.0 / N)
    assert abs(stats.skew(x)) < 5 * math.sqrt(6.0 / N)
    assert abs(stats.kurtosis(x)) < 5 * math.sqrt(24.0 / N)
    assert stats.kstest(x, "norm").pvalue > 1e-4
    # a Box-Muller pair has radius^2 ~ chi-square(2): exp(-r^2 / 2) is uniform
    for a in (0, 2):
        r2 = z[a] ** 2 + z[a + 1] ** 2
        assert stats.kstest(np.exp(-r2 / 2.0), "uniform").pvalue > 1e-4
    # the tail is cut at sqrt(-2 ln 2^-24)
    assert np.abs(x).max() <= math.sqrt(-2.0 * math.log(2.0 ** -24)) + 1e-5


def test_normals_independence(oracle):
    n = 200_000
    lim = 5 / math.sqrt(n)
    a = P.normals(7, 0, n, 0, 8).astype(np.float64)
    b = P.normals(7, 0, n, 1, 8).astype(np.float64)
    c = P.normals(7, 1, n, 0, 8).astype(np.float64)  # the neighbouring streams
    d = P.normals(8, 0, n, 0, 8).astype(np.float64)  # another seed
    for x, y in ((a[0], a[1]), (a[0], a[2]), (a[0], a[4]), (a[3], a[7]), (a[0], b[0]), (a[0], c[0]),
                 (a[0], d[0]), (a[5], b[5])):
        assert abs(np.corrcoef(x, y)[0, 1]) < lim
    # the same (seed, stream, step) gives the same numbers
    assert np.array_equal(a, P.normals(7, 0, n, 0, 8))
    # a stream's numbers do not depend on where a call starts
    assert np.array_equal(a[:, 1000:1010], P.normals(7, 1000, 10, 0, 8))


def test_normals_counter_layout(oracle):
    """normal c of step s, stream id = word pair (c % 4) / 2 of Philox(key = seed, counter =
    (s, c / 4, id low, id high)), cosine branch for even c."""
    seed = 0x0123456789ABCDEF
    for step, id0, m in ((0, 0, 5), (77, 2 ** 32 + 5, 8), (2 ** 31, 123456789, 3)):
        z = P.normals(seed, id0, 3, step, m)
        for i in range(3):
            sid = id0 + i
            for c in range(m):
                w = P.philox((step, c // 4, sid & 0xFFFFFFFF, sid >> 32), (seed & 0xFFFFFFFF, seed >> 32))
                j = (c % 4) // 2
                rr = np.sqrt(np.float32(P.bm_m2log(int(w[2 * j]))))
                co, si = P.bm_sincos(int(w[2 * j + 1]))
                want = np.float32(rr * np.float32(co if c % 2 == 0 else si))
                assert z[c, i] == want, (step, id0, i, c)


def _nsteps(t0, tf, dt):
    return int(np.floor(np.float32(np.float32(abs(np.float32(tf) - np.float32(t0))) / np.float32(abs(dt)))))


@pytest.mark.parametrize("dt", [0.1, 0.3, 1.0 / 64])
def test_em_without_noise_is_euler(oracle, dt):
    N = 4
    u0 = np.array([[0.5, -1.0, 2.0, 0.1]], np.float32)
    p = np.array([[1.5, -0.7, 0.3, 2.0], [0.0] * 4], np.float32)
    n = _nsteps(0.0, 1.0, dt)
    r = P.sde_solve("gbm", "em", u0, p, (0.0, 1.0), dt, seed=5, nsave=n + 1)
    d = np.float32(dt)
    for i in range(N):
        u, t = np.float32(u0[0, i]), np.float32(0.0)
        a = p[0, i]
        for j in range(n):
            u = np.float32(u + np.float32(np.float32(a * u) * d))
            t = np.float32(t + d)
            assert r["us"][j + 1, 0, i] == u and r["ts"][j + 1, i] == t
    assert (r["naccept"] == n).all() and (r["retcode"] == 0).all()


def test_siea_without_noise_is_heun(oracle):
    N = 3
    u0 = np.array([[0.5, -1.0, 2.0]], np.float32)
    p = np.array([[1.5, -0.7, 0.3], [0.0] * N], np.float32)
    dt = np.float32(1.0 / 32)
    r = P.sde_solve("gbm", "siea", u0, p, (0.0, 1.0), dt, seed=5, nsave=33)
    for i in range(N):
        u, a = np.float32(u0[0, i]), p[0, i]
        for j in range(32):
            k0 = np.float32(a * u)
            k1 = np.float32(a * np.float32(u + np.float32(k0 * dt)))
            u = np.float32(u + np.float32(np.float32(np.float32(0.5 * k0) + np.float32(0.5 * k1)) * dt))
            assert r["us"][j + 1, 0, i] == u


def test_em_additive_noise_path(oracle):
    """f = 0, g = sigma: u_n = u0 + sum_j (sqrt(dt) sigma) w_j with w_j the normals of step j and
    stream traj_base + i."""
    N, dt, base = 6, np.float32(0.05), 1000
    u0 = np.linspace(-1, 1, N, dtype=np.float32)[None, :]
    p = np.linspace(0.5, 2.0, N, dtype=np.float32)[None, :]
    n = _nsteps(0.0, 1.0, dt)
    r = P.sde_solve("bm", "em", u0, p, (0.0, 1.0), dt, seed=99, traj_base=base, nsave=n + 1)
    sq = np.sqrt(dt)
    for i in range(N):
        u = u0[0, i]
        for j in range(n):
            w = P.normals(99, base + i, 1, j, 1)[0, 0]
            u = np.float32(u + np.float32(np.float32(sq * p[0, i]) * w))
            assert r["us"][j + 1, 0, i] == u
    # SIEA on additive noise: g1 = g2 = g0, the extra terms cancel and the path is the same
    s = P.sde_solve("bm", "siea", u0, p, (0.0, 1.0), dt, seed=99, traj_base=base, nsave=n + 1)
    np.testing.assert_allclose(s["us"], r["us"], rtol=0, atol=2e-5 * float(p.max()))


def test_siea_gbm_one_step_map(oracle):
    """SIEA on du = a u dt + b u dW is u_{k+1} = u_k R(w_k) with the weak order-2 Taylor map
    R = 1 + a h + a^2 h^2 / 2 + b (1 + a h) dW + b^2 (dW^2 - h) / 2, dW = sqrt(h) w."""
    N, h, n = 8, np.float32(1.0 / 16), 16
    u0 = np.full((1, N), 0.7, np.float32)
    a = np.linspace(-1.0, 1.5, N).astype(np.float32)
    b = np.linspace(0.1, 0.9, N).astype(np.float32)
    r = P.sde_solve("gbm", "siea", u0, np.stack([a, b]), (0.0, 1.0), h, seed=3, nsave=n + 1)
    e = P.sde_solve("gbm", "em", u0, np.stack([a, b]), (0.0, 1.0), h, seed=3, nsave=n + 1)
    H = float(h)
    sq = float(np.sqrt(h))
    for i in range(N):
        A, B = float(a[i]), float(b[i])
        u = uem = float(u0[0, i])
        for j in range(n):
            w = float(P.normals(3, i, 1, j, 1)[0, 0])
            dW = sq * w
            u *= 1 + A * H + A * A * H * H / 2 + B * (1 + A * H) * dW + B * B * (dW * dW - H) / 2
            uem *= 1 + A * H + B * dW
            assert r["us"][j + 1, 0, i] == pytest.approx(u, rel=1e-5)
            assert e["us"][j + 1, 0, i] == pytest.approx(uem, rel=1e-5)


def _moments(R, h):
    x, wts = np.polynomial.hermite_e.hermegauss(20)
    wts = wts / wts.sum()
    v = R(x, h)
    return float((wts * v).sum()), float((wts * v * v).sum())


@pytest.mark.parametrize("order,R", [
    (1, lambda w, h, a=1.3, b=0.6: 1 + a * h + b * math.sqrt(h) * w),
    (2, lambda w, h, a=1.3, b=0.6: (1 + a * h + a * a * h * h / 2 + b * (1 + a * h) * math.sqrt(h) * w
                                    + b * b * (h * w * w - h) / 2)),
])
def test_one_step_maps_weak_order(order, R):
    """The local errors of E[R] and E[R^2] against exp(a h), exp((2a + b^2) h) are O(h^(order+1))."""
    a, b = 1.3, 0.6
    errs = []
    for h in (0.02, 0.01, 0.005):
        m1, m2 = _moments(R, h)
        errs.append((abs(m1 - math.exp(a * h)), abs(m2 - math.exp((2 * a + b * b) * h))))
    for k in range(2):
        for q in range(2):
            ratio = errs[q][k] / errs[q + 1][k]
            assert 2 ** (order + 1) * 0.8 < ratio < 2 ** (order + 1) * 1.25, (order, k, ratio)


def test_em_nondiagonal_noise(oracle):
    N, h, n = 5, np.float32(1.0 / 64), 64
    u0 = np.ones((2, N), np.float32) * np.linspace(0.5, 1.5, N, dtype=np.float32)
    r = P.sde_solve("nondiag", "em", u0, None, (0.0, 1.0), h, seed=17, nsave=n + 1)
    sq = float(np.sqrt(h))
    for i in range(N):
        u = u0[:, i].astype(np.float64)
        for j in range(n):
            w = P.normals(17, i, 1, j, 4)[:, 0].astype(np.float64)
            G = np.array([[0.3 * u[0], 0.9 * u[0], 1.2 * u[1], 0.3 * u[1]],
                          [0.6 * u[0], 0.12 * u[0], 0.2 * u[1], 1.8 * u[1]]])
            u = u + 1.01 * u * float(h) + G @ (sq * w)
            np.testing.assert_allclose(r["us"][j + 1, :, i], u, rtol=2e-5, atol=1e-6)


def test_em_ou_moments(oracle):
    """EM on du = theta (mu - u) dt + sigma dW: E u_k = mu + (u0 - mu) (1 - theta h)^k and
    V_{k+1} = (1 - theta h)^2 V_k + sigma^2 h."""
    N, h = 200_000, np.float32(0.05)
    th, mu, sg = 2.0, 1.0, 0.8
    u0 = np.full((1, N), 3.0, np.float32)
    p = np.tile(np.array([[th], [mu], [sg]], np.float32), (1, N))
    r = P.sde_solve("ou", "em", u0, p, (0.0, 1.0), h, seed=2024, save_everystep=False)
    n = _nsteps(0.0, 1.0, h)
    q = 1 - th * float(h)
    mean = mu + (3.0 - mu) * q ** n
    V = 0.0
    for _ in range(n):
        V = q * q * V + sg * sg * float(h)
    x = r["us"][1, 0].astype(np.float64)
    assert abs(x.mean() - mean) < 5 * math.sqrt(V / N)
    assert abs(x.var() - V) < 5 * V * math.sqrt(2.0 / N)


def test_oracle_step_count_and_saveat(oracle):
    """n = floor(|tf - t0| / dt) + 1 points; saveat rows are the linear interpolation between the
    steps around each save time (gpu_em_perform_step.jl:43-44, :58-66)."""
    u0 = np.array([[0.5]], np.float32)
    p = np.array([[1.0], [0.5]], np.float32)
    h = np.float32(0.3)
    full = P.sde_solve("gbm", "em", u0, p, (0.0, 1.0), h, seed=1, nsave=4)
    assert np.array_equal(full["ts"][:, 0], np.array([0, 0.3, 0.6, 0.90000004], np.float32))
    sv = np.array([0.0, 0.45, 0.9, 1.0], np.float32)
    s = P.sde_solve("gbm", "em", u0, p, (0.0, 1.0), h, seed=1, saveat=sv)
    th = np.float32(np.float32(0.45 - np.float32(np.float32(0.6) - h)) / h)
    u1, u2 = full["us"][1, 0, 0], full["us"][2, 0, 0]
    assert s["us"][1, 0, 0] == np.float32(u1 + np.float32(np.float32(u2 - u1) * th))
    assert s["us"][0, 0, 0] == np.float32(0.5)
    # 1.0 lies after the last step (t = 0.9): the row keeps the t0 prefill
    assert s["ts"][3, 0] == np.float32(0.0)
    e = P.sde_solve("gbm", "em", u0, p, (0.0, 1.0), h, seed=1, save_everystep=False)
    assert e["ts"][1, 0] == full["ts"][3, 0] and e["us"][1, 0, 0] == full["us"][3, 0, 0]


# ---- host API ------------------------------------------------------------------------------


def test_sde_problem_and_algorithms(deg):
    f, g = (lambda u, p, t: [p[0] * u[0]]), (lambda u, p, t: [p[1] * u[0]])
    prob = deg.SDEProblem(f, g, [0.1], (0.0, 1.0), [1.5, 0.01], seed=1234)
    assert prob.seed == 1234 and deg.is_diagonal_noise(prob) and prob.noise_m is None
    ip = deg.make_prob_compatible(prob)
    assert isinstance(ip, deg.ImmutableSDEProblem) and ip.seed == 1234
    q = deg.remake(prob, seed=7, u0=[0.2])
    assert q.seed == 7 and q.u0[0] == np.float32(0.2) and q.f.g is g
    nd = deg.SDEProblem(f, g, [1.0, 1.0], (0.0, 1.0), noise_rate_prototype=np.zeros((2, 4)))
    assert not deg.is_diagonal_noise(nd) and nd.noise_m == 4
    with pytest.raises(ValueError):
        deg.SDEProblem(f, g, [1.0, 1.0], (0.0, 1.0), noise_rate_prototype=np.zeros((3, 4)))
    assert deg.alg_order(deg.GPUEM()) == 1 and deg.alg_order(deg.GPUSIEA()) == 2
    assert deg.GPUEM() == deg.GPUEM() and deg.GPUEM() != deg.GPUSIEA()


def test_sde_trace_noise_shapes(deg):
    from diffeqgpu_amd import systems

    f, n, np_, m, _ = systems.SDE_SYSTEMS["nondiag"]
    src, _ = f.hip_source(n, np_, m)
    assert "#define DEG_NM 4" in src and "#define DEG_DIAG_NOISE 0" in src and "g[7] =" in src
    f, n, np_, m, _ = systems.SDE_SYSTEMS["lorenz"]
    src, _ = f.hip_source(n, np_, m)
    assert "#define DEG_NM 3" in src and "#define DEG_DIAG_NOISE 1" in src
    from diffeqgpu_amd import tracer

    with pytest.raises(tracer.TraceError):
        tracer.emit_noise(lambda u, p, t: [u[0]], lambda u, p, t: [u[0], u[0]], 1, 0)
    with pytest.raises(tracer.TraceError):
        tracer.emit_noise(lambda u, p, t: [u[0]], lambda u, p, t: [[u[0], u[0]]], 1, 0, 3)


def test_sde_jit_compile_all(deg):
    """Both SDE kernels compile for every test system (hiprtc, no GPU needed); SIEA refuses
    general noise and the SDE kernels refuse adaptive and ODE-only flags."""
    from diffeqgpu_amd import _lib, systems

    for name, (f, n, np_, m, _) in systems.SDE_SYSTEMS.items():
        src, _ = f.hip_source(n, np_, m)
        for meth in (_lib.DEG_EM, _lib.DEG_SIEA):
            if meth == _lib.DEG_SIEA and m is not None:
                with pytest.raises(_lib.DegError, match="noise type"):
                    _lib.precompile(src, n, np_, meth, _lib.F_PARITY)
                continue
            assert len(_lib.precompile(src, n, np_, meth, _lib.F_PARITY)) == 16
    f, n, np_, m, _ = systems.SDE_SYSTEMS["gbm"]
    src, _ = f.hip_source(n, np_, m)
    assert len(_lib.precompile(src, n, np_, _lib.DEG_EM, 0)) == 16  # contraction allowed
    with pytest.raises(_lib.DegError, match="Adaptive"):
        _lib.precompile(src, n, np_, _lib.DEG_EM, _lib.F_ADAPTIVE)
    with pytest.raises(_lib.DegError, match="SDE kernels"):
        _lib.precompile(src, n, np_, _lib.DEG_EM, _lib.F_PARITY | _lib.F_PERSISTENT)
    ode_src, _ = deg.ODEFunction(systems.decay).hip_source(1, 1, False)
    with pytest.raises(_lib.DegError, match="deg_noise"):
        _lib.precompile(ode_src, 1, 1, _lib.DEG_EM, _lib.F_PARITY)


def test_sde_host_errors(deg):
    from diffeqgpu_amd import systems

    f, n, np_, m, _ = systems.SDE_SYSTEMS["gbm"]
    prob = deg.SDEProblem(f, None, [0.1], (0.0, 1.0), [1.5, 0.01])
    ep = deg.EnsembleProblem(prob)
    with pytest.raises(NotImplementedError, match="Adaptive"):
        deg.solve(ep, deg.GPUEM(), deg.EnsembleGPUKernel(), trajectories=4, dt=0.1)
    with pytest.raises(NotImplementedError, match="Adaptive"):
        deg.vectorized_asolve(None, prob, deg.GPUEM(), dt=0.1)
    with pytest.raises(TypeError):
        deg.solve(ep, deg.ODEFunction, deg.EnsembleGPUKernel(), trajectories=4, dt=0.1)
    with pytest.raises(ValueError, match="kahan"):
        deg.HIPDevice(kahan=True).sde_flags()
    assert deg.HIPDevice().sde_flags() == 2 and deg.HIPDevice(contract="fast", waves=2).sde_flags() == 2 << 9
