"""Mass matrices (M u' = f) for the implicit methods: the host API, the JIT, and the oracle pinned
by what fixes it.

  explicit identity M    Rodas4, Rodas5P, Kvaerno3, Kvaerno5: bit-identical with no mass matrix
                         (M v = v exactly); Rosenbrock23 takes the reference's other association
                         for its M branch (gpu_rosenbrock23_perform_step.jl:54-58, :140-148) and
                         agrees to rounding
  2 x 2 linear DAE       the reference's fixed-step DAE test (gpu_ode_modelingtoolkit_dae.jl:22-106):
                         u1 + u2 = 1 and u1 -> 1e4 / 10000.04 (closed form) for all five methods
  Robertson DAE          M = diag(1, 1, 0): the reference's test settings for Rosenbrock23
                         (gpu_ode_mass_matrix.jl) against an FP64 Radau solution of the ODE form
                         (the same solution); Rodas4 / Rodas5P on the stiff workload's fixture,
                         and y1 + y2 + y3 = 1 at every save
GPU parity with the oracle is tested in test_gpu_mass.py.
"""
import numpy as np
import pytest

import pyoracle as P

F32 = np.float32
IMPLICIT = ("rb23", "rodas4", "rodas5p", "kvaerno3", "kvaerno5")


def test_mass_matrix_api(deg):
    from diffeqgpu_amd import systems, tracer

    f = deg.ODEFunction(systems.decay)
    assert f.mass_matrix is None
    assert deg.ODEFunction(systems.decay, mass_matrix="I").mass_matrix is None
    g = deg.ODEFunction(systems.decay, mass_matrix=[[2.0]])
    assert g.mass_matrix.dtype == np.float32 and g.mass_matrix[0, 0] == 2.0
    assert deg.ODEFunction(g).mass_matrix[0, 0] == 2.0  # kept when wrapped again
    src, _ = g.hip_source(1, 1, True)
    assert "#define DEG_MASS 1" in src and "0x40000000u" in src
    assert "DEG_MASS" not in f.hip_source(1, 1, True)[0]
    with pytest.raises(ValueError):
        deg.ODEFunction(systems.decay, mass_matrix=[[1.0, 0.0]])
    with pytest.raises(ValueError):
        deg.ODEFunction(systems.decay, mass_matrix="J")
    with pytest.raises(tracer.TraceError):
        deg.ODEFunction(systems.decay, mass_matrix=np.eye(2)).hip_source(1, 1, True)


def test_mass_matrix_jit(deg):
    from diffeqgpu_amd import _lib, systems

    for name in systems.MASS_SYSTEMS:
        f, n, np_, _ = systems.SYSTEMS[name]
        src, _ = f.hip_source(n, np_, True)
        for m in (_lib.DEG_RB23, _lib.DEG_RODAS4, _lib.DEG_RODAS5P, _lib.DEG_KVAERNO3, _lib.DEG_KVAERNO5):
            for flags in (_lib.F_PARITY, _lib.F_ADAPTIVE | _lib.F_PARITY | _lib.F_DETPOW | _lib.F_PERSISTENT):
                assert len(_lib.precompile(src, n, np_, m, flags)) == 16
        for m in (_lib.DEG_TSIT5, _lib.DEG_VERN7, _lib.DEG_VERN9):
            with pytest.raises(_lib.DegError, match="implicit"):
                _lib.precompile(src, n, np_, m, _lib.F_PARITY)


def test_mass_matrix_explicit_method_raises(deg):
    from diffeqgpu_amd import systems

    f = systems.SYSTEMS["dae2"][0]
    prob = deg.ODEProblem(f, [1.0, 0.0], (0.0, 0.1))
    with pytest.raises(ValueError, match="implicit"):
        deg.solve(deg.EnsembleProblem(prob), deg.GPUTsit5(), deg.EnsembleGPUKernel(), trajectories=2, dt=0.001,
                  adaptive=False)


@pytest.mark.parametrize("method", IMPLICIT)
@pytest.mark.parametrize("adaptive", [False, True])
def test_identity_mass_matrix(oracle, method, adaptive):
    g = np.load("tests/golden/robertson_radau.npz")
    N = 24
    u0, p, sv = g["u0"][:, :N].astype(F32), g["p"][:, :N].astype(F32), g["t"].astype(F32)
    if adaptive:
        kw = dict(abstol=1e-8, reltol=1e-4, saveat=sv, detpow=True)
        args = (True, u0, p, (0.0, 1e5), 1e-6)
    else:
        kw = dict(save_everystep=False)
        args = (False, u0, p, (0.0, 1.0), 1e-3)
    a = P.solve("robertson", method, *args, **kw)
    b = P.solve("robertson_eye", method, *args, **kw)
    if method == "rb23":
        np.testing.assert_allclose(b["us"], a["us"], rtol=1e-3, atol=1e-9)
    else:
        for key in a:
            np.testing.assert_array_equal(a[key], b[key])


@pytest.mark.parametrize("method", IMPLICIT)
def test_linear_dae_closed_form(oracle, method):
    """u1' = -0.04 u1 + 1e4 u2, 0 = u1 + u2 - 1: u1 -> 1e4 / 10000.04 with rate 10000.04, so after
    100 steps of 0.001 (lambda dt = 10) only the limit is left."""
    u0 = np.array([[1.0, 1.0], [0.0, 0.0]], F32)
    r = P.solve("dae2", method, False, u0, None, (0.0, 0.1), 0.001, nsave=101, save_everystep=True)
    us = r["us"].astype(np.float64)
    assert (r["retcode"] == 0).all() and np.isfinite(us).all()
    np.testing.assert_allclose(us[:, 0] + us[:, 1], 1.0, atol=1e-6)
    np.testing.assert_allclose(us[-1, 0], 1e4 / 10000.04, rtol=2e-7)
    assert np.all(np.abs(us[-1, 1] - (1.0 - 1e4 / 10000.04)) < 1e-7)


def _radau_robertson(k, ts):
    from scipy.integrate import solve_ivp

    def f(t, y):
        return [-k[0] * y[0] + k[2] * y[1] * y[2], k[0] * y[0] - k[1] * y[1] ** 2 - k[2] * y[1] * y[2],
                k[1] * y[1] ** 2]

    return solve_ivp(f, (0.0, ts[-1]), [1.0, 0.0, 0.0], method="Radau", rtol=1e-11, atol=1e-16, t_eval=ts).y


def test_robertson_dae_rb23_reference_settings(oracle):
    """gpu_ode_mass_matrix.jl:33-54: Rosenbrock23, abstol = reltol = 1e-5, dt = 0.1 over (0, 1e5),
    k = (0.04, 3e7, 1e4); the reference bound on the first and last states is 8e-4."""
    k = np.array([0.04, 3e7, 1e4])
    u0 = np.array([[1.0], [0.0], [0.0]], F32)
    r = P.solve("rober_dae", "rb23", True, u0, k.astype(F32)[:, None], (0.0, 1e5), 0.1, abstol=1e-5, reltol=1e-5,
                save_everystep=False, detpow=True)
    assert r["retcode"][0] == 0 and r["ts"][1, 0] == F32(1e5)
    ref = _radau_robertson(k, np.array([1e5]))[:, -1]
    assert np.linalg.norm(r["us"][1, :, 0].astype(np.float64) - ref) < 1e-4
    assert np.linalg.norm(r["us"][0, :, 0] - np.array([1.0, 0.0, 0.0])) == 0.0


@pytest.mark.parametrize("method", ["rodas4", "rodas5p"])
def test_robertson_dae_rodas_fixture(oracle, method):
    g = np.load("tests/golden/robertson_radau.npz")
    N = 64
    u0, p, sv = g["u0"][:, :N].astype(F32), g["p"][:, :N].astype(F32), g["t"].astype(F32)
    r = P.solve("rober_dae", method, True, u0, p, (0.0, 1e5), 1e-6, abstol=1e-8, reltol=1e-4, saveat=sv, detpow=True)
    assert (r["retcode"] == 0).all()
    us = r["us"].astype(np.float64)
    assert np.max(np.abs(us - g["u"][:, :, :N])) < 2e-5
    assert np.max(np.abs(us.sum(axis=1) - 1.0)) < 2e-6
