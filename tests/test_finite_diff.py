"""autodiff = false: the implicit methods' Jacobian and time gradient by forward differences
(finite_diff_jac, alg_utils.jl:17-27; build_J_W / build_tgrad, nlsolve/type.jl:129-157).

  source            the emitted deg_jac / deg_tgrad difference deg_rhs with dx = sqrt(eps(Float32));
                    a user jac / tgrad still wins (has_jac is checked first)
  oracle J          u' = -p u^2 at u = 0.01: the forward difference is -p (2u + dx) (a central
                    difference or the exact derivative miss it by p dx = 3.5e-4); Lorenz within the
                    forward-difference bound of the analytic J (a transposed J is off by O(10));
                    the time gradient of u' = p t u is p u to rounding
  reference test    finite_diff.jl: u' = -u^2 from 10 on (0, 10), adaptive dt = 0.01, all five
                    methods, within 2e-4 of the solution 10 / (1 + 10 t)
  against AD        the same solves with the analytic J agree to 1e-4 relative (nonstiff systems)
GPU parity with the oracle is tested in test_gpu_finite_diff.py.
"""
import numpy as np
import pytest

import pyoracle as P

F32 = np.float32
IMPLICIT = ("rb23", "rodas4", "rodas5p", "kvaerno3", "kvaerno5")
DX = float(np.sqrt(F32(np.finfo(F32).eps)))


def test_fd_step_bits():
    from diffeqgpu_amd import tracer

    assert np.array([tracer.FD_STEP_BITS], np.uint32).view(F32)[0] == F32(DX)
    assert F32(DX) == F32(2.0 ** -11.5)


def test_fd_source(deg):
    from diffeqgpu_amd import systems

    f = deg.ODEFunction(systems.robertson)
    ad, _ = f.hip_source(3, 3, True)
    fd, _ = f.hip_source(3, 3, True, fd=True)
    assert "0x39b504f3u" in fd and "0x39b504f3u" not in ad
    assert fd.count("deg_rhs(") == 1 + 2 + 2  # the definition, two calls in deg_jac, two in deg_tgrad
    # fd only matters where a Jacobian is emitted
    assert f.hip_source(3, 3, False, fd=True)[0] == f.hip_source(3, 3, False)[0]
    # a user jac wins over the differences (has_jac first); the missing tgrad is still differenced
    g = deg.ODEFunction(systems.robertson, jac=systems.robertson_jac)
    src, _ = g.hip_source(3, 3, True, fd=True)
    jac_part = src[src.index("void deg_jac"):src.index("void deg_tgrad")]
    assert "deg_rhs(" not in jac_part and "0x39b504f3u" in src[src.index("void deg_tgrad"):]
    full = deg.ODEFunction(systems.robertson, jac=systems.robertson_jac, tgrad=systems.zero_tgrad3)
    assert "0x39b504f3u" not in full.hip_source(3, 3, True, fd=True)[0]


def test_fd_algorithms(deg):
    for name in ("GPURosenbrock23", "GPURodas4", "GPURodas5P", "GPUKvaerno3", "GPUKvaerno5"):
        a = getattr(deg, name)(autodiff=False)
        assert a.autodiff is False and "autodiff=False" in repr(a)
        assert getattr(deg, name)().autodiff is True


def test_fd_jit(deg):
    from diffeqgpu_amd import _lib, systems

    for name, (n, np_) in (("robertson", (3, 3)), ("tdep", (1, 1))):
        f = deg.ODEFunction(getattr(systems, name))
        src, _ = f.hip_source(n, np_, True, fd=True)
        for m in (_lib.DEG_RB23, _lib.DEG_RODAS4, _lib.DEG_RODAS5P, _lib.DEG_KVAERNO3, _lib.DEG_KVAERNO5):
            assert len(_lib.precompile(src, n, np_, m, _lib.F_PARITY)) == 16
    dae = systems.SYSTEMS["dae2"][0]
    src, _ = deg.ODEFunction(dae.f, mass_matrix=dae.mass_matrix).hip_source(2, 0, True, fd=True)
    assert len(_lib.precompile(src, 2, 0, _lib.DEG_RODAS5P, _lib.F_PARITY)) == 16


def test_fd_jacobian_quadratic(oracle):
    """f = -p u^2: (f(u + dx) - f(u)) / dx = -p (2u + dx) exactly up to rounding (about 1e-8 here)."""
    for u, p in ((0.01, 1.0), (0.02, 2.0), (-0.01, 0.5)):
        J, dT = P.jac("quad", [u], [p], 0.0, fd=True)
        expect = -p * (2.0 * F32(u) + DX)
        assert abs(float(J[0, 0]) - expect) < 1e-6, (u, p, J, expect)
        assert abs(float(J[0, 0]) - (-2.0 * p * u)) > 0.5 * p * DX  # not the exact derivative
        assert dT[0] == 0.0


def test_fd_jacobian_lorenz(oracle):
    rng = np.random.default_rng(3)
    eps = float(np.finfo(F32).eps)
    for _ in range(20):
        u = rng.uniform(-15, 15, 3).astype(F32)
        p = np.array([10.0, 28.0, 8.0 / 3.0], F32) * rng.uniform(0.5, 1.5, 3).astype(F32)
        Ja, _ = P.jac("lorenz", u, p, 0.0)
        Jf, dT = P.jac("lorenz", u, p, 0.0, fd=True)
        s, r, b = (float(x) for x in p)
        fmax = max(abs(s * (u[1] - u[0])), abs(u[0] * (r - u[2]) - u[1]), abs(u[0] * u[1] - b * u[2])) + 50.0
        # truncation: the Hessian entries are 0 or +-1, so dx / 2; rounding of f at +-|f| eps / dx twice
        bound = DX + 4.0 * eps * fmax / DX
        assert np.max(np.abs(Jf.astype(np.float64) - Ja)) < bound
        assert not np.any(dT)


def test_fd_time_gradient(oracle):
    for u, p, t in ((2.0, 0.5, 0.3), (1.0, 2.0, 0.9), (3.0, 1.5, 0.0)):
        J, dT = P.jac("tdep", [u], [p], t, fd=True)
        Ja, dTa = P.jac("tdep", [u], [p], t)
        assert abs(float(dT[0]) - p * u) < 4e-4 * p * u
        assert abs(float(J[0, 0]) - float(Ja[0, 0])) < 4e-4


@pytest.mark.parametrize("method", IMPLICIT)
def test_reference_finite_diff(oracle, method):
    """finite_diff.jl:1-46 on the oracle: 2 trajectories, save_everystep = false, adaptive, dt = 0.01,
    default tolerances; the reference bound 2e-4 against Rodas5P, here against 10 / (1 + 10 t)."""
    u0 = np.full((1, 2), 10.0, F32)
    p = np.ones((1, 2), F32)
    r = P.solve("quad", method, True, u0, p, (0.0, 10.0), 0.01, save_everystep=False, fdjac=True)
    assert (r["retcode"] == 0).all()
    end = r["us"][1, 0].astype(np.float64)
    assert np.all(np.abs(end - 10.0 / 101.0) < 2e-4)
    assert np.all(r["us"][0, 0] == 10.0)


@pytest.mark.parametrize("method", IMPLICIT)
@pytest.mark.parametrize("adaptive", [False, True])
def test_fd_against_analytic(oracle, method, adaptive):
    rng = np.random.default_rng(1)
    N = 32
    cases = [("vdp", rng.uniform(-2, 2, (2, N)), rng.uniform(0.1, 3, (1, N)), (0.0, 5.0), 0.01),
             ("tdep", np.ones((1, N)), rng.uniform(0.5, 2, (1, N)), (0.0, 1.0), 1 / 64)]
    for system, u0, p, tspan, dt in cases:
        kw = dict(abstol=1e-6, reltol=1e-6, save_everystep=False)
        r = P.solve(system, method, adaptive, u0, p, tspan, dt, fdjac=True, **kw)
        a = P.solve(system, method, adaptive, u0, p, tspan, dt, **kw)
        assert (r["retcode"] == 0).all()
        rel = np.abs(r["us"][1] - a["us"][1]) / (1.0 + np.abs(a["us"][1]))
        assert np.max(rel) < 1e-4, (system, np.max(rel))
