"""GPU parity of the implicit methods with a mass matrix (M u' = f) against the oracle, and the
reference's DAE tests replayed through the high-level API.

Bars:
  * no contraction, deterministic controller pow: bit-identical ts, us, return codes and step
    counts with the oracle (fixed and adaptive, persistent grid), for an explicit identity M, the
    Robertson DAE and the 2 x 2 linear DAE;
  * gpu_ode_mass_matrix.jl: GPURosenbrock23 on the Robertson DAE, abstol = reltol = 1e-5, dt = 0.1,
    first and last states within 8e-4 of an FP64 Radau solution;
  * gpu_ode_modelingtoolkit_dae.jl (direct mass-matrix DAE): all five methods, fixed dt = 0.001,
    no NaN and |u1 + u2 - 1| < 0.01 at the end.
"""
import numpy as np
import pytest

import pyoracle
from test_gpu_parity import assert_bitwise, gpu_solve

pytestmark = pytest.mark.gpu
F32 = np.float32
IMPLICIT = ("rb23", "rodas4", "rodas5p", "kvaerno3", "kvaerno5")
ALGS = {"rb23": "GPURosenbrock23", "rodas4": "GPURodas4", "rodas5p": "GPURodas5P", "kvaerno3": "GPUKvaerno3",
        "kvaerno5": "GPUKvaerno5"}


@pytest.fixture(scope="module")
def torch_cuda():
    import torch

    if not torch.cuda.is_available():
        pytest.fail("no HIP device visible: the GPU tests need an MI355X")
    return torch


def _inputs(system, N):
    if system == "dae2":
        u0 = np.zeros((2, N), F32)
        u0[0] = np.linspace(0.5, 1.0, N, dtype=F32)
        u0[1] = (F32(1.0) - u0[0]).astype(F32)
        return u0, np.zeros((0, N), F32)
    g = np.load("tests/golden/robertson_radau.npz")
    return g["u0"][:, :N].astype(F32).copy(), g["p"][:, :N].astype(F32).copy()


CASES = [("robertson_eye", m) for m in IMPLICIT] + [("dae2", m) for m in IMPLICIT] + \
        [("rober_dae", m) for m in IMPLICIT]


@pytest.mark.parametrize("system,method", CASES)
@pytest.mark.parametrize("adaptive", [False, True])
def test_mass_matrix_parity_bitwise(deg, torch_cuda, system, method, adaptive):
    N = 256
    u0, p = _inputs(system, N)
    if system == "dae2":
        tspan, dt = (0.0, 0.1), 0.001
        akw = dict(abstol=1e-6, reltol=1e-6)
    elif adaptive:
        tspan, dt = (0.0, 1e5), 1e-3
        akw = dict(abstol=1e-5, reltol=1e-5)
    else:
        tspan, dt = (0.0, 1.0), 1e-3
        akw = {}
    if adaptive:
        dev = deg.HIPDevice(contract="off", powmode="det", persistent=True, max_attempts=20000)
        g = gpu_solve(deg, torch_cuda, system, method, True, u0, p, tspan, dt, dev, save_everystep=False, **akw)
        o = pyoracle.solve(system, method, True, u0, p, tspan, dt, save_everystep=False, detpow=True,
                           max_attempts=20000, **akw)
    else:
        dev = deg.HIPDevice(contract="off")
        g = gpu_solve(deg, torch_cuda, system, method, False, u0, p, tspan, dt, dev, save_everystep=False)
        o = pyoracle.solve(system, method, False, u0, p, tspan, dt, save_everystep=False)
    assert_bitwise(g, o)
    # the DAEs are solved where the method applies: every method at fixed step on the linear DAE,
    # the Rosenbrock methods adaptively on both (the Kvaerno error estimate W \ (btilde . z) divides
    # the algebraic rows by gamma dt, so adaptive Kvaerno runs end at the step-size floor, in the
    # oracle as on the device)
    if (system == "dae2" and not adaptive) or (system != "robertson_eye" and adaptive and method in ("rb23", "rodas4", "rodas5p")):
        assert (g["retcode"] == 0).all()
        end = g["us"][1].astype(np.float64)
        assert np.max(np.abs(end.sum(axis=0) - 1.0)) < 1e-5


def test_reference_mass_matrix_robertson(deg, torch_cuda):
    """gpu_ode_mass_matrix.jl:5-54 through solve(...; adaptive = true): 2 trajectories."""
    from scipy.integrate import solve_ivp
    from diffeqgpu_amd import systems

    f = systems.SYSTEMS["rober_dae"][0]
    k = (0.04, 3e7, 1e4)
    prob = deg.ODEProblem(f, [1.0, 0.0, 0.0], (0.0, 1e5), list(k))
    print("[gpu] reference mass-matrix Robertson", flush=True)
    sol = deg.solve(deg.EnsembleProblem(prob, safetycopy=False), deg.GPURosenbrock23(), deg.EnsembleGPUKernel(),
                    trajectories=2, dt=0.1, adaptive=True, abstol=1e-5, reltol=1e-5, saveat=[0.0, 1e5])

    def rober(t, y):
        return [-k[0] * y[0] + k[2] * y[1] * y[2], k[0] * y[0] - k[1] * y[1] ** 2 - k[2] * y[1] * y[2],
                k[1] * y[1] ** 2]

    ref = solve_ivp(rober, (0.0, 1e5), [1.0, 0.0, 0.0], method="Radau", rtol=1e-11, atol=1e-16, t_eval=[0.0, 1e5]).y
    for s in sol.u:
        assert s.retcode == "Success"
        assert np.linalg.norm(s.u[0] - ref[:, 0]) < 8e-4
        assert np.linalg.norm(s.u[-1] - ref[:, -1]) < 8e-4


@pytest.mark.parametrize("method", IMPLICIT)
def test_reference_direct_mass_matrix_dae(deg, torch_cuda, method):
    """gpu_ode_modelingtoolkit_dae.jl:22-106 (direct mass-matrix DAE), fixed dt = 0.001 on (0, 0.1)."""
    from diffeqgpu_amd import systems

    f = systems.SYSTEMS["dae2"][0]
    prob = deg.ODEProblem(f, [1.0, 0.0], (0.0, 0.1))
    print(f"[gpu] reference direct DAE {method}", flush=True)
    sol = deg.solve(deg.EnsembleProblem(prob, safetycopy=False), getattr(deg, ALGS[method])(),
                    deg.EnsembleGPUKernel(), trajectories=2, dt=0.001, adaptive=False)
    assert len(sol.u) == 2
    end = sol.u[0].u[-1]
    assert not np.isnan(end).any()
    assert abs(float(end[0]) + float(end[1]) - 1.0) < 0.01
