"""GPU parity of dual-number (forward-mode) solves against the oracle, and the reference's
forward_diff.jl replayed through the high-level API.

Bars:
  * no contraction, deterministic controller pow: bit-identical ts, us (values and partials), return
    codes and step counts with the oracle's dual Lorenz (6 partials), fixed and adaptive
    (persistent grid), endpoint and saveat, for GPUTsit5, GPUVern7 and GPUVern9, and for the implicit
    methods GPURosenbrock23, GPURodas4, GPURodas5P, GPUKvaerno3 and GPUKvaerno5 with autodiff = false
    (the dual W factor: the value LU plus one solve per partial);
  * autodiff = true for the implicit methods (the dual Jacobian by forward mode over the traced dual
    graph, which the oracle does not restate): the partials at t = 1 within 1e-3 of the exact flow's
    d u / d (u0, p), as the oracle's own dual solves are pinned;
  * forward_diff.jl: u0 and p of Lorenz as duals seeded with the identity, trajectories = 2, fixed
    dt = 0.01 (endpoint) and adaptive dt = 0.01 for all eight methods (the implicit ones with
    autodiff = false): every trajectory succeeds, the values equal the plain solve's (fixed step: bit
    for bit in parity mode), the partials are finite and d u(t0) / d u0 is the identity.
"""
import numpy as np
import pytest

import pyoracle
from test_gpu_parity import assert_bitwise

pytestmark = pytest.mark.gpu
F32 = np.float32
ALGS = {"tsit5": "GPUTsit5", "vern7": "GPUVern7", "vern9": "GPUVern9"}
IMPLICIT = {"rb23": "GPURosenbrock23", "rodas4": "GPURodas4", "rodas5p": "GPURodas5P", "kvaerno3": "GPUKvaerno3",
            "kvaerno5": "GPUKvaerno5"}


@pytest.fixture(scope="module")
def torch_cuda():
    import torch

    if not torch.cuda.is_available():
        pytest.fail("no HIP device visible: the GPU tests need an MI355X")
    return torch


def _inputs(N):
    rng = np.random.default_rng(17)
    u0v = (np.array([[1.0], [0.0], [0.0]], F32) + rng.uniform(0, 0.5, (3, N)).astype(F32)).astype(F32)
    pv = (np.array([[10.0], [28.0], [8.0 / 3.0]], F32) * rng.uniform(0.3, 1.0, (1, N)).astype(F32)).astype(F32)
    seeds = np.eye(6, dtype=F32)
    u0 = np.concatenate([u0v, np.repeat(seeds[:3].reshape(-1, 1), N, axis=1)]).astype(F32)
    p = np.concatenate([pv, np.repeat(seeds[3:].reshape(-1, 1), N, axis=1)]).astype(F32)
    return u0, p


@pytest.mark.parametrize("method", ["tsit5", "vern7", "vern9"])
@pytest.mark.parametrize("adaptive", [False, True])
@pytest.mark.parametrize("mode", ["endpoint", "saveat"])
def test_dual_parity_bitwise(deg, torch_cuda, method, adaptive, mode):
    from diffeqgpu_amd import systems

    N = 256
    u0, p = _inputs(N)
    f, n, np_, _ = systems.SYSTEMS["lorenz_dual"]
    tspan, dt = (0.0, 2.0), 0.01
    kw = dict(abstol=1e-6, reltol=1e-6) if adaptive else {}
    okw = dict(kw)
    if mode == "saveat":
        sv = np.linspace(0.0, 2.0, 9).astype(F32)
        kw["saveat"] = sv
        okw["saveat"] = sv
    else:
        kw["save_everystep"] = False
    if adaptive:
        dev = deg.HIPDevice(contract="off", powmode="det", persistent=True)
    else:
        dev = deg.HIPDevice(contract="off")
    probs = deg.HIPProbVector.from_arrays(f, u0, p, tspan, dev)
    prob = deg.ODEProblem(systems.lorenz, u0[:3, 0], tspan, p[:3, 0])
    alg = getattr(deg, ALGS[method])()
    print(f"[gpu] dual lorenz {method} adaptive={adaptive} {mode} N={N}", flush=True)
    fn = deg.vectorized_asolve if adaptive else deg.vectorized_solve
    res = fn(probs, prob, alg, dt=dt, **kw)
    torch_cuda.cuda.synchronize()
    g = dict(ts=res.ts.cpu().numpy(), us=res.us_soa.cpu().numpy(), retcode=res.retcode.cpu().numpy(),
             naccept=res.naccept.cpu().numpy(), nreject=res.nreject.cpu().numpy())
    o = pyoracle.solve("lorenz_dual", method, adaptive, u0, p, tspan, dt, detpow=adaptive, **okw)
    assert_bitwise(g, o)
    assert (g["retcode"] == 0).all()


@pytest.mark.parametrize("method", list(IMPLICIT))
@pytest.mark.parametrize("adaptive", [False, True])
def test_dual_implicit_parity_bitwise(deg, torch_cuda, method, adaptive):
    from diffeqgpu_amd import systems

    N = 128
    u0, p = _inputs(N)
    f, n, np_, _ = systems.SYSTEMS["lorenz_dual"]
    tspan, dt = (0.0, 1.0), 0.01
    sv = np.linspace(0.0, 1.0, 5).astype(F32)
    kw = dict(abstol=1e-6, reltol=1e-6) if adaptive else {}
    dev = deg.HIPDevice(contract="off", powmode="det", persistent=True) if adaptive else deg.HIPDevice(contract="off")
    probs = deg.HIPProbVector.from_arrays(f, u0, p, tspan, dev)
    prob = deg.ODEProblem(systems.lorenz, u0[:3, 0], tspan, p[:3, 0])
    alg = getattr(deg, IMPLICIT[method])(autodiff=False)
    print(f"[gpu] dual lorenz implicit {method} adaptive={adaptive} N={N}", flush=True)
    fn = deg.vectorized_asolve if adaptive else deg.vectorized_solve
    res = fn(probs, prob, alg, dt=dt, saveat=sv, **kw)
    torch_cuda.cuda.synchronize()
    g = dict(ts=res.ts.cpu().numpy(), us=res.us_soa.cpu().numpy(), retcode=res.retcode.cpu().numpy(),
             naccept=res.naccept.cpu().numpy(), nreject=res.nreject.cpu().numpy())
    o = pyoracle.solve("lorenz_dual", method, adaptive, u0, p, tspan, dt, detpow=adaptive, saveat=sv, fdjac=True, **kw)
    assert_bitwise(g, o)
    assert (g["retcode"] == 0).all()


@pytest.mark.parametrize("method", ["rb23", "rodas5p", "kvaerno5"])
def test_dual_implicit_autodiff_against_flow_derivative(deg, torch_cuda, method):
    from diffeqgpu_amd import systems
    from test_dual import _fd_jacobian

    N = 16
    u0, p = _inputs(N)
    f, n, np_, _ = systems.SYSTEMS["lorenz_dual"]
    dev = deg.HIPDevice(contract="off")
    probs = deg.HIPProbVector.from_arrays(f, u0, p, (0.0, 1.0), dev)
    prob = deg.ODEProblem(systems.lorenz, u0[:3, 0], (0.0, 1.0), p[:3, 0])
    print(f"[gpu] dual lorenz implicit autodiff {method}", flush=True)
    res = deg.vectorized_solve(probs, prob, getattr(deg, IMPLICIT[method])(), dt=0.001, save_everystep=False)
    torch_cuda.cuda.synchronize()
    us = res.us_soa.cpu().numpy()
    te = float(res.ts.cpu().numpy()[-1, 0])
    assert (res.retcode.cpu().numpy() == 0).all()
    J = _fd_jacobian(u0[:3], p[:3], te, h=1e-5)
    parts = us[-1, 3:].astype(np.float64).reshape(3, 6, N)
    err = np.abs(parts - J).max(axis=(0, 1)) / np.abs(J).max(axis=(0, 1))
    assert err.max() < 1e-3, err.max()


def test_reference_forward_diff(deg, torch_cuda):
    """forward_diff.jl:1-58: the explicit methods, and the implicit ones with autodiff = false."""
    from diffeqgpu_amd import systems

    D = deg.Dual
    u0 = [D(1.0, (1, 0, 0, 0, 0, 0)), D(0.0, (0, 1, 0, 0, 0, 0)), D(0.0, (0, 0, 1, 0, 0, 0))]
    p = [D(10.0, (0, 0, 0, 1, 0, 0)), D(28.0, (0, 0, 0, 0, 1, 0)), D(8.0 / 3.0, (0, 0, 0, 0, 0, 1))]
    prob = deg.ODEProblem(systems.lorenz, u0, (0.0, 10.0), p)
    ep = deg.EnsembleProblem(prob, prob_func=lambda pr, ctx: deg.remake(pr, p=p), safetycopy=False)
    plain = deg.ODEProblem(systems.lorenz, [1.0, 0.0, 0.0], (0.0, 10.0), [10.0, 28.0, 8.0 / 3.0])
    algs = [(m, getattr(deg, nm)()) for m, nm in ALGS.items()]
    algs += [(m, getattr(deg, nm)(autodiff=False)) for m, nm in IMPLICIT.items()]
    for method, alg in algs:
        print(f"[gpu] reference forward_diff {method}", flush=True)
        dev = deg.HIPDevice(contract="off")
        sol = deg.solve(ep, alg, deg.EnsembleGPUKernel(dev), trajectories=2, save_everystep=False,
                        adaptive=False, dt=0.01)
        ref = deg.solve(deg.EnsembleProblem(plain), alg, deg.EnsembleGPUKernel(dev), trajectories=2,
                        save_everystep=False, adaptive=False, dt=0.01)
        for s, r in zip(sol.u, ref.u):
            assert s.retcode == "Success" and s.u.shape == (2, 3, 7)
            np.testing.assert_array_equal(s.u[:, :, 0], r.u)
            np.testing.assert_array_equal(s.u[0, :, 1:4], np.eye(3, dtype=F32))
            assert np.all(np.isfinite(s.u))
        asol = deg.solve(ep, alg, deg.EnsembleGPUKernel(), trajectories=2, adaptive=True, dt=0.01)
        for s in asol.u:
            assert s.retcode == "Success" and s.u.shape[1:] == (3, 7) and np.all(np.isfinite(s.u))


def test_dual_hill_at_zero(deg, torch_cuda):
    """A Hill feedback u' = -u + u^p / (1 + u^p) from u0 = 0 with p = 2, u0 and p seeded: u stays 0,
    so d u / d u0 = exp(-t) and d u / d p = 0. The power's partials at a zero base go through
    ForwardDiff's guards (the log term is 0 there); without them they are NaN."""
    from test_dual import _hill

    D = deg.Dual
    prob = deg.ODEProblem(_hill, [D(0.0, (1.0, 0.0))], (0.0, 2.0), [D(2.0, (0.0, 1.0))])
    sv = np.linspace(0.0, 2.0, 5).astype(F32)
    for name in ("GPUTsit5", "GPUVern7"):
        print(f"[gpu] dual hill at zero {name}", flush=True)
        sol = deg.solve(deg.EnsembleProblem(prob), getattr(deg, name)(), deg.EnsembleGPUKernel(), trajectories=3,
                        adaptive=True, dt=0.01, abstol=1e-8, reltol=1e-8, saveat=sv)
        for s in sol.u:
            assert s.retcode == "Success" and s.u.shape == (5, 1, 3)
            assert np.all(np.isfinite(s.u))
            np.testing.assert_array_equal(s.u[:, 0, 0], 0.0)
            np.testing.assert_allclose(s.u[:, 0, 1], np.exp(-sv.astype(np.float64)), rtol=1e-5)
            np.testing.assert_array_equal(s.u[:, 0, 2], 0.0)
