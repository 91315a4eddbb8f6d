"""Seeded synthetic ensembles for the benchmark configurations (structure-of-arrays inputs).

C1 lorenz_fixed    10,000 Lorenz, p = [10, 28, 8/3] .* U^3, tspan (0, 10), Tsit5 dt = 0.1
C2 lorenz_adaptive 1,000,000 Lorenz, same p law, Tsit5 abstol = reltol = 1e-6, saveat 0:1:10
C3 lorenz_uneven   4,000,000 Lorenz, rho~U(0,60), sigma~U(5,15), beta~U(1,4), tend~U(1,50),
                   Vern7 abstol = reltol = 1e-8, compensated t, endpoint only
C4 robertson       1,000,000 Robertson, k = (0.04, 3e7, 1e4) .* exp(0.3 Z), tspan (0, 1e5),
                   Rosenbrock23 reltol = 1e-4, abstol = 1e-8, 20 log-spaced saves
C5 vdp             2,000,000 Van der Pol, mu ~ logU(0.1, 30), u0 ~ U(-2, 2)^2, tspan (0, 50),
                   Tsit5 abstol = reltol = 1e-6
The parameter law of C1/C2 is the reference's random-parameter Lorenz ensemble
(test/gpu_kernel_de/gpu_ode_regression.jl:128: p = rand(Float32, 3) .* p).
"""
from __future__ import annotations

import numpy as np

from .rng import normal_table, uniform_table

LORENZ_P = np.array([10.0, 28.0, np.float32(8.0) / np.float32(3.0)], dtype=np.float32)


def lorenz_random(N: int, seed: int):
    U = uniform_table(seed, 3, N)
    p = (LORENZ_P[:, None] * U).astype(np.float32)
    u0 = np.zeros((3, N), np.float32)
    u0[0] = 1.0
    return u0, p


def config(name: str, N: int | None = None, seed: int | None = None) -> dict:
    if name == "lorenz_fixed":
        N = N or 10_000
        u0, p = lorenz_random(N, 101 if seed is None else seed)
        return dict(system="lorenz", method="tsit5", adaptive=False, u0=u0, p=p, tspan=(0.0, 10.0), dt=0.1,
                    saveat=None, save_everystep=False, N=N)
    if name == "lorenz_adaptive":
        N = N or 1_000_000
        u0, p = lorenz_random(N, 202 if seed is None else seed)
        return dict(system="lorenz", method="tsit5", adaptive=True, u0=u0, p=p, tspan=(0.0, 10.0), dt=0.1,
                    abstol=1e-6, reltol=1e-6, saveat=1.0, N=N)
    if name == "lorenz_uneven":
        N = N or 4_000_000
        s = 303 if seed is None else seed
        U = uniform_table(s, 4, N)
        p = np.stack([5.0 + 10.0 * U[0], 60.0 * U[1], 1.0 + 3.0 * U[2]]).astype(np.float32)
        tfs = (1.0 + 49.0 * U[3]).astype(np.float32)
        u0 = np.zeros((3, N), np.float32)
        u0[0] = 1.0
        return dict(system="lorenz", method="vern7", adaptive=True, u0=u0, p=p, tspan=(0.0, 50.0), tfs=tfs,
                    dt=0.01, abstol=1e-8, reltol=1e-8, saveat=None, kahan=True, N=N)
    if name == "robertson":
        N = N or 1_000_000
        s = 404 if seed is None else seed
        Z = normal_table(s, 3, N)
        p = (np.array([0.04, 3e7, 1e4])[:, None] * np.exp(0.3 * Z)).astype(np.float32)
        u0 = np.zeros((3, N), np.float32)
        u0[0] = 1.0
        saveat = np.array([10.0 ** (-4 + 9 * k / 19) for k in range(20)], dtype=np.float64)
        return dict(system="robertson", method="rb23", adaptive=True, u0=u0, p=p, tspan=(0.0, 1e5), dt=1e-6,
                    abstol=1e-8, reltol=1e-4, saveat=saveat, N=N)
    if name == "vdp":
        N = N or 2_000_000
        s = 505 if seed is None else seed
        U = uniform_table(s, 3, N).astype(np.float64)
        mu = np.exp(np.log(0.1) + (np.log(30.0) - np.log(0.1)) * U[0]).astype(np.float32)[None, :]
        u0 = (-2.0 + 4.0 * U[1:3]).astype(np.float32)
        return dict(system="vdp", method="tsit5", adaptive=True, u0=u0, p=mu, tspan=(0.0, 50.0), dt=0.01,
                    abstol=1e-6, reltol=1e-6, saveat=None, N=N)
    raise KeyError(name)
