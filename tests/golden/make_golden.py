"""Writes the FP64 reference fixtures under tests/golden/ (run from the repo root).

The reference solutions come from the oracle's FP64 Vern9 (abstol = reltol = 1e-12, the
same PI controller), itself pinned against scipy DOP853/Radau in tests/test_oracle_pins.py.
Inputs are the seeded workloads of diffeqgpu_amd.workloads, so the fixtures match the
benchmark's inputs exactly.
  lorenz_adaptive_vern9.npz  first 1,000 trajectories of the 1M adaptive Lorenz workload:
                             u(t) at t = 0, 1, ..., 10 (float64, (11, 3, 1000))
  lorenz_uneven_vern9.npz    first 1,000 trajectories of the 4M uneven Lorenz workload (sweep 0):
                             u(tend_i) (float64, (1, 3, 1000))
  vdp_vern9.npz              first 1,000 trajectories of the 2M Van der Pol workload: u(50)
  robertson_radau.npz        first 1,000 trajectories of the 1M Robertson workload at the 20
                             log-spaced save times; scipy Radau (rtol 1e-11, atol 1e-16), since an
                             explicit FP64 method cannot cross this stiff problem
Usage: python tests/golden/make_golden.py [name ...] (default: all).
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import __graft_entry__ as g  # noqa: E402
import pyoracle  # noqa: E402


HERE = os.path.dirname(os.path.abspath(__file__))


def lorenz_adaptive(workloads):
    cfg = workloads.config("lorenz_adaptive")
    u0, p = cfg["u0"][:, :1000], cfg["p"][:, :1000]
    tsave = np.arange(11, dtype=np.float64)
    out, st, ns = pyoracle.vern9("lorenz", u0, p, (0.0, 10.0), saveat=tsave, abstol=1e-12, reltol=1e-12)
    assert np.all(st == 0)
    np.savez_compressed(os.path.join(os.path.dirname(__file__), "lorenz_adaptive_vern9.npz"), u=out, t=tsave,
                        u0=u0, p=p, note="FP64 Vern9 1e-12, first 1000 trajectories of workloads.config('lorenz_adaptive')")
    print("lorenz_adaptive_vern9.npz", out.shape, int(ns.mean()), "steps/trajectory")


def lorenz_uneven(workloads):
    cfg = workloads.config("lorenz_uneven", N=1000)
    out, st, ns = pyoracle.vern9("lorenz", cfg["u0"], cfg["p"], cfg["tspan"], tfs=cfg["tfs"], abstol=1e-12,
                                 reltol=1e-12)
    assert np.all(st == 0)
    np.savez_compressed(os.path.join(HERE, "lorenz_uneven_vern9.npz"), u=out, u0=cfg["u0"], p=cfg["p"],
                        tfs=cfg["tfs"], note="FP64 Vern9 1e-12 at tend, first 1000 of workloads.config('lorenz_uneven')")
    print("lorenz_uneven_vern9.npz", out.shape, int(ns.mean()), "steps/trajectory")


def vdp(workloads):
    cfg = workloads.config("vdp", N=1000)
    out, st, ns = pyoracle.vern9("vdp", cfg["u0"], cfg["p"], cfg["tspan"], abstol=1e-12, reltol=1e-12)
    assert np.all(st == 0)
    np.savez_compressed(os.path.join(HERE, "vdp_vern9.npz"), u=out, u0=cfg["u0"], p=cfg["p"],
                        note="FP64 Vern9 1e-12 at t = 50, first 1000 of workloads.config('vdp')")
    print("vdp_vern9.npz", out.shape, int(ns.mean()), "steps/trajectory")


def _radau_one(args):
    from scipy.integrate import solve_ivp

    k, sv = args

    def f(t, y):
        return [-k[0] * y[0] + k[2] * y[1] * y[2], k[0] * y[0] - k[1] * y[1] ** 2 - k[2] * y[1] * y[2], k[1] * y[1] ** 2]

    def jac(t, y):
        return [[-k[0], k[2] * y[2], k[2] * y[1]], [k[0], -2 * k[1] * y[1] - k[2] * y[2], -k[2] * y[1]],
                [0.0, 2 * k[1] * y[1], 0.0]]

    s = solve_ivp(f, (0.0, 1e5), [1.0, 0.0, 0.0], method="Radau", rtol=1e-11, atol=1e-16, t_eval=sv, jac=jac)
    assert s.success
    return s.y.T


def robertson(workloads):
    from multiprocessing import Pool

    cfg = workloads.config("robertson", N=1000)
    sv = np.asarray(cfg["saveat"], dtype=np.float32).astype(np.float64)   # the Float32 save times the solver uses
    with Pool(min(8, os.cpu_count() or 1)) as pool:
        ys = pool.map(_radau_one, [(cfg["p"][:, i].astype(np.float64), sv) for i in range(1000)])
    out = np.stack(ys, axis=-1)  # (20, 3, 1000)
    np.savez_compressed(os.path.join(HERE, "robertson_radau.npz"), u=out, t=sv, u0=cfg["u0"], p=cfg["p"],
                        note="scipy Radau rtol 1e-11 atol 1e-16, first 1000 of workloads.config('robertson')")
    print("robertson_radau.npz", out.shape)


FIXTURES = {"lorenz_adaptive": lorenz_adaptive, "lorenz_uneven": lorenz_uneven, "vdp": vdp, "robertson": robertson}


def main(names):
    g.load_package()
    from diffeqgpu_amd import workloads

    for name in names or list(FIXTURES):
        FIXTURES[name](workloads)


if __name__ == "__main__":
    main(sys.argv[1:])
