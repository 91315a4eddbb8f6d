"""Writes the FP64 reference fixtures under tests/golden/ (run from the repo root).

The reference solutions come from the oracle's FP64 Vern9 (abstol = reltol = 1e-12, the
same PI controller), itself pinned against scipy DOP853/Radau in tests/test_oracle_pins.py.
Inputs are the seeded workloads of diffeqgpu_amd.workloads, so the fixtures match the
benchmark's inputs exactly.
  lorenz_adaptive_vern9.npz  first 1,000 trajectories of the 1M adaptive Lorenz workload:
                             u(t) at t = 0, 1, ..., 10 (float64, (11, 3, 1000))
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import __graft_entry__ as g  # noqa: E402
import pyoracle  # noqa: E402


def main():
    g.load_package()
    from diffeqgpu_amd import workloads

    cfg = workloads.config("lorenz_adaptive")
    u0, p = cfg["u0"][:, :1000], cfg["p"][:, :1000]
    tsave = np.arange(11, dtype=np.float64)
    out, st, ns = pyoracle.vern9("lorenz", u0, p, (0.0, 10.0), saveat=tsave, abstol=1e-12, reltol=1e-12)
    assert np.all(st == 0)
    np.savez_compressed(os.path.join(os.path.dirname(__file__), "lorenz_adaptive_vern9.npz"), u=out, t=tsave,
                        u0=u0, p=p, note="FP64 Vern9 1e-12, first 1000 trajectories of workloads.config('lorenz_adaptive')")
    print("lorenz_adaptive_vern9.npz", out.shape, int(ns.mean()), "steps/trajectory")


if __name__ == "__main__":
    main()
