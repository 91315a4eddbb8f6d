"""GPU: tail compaction (DEG_F_COMPACT) moves trajectory states between waves of a workgroup
through LDS; the results must be bit-identical to the solve without it, for every method, with
and without saveat, for block sizes from one wave (nothing to merge) to four, and for ensembles
smaller than one block."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch

    if not torch.cuda.is_available():
        pytest.fail("no HIP device visible: the GPU tests need an MI355X")
    return torch


CASES = [("lorenz_adaptive", 40_000), ("vdp", 40_000), ("robertson", 30_000), ("lorenz_uneven", 20_000),
         ("lorenz_adaptive", 100), ("vdp", 1)]


@pytest.mark.parametrize("name,N", CASES)
def test_compaction_bitwise(deg, torch_cuda, name, N):
    torch = torch_cuda
    from diffeqgpu_amd import systems, workloads

    cfg = workloads.config(name, N=N)
    f, n, np_, _ = systems.SYSTEMS[cfg["system"]]
    prob = deg.ODEProblem(f, cfg["u0"][:, 0], cfg["tspan"], cfg["p"][:, 0])
    alg = {"tsit5": deg.GPUTsit5, "vern7": deg.GPUVern7, "rb23": deg.GPURosenbrock23}[cfg["method"]]()
    kw = dict(dt=cfg["dt"], abstol=cfg["abstol"], reltol=cfg["reltol"])
    if cfg.get("saveat") is not None:
        kw["saveat"] = cfg["saveat"]

    def run(**devkw):
        dev = deg.HIPDevice(kahan=bool(cfg.get("kahan")), **devkw)
        probs = deg.HIPProbVector.from_arrays(f, cfg["u0"], cfg["p"], cfg["tspan"], dev, tfs=cfg.get("tfs"))
        print(f"[gpu] compact {name} N={N} {devkw}", flush=True)
        r = deg.vectorized_asolve(probs, prob, alg, **kw)
        torch.cuda.synchronize()
        return [x.cpu().numpy() for x in (r.ts, r.us_soa, r.retcode, r.naccept, r.nreject)]

    ref = run()
    for devkw in (dict(compact=True), dict(compact=True, block_size=64), dict(compact=True, block_size=128, refill=1),
                  dict(compact=True, blocks_per_cu=2, chunk=4)):
        got = run(**devkw)
        for a, b in zip(ref, got):
            np.testing.assert_array_equal(a, b, err_msg=str(devkw))
