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


@pytest.mark.parametrize("name,N,devkw", [("lorenz_adaptive", 40_000, {}), ("lorenz_adaptive", 3000, dict(contract="off", powmode="det")),
                                          ("lorenz_adaptive", 5000, dict(persistent=False)), ("lorenz_adaptive", 70, {})])
def test_batched_saves_bitwise(deg, torch_cuda, name, N, devkw):
    """Batched dense output (DEG_F_BATCH_SAVES) writes exactly what the per-step saves write."""
    torch = torch_cuda
    from diffeqgpu_amd import systems, workloads

    cfg = workloads.config(name, N=N)
    f, n, np_, _ = systems.SYSTEMS[cfg["system"]]
    prob = deg.ODEProblem(f, cfg["u0"][:, 0], cfg["tspan"], cfg["p"][:, 0])
    outs = []
    for bs in (False, True):
        dev = deg.HIPDevice(batch_saves=bs, **devkw)
        probs = deg.HIPProbVector.from_arrays(f, cfg["u0"], cfg["p"], cfg["tspan"], dev)
        for saveat in (1.0, np.array([0.05, 0.5, 0.51, 0.52, 3.3, 9.99, 10.0], np.float32)):
            print(f"[gpu] batch_saves={bs} N={N} {devkw} saveat={saveat if np.ndim(saveat) == 0 else 'list'}", flush=True)
            r = deg.vectorized_asolve(probs, prob, deg.GPUTsit5(), dt=0.1, abstol=1e-6, reltol=1e-6, saveat=saveat)
            torch.cuda.synchronize()
            outs.append([x.cpu().numpy() for x in (r.ts, r.us_soa, r.retcode, r.naccept, r.nreject)])
    for a_, b_ in zip(outs[:2], outs[2:]):
        for x, y in zip(a_, b_):
            np.testing.assert_array_equal(x, y)
