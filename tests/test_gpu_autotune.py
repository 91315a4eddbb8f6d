"""GPU: the launch-configuration autotuner, and configuration independence of the results.

Bars: every configuration the tuner can choose (block size, register cap, refill batch,
chunk) gives bit-identical outputs (each lane's arithmetic does not depend on where or when it
runs); the tuned configuration is one of the candidates and is no slower than the default by
more than timing noise (10 %).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch

    if not torch.cuda.is_available():
        pytest.fail("no HIP device visible: the GPU tests need an MI355X")
    return torch


def test_configurations_bitwise_equal_and_tuner(deg, torch_cuda, tmp_path, monkeypatch):
    torch = torch_cuda
    from diffeqgpu_amd import autotune, systems, workloads

    monkeypatch.setattr(autotune, "CACHE_PATH", str(tmp_path / "autotune.json"))
    monkeypatch.setattr(autotune, "SHIPPED_PATH", str(tmp_path / "none.json"))
    f, n, np_, _ = systems.SYSTEMS["lorenz"]
    cfg = workloads.config("lorenz_adaptive", N=30_000)
    prob = deg.ODEProblem(f, cfg["u0"][:, 0], cfg["tspan"], cfg["p"][:, 0])
    kw = dict(dt=0.1, abstol=1e-6, reltol=1e-6, saveat=1.0)

    def run(**devkw):
        dev = deg.HIPDevice(**devkw)
        probs = deg.HIPProbVector.from_arrays(f, cfg["u0"], cfg["p"], cfg["tspan"], dev)
        print(f"[gpu] lorenz 30000 {devkw}", flush=True)
        r = deg.vectorized_asolve(probs, prob, deg.GPUTsit5(), **kw)
        torch.cuda.synchronize()
        return [x.cpu().numpy() for x in (r.ts, r.us_soa, r.retcode, r.naccept, r.nreject)]

    ref = run()
    for devkw in (dict(block_size=64), dict(block_size=192, refill=1), dict(waves=7), dict(waves=5, chunk=4),
                  dict(refill=16, chunk=64), dict(blocks_per_cu=1), dict(persistent=False, block_size=128)):
        got = run(**devkw)
        for a, b in zip(ref, got):
            np.testing.assert_array_equal(a, b, err_msg=str(devkw))

    probs = deg.HIPProbVector.from_arrays(f, cfg["u0"], cfg["p"], cfg["tspan"], deg.HIPDevice())
    tuned = autotune.tune(probs, prob, deg.GPUTsit5(), log=print, **kw)
    assert tuned.block_size in autotune.BLOCKS and tuned.waves in autotune.WAVES
    assert tuned.refill in autotune.REFILL and tuned.chunk in autotune.CHUNK
    assert tuned.blocks_per_cu in autotune.blocks_per_cu_candidates(tuned.block_size)
    entry = next(iter(autotune._load().values()))
    assert entry["kernel_ms"] <= entry["default_ms"] * 1.10
    # cached: the second call returns the same configuration without timing
    assert autotune.tune(probs, prob, deg.GPUTsit5(), **kw) == tuned
    got = run(block_size=tuned.block_size, waves=tuned.waves, refill=tuned.refill, chunk=tuned.chunk,
              blocks_per_cu=tuned.blocks_per_cu)
    for a, b in zip(ref, got):
        np.testing.assert_array_equal(a, b)
