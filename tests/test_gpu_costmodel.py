"""GPU: the start-order cost model (deg_start_order_model), through the C ABI.

Bars:
  * the device key is 2^(w . x) with x the host features of the same trajectory (span / h0 from the
    kernel's own key) to 1e-4 relative (hardware log2 / exp2);
  * the order is a permutation in descending key buckets (the range-adaptive buckets of deg_order_by_span), and a persistent solve started in it equals
    the index-order solve bit for bit;
  * fitted on a held-out draw of the Lorenz law of the 1M adaptive workload, the model ranks the attempt
    counts of another draw better than span / h0 does (Spearman >= 0.9, against about 0.8);
  * a model of the wrong length is refused.
"""
import _orderbuckets as OB
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
F32 = np.float32


@pytest.fixture(scope="module")
def torch_cuda():
    import torch

    if not torch.cuda.is_available():
        pytest.fail("no HIP device visible: the GPU tests need an MI355X")
    return torch


def _rank(x):
    r = np.empty(len(x))
    r[np.argsort(x, kind="stable")] = np.arange(len(x))
    return r


def _spearman(a, b):
    return float(np.corrcoef(_rank(a), _rank(b))[0, 1])


def test_cost_model_key_order_and_results(deg, torch_cuda):
    torch = torch_cuda
    from diffeqgpu_amd import _lib, costmodel, systems, workloads

    f, n, np_, _ = systems.SYSTEMS["lorenz"]
    kw = dict(dt=0.1, abstol=1e-6, reltol=1e-6, saveat=1.0)
    pilot = workloads.config("lorenz_adaptive", N=20_000, seed=9001)
    test = workloads.config("lorenz_adaptive", N=20_000, seed=77)
    prob = deg.ODEProblem(f, test["u0"][:, 0], test["tspan"], test["p"][:, 0])
    print("[gpu] cost model: fit on the pilot", flush=True)
    pprobs = deg.HIPProbVector.from_arrays(f, pilot["u0"], pilot["p"], pilot["tspan"], deg.HIPDevice())
    w = costmodel.fit(pprobs, prob, deg.GPUTsit5(), **kw)
    assert len(w) == 3 + n + np_

    outs = {}
    for mode in ("none", "cost", "model"):
        print(f"[gpu] cost model: solve in {mode} order", flush=True)
        dev = deg.HIPDevice(persistent=True, start_order=mode, cost_model=w if mode == "model" else None)
        probs = deg.HIPProbVector.from_arrays(f, test["u0"], test["p"], test["tspan"], dev)
        r = deg.vectorized_asolve(probs, prob, deg.GPUTsit5(), **kw)
        torch.cuda.synchronize()
        outs[mode] = r
    # the device key against the host features of the same trajectories
    h0key = outs["cost"].order_key.cpu().numpy().astype(np.float64)
    mkey = outs["model"].order_key.cpu().numpy().astype(np.float64)
    X = costmodel.features(h0key, test["u0"], test["p"], np.full(test["N"], 10.0))
    np.testing.assert_allclose(mkey, 2.0 ** (X @ np.asarray(w, np.float64)), rtol=1e-4)
    o = outs["model"].order_auto.cpu().numpy()
    OB.assert_descending(outs["model"].order_key.cpu().numpy(), o)
    # results do not depend on the start order
    for x, y in ((outs["none"].ts, outs["model"].ts), (outs["none"].us_soa, outs["model"].us_soa),
                 (outs["none"].naccept, outs["model"].naccept), (outs["none"].nreject, outs["model"].nreject)):
        np.testing.assert_array_equal(x.cpu().numpy(), y.cpu().numpy())
    # the model ranks the attempt counts better than span / h0
    att = (outs["none"].naccept.to(torch.int64) + outs["none"].nreject.to(torch.int64)).cpu().numpy()
    s_model, s_h0 = _spearman(mkey, att), _spearman(h0key, att)
    print(f"[cost model] Spearman with attempts: model {s_model:.3f}, span / h0 {s_h0:.3f}", flush=True)
    assert s_model >= 0.9 and s_model > s_h0 + 0.05, (s_model, s_h0)
    # a model of the wrong length is refused
    r = outs["model"]
    with pytest.raises(_lib.DegError, match="3 \\+ n \\+ np"):
        r.kernel.start_order_model(r.args, r.cost_model.data_ptr(), 4, r.order_key.data_ptr(),
                                   r.order_auto.data_ptr(), r.order_work.data_ptr())


def test_ordered_inputs_layout_and_results(deg, torch_cuda):
    """With a cost order the ordering pass also writes u0 and p in queue-major order position
    (deg_solve_args.inputs_ordered): entry k of the order at column index (k mod 8) * ceil(N / 8) + k / 8.
    The solve that reads them equals, bit for bit, the one that gathers from u0 and p; N not a
    multiple of 8 exercises the padded columns. An inputs_ordered without an order is refused."""
    torch = torch_cuda
    from diffeqgpu_amd import _lib, systems, workloads

    f, n, np_, _ = systems.SYSTEMS["lorenz"]
    kw = dict(dt=0.1, abstol=1e-6, reltol=1e-6, saveat=1.0)
    for N in (20_003, 64):
        cfg = workloads.config("lorenz_adaptive", N=N, seed=31)
        prob = deg.ODEProblem(f, cfg["u0"][:, 0], cfg["tspan"], cfg["p"][:, 0])
        outs = {}
        for ordered in (True, False):
            print(f"[gpu] ordered inputs N={N} {ordered}", flush=True)
            dev = deg.HIPDevice(persistent=True, start_order="cost", ordered_inputs=ordered)
            probs = deg.HIPProbVector.from_arrays(f, cfg["u0"], cfg["p"], cfg["tspan"], dev)
            r = deg.vectorized_asolve(probs, prob, deg.GPUTsit5(), **kw)
            torch.cuda.synchronize()
            outs[ordered] = r
        r = outs[True]
        assert outs[False].inputs_ordered is None and r.inputs_ordered is not None
        slab = (N + 7) // 8
        inp = r.inputs_ordered.cpu().numpy().reshape(n + np_, 8 * slab)
        o = r.order_auto.cpu().numpy()
        k = np.arange(N)
        col = (k % 8) * slab + k // 8
        np.testing.assert_array_equal(inp[:n, col], cfg["u0"][:, o])
        np.testing.assert_array_equal(inp[n:, col], cfg["p"][:, o])
        for name in ("ts", "us_soa", "retcode", "naccept", "nreject"):
            np.testing.assert_array_equal(getattr(outs[True], name).cpu().numpy(),
                                          getattr(outs[False], name).cpu().numpy())
    # refused without an order
    a = r.args
    saved = a.order
    a.order = None
    with pytest.raises(_lib.DegError, match="inputs_ordered needs an order"):
        r.kernel.solve(a)
    a.order = saved
