"""GPU: device-side parameter draws and graph-captured sweeps (the uneven 4M Lorenz Vern7
workload's 10-sweep replay, at a size the oracle finishes in seconds).

Bars:
  * device draws equal the host generator (rng.uniform_table, Float32 lo + scale * U) bit for
    bit, for every row of a 4-row table and with an id offset;
  * every graph replay equals the eager run of the same sweeps bit for bit (outputs and drawn
    parameters). The replays are separated by other eager GPU work and synchronisations;
  * with the deterministic controller pow and no contraction, each sweep's results equal the
    CPU oracle on the host-drawn parameters of that sweep, bit for bit.
"""
import _orderbuckets as OB
import numpy as np
import pytest

import pyoracle

pytestmark = pytest.mark.gpu
F32 = np.float32

UNEVEN_DRAWS = [("p", 0, 5.0, 10.0), ("p", 1, 0.0, 60.0), ("p", 2, 1.0, 3.0), ("tfs", 0, 1.0, 49.0)]


@pytest.fixture(scope="module")
def torch_cuda():
    import torch

    if not torch.cuda.is_available():
        pytest.fail("no HIP device visible: the GPU tests need an MI355X")
    return torch


def host_uneven(seed, N):
    from diffeqgpu_amd import rng

    U = rng.uniform_table(seed, 4, N)
    p = np.stack([F32(5.0) + F32(10.0) * U[0], F32(60.0) * U[1], F32(1.0) + F32(3.0) * U[2]]).astype(F32)
    tfs = (F32(1.0) + F32(49.0) * U[3]).astype(F32)
    return p, tfs


def test_device_draws_match_host(deg, torch_cuda):
    torch = torch_cuda
    from diffeqgpu_amd import _lib, rng

    N, off = 100_003, 777
    out = torch.empty(N, dtype=torch.float32, device="cuda")
    st = torch.cuda.current_stream()
    for seed in (0, 303, 2 ** 63 + 12345):
        for row, (lo, scale) in enumerate([(5.0, 10.0), (0.0, 60.0), (1.0, 3.0), (-2.0, 4.0)]):
            print(f"[gpu] draw seed={seed} row={row}", flush=True)
            _lib.draw_uniform(out.data_ptr(), N, seed, 4, row, lo, scale, st.cuda_stream, id_offset=off)
            torch.cuda.synchronize()
            U = rng.uniform_table(seed, 4, off + N)[row, off:]
            want = (F32(lo) + F32(scale) * U).astype(F32)
            np.testing.assert_array_equal(out.cpu().numpy(), want)


def _snapshot(sw):
    # the start order's key, permutation and work counters are scratch (the order inside a key bucket
    # follows the atomics of the sort, and the solve's results do not depend on it)
    return [{k: v.cpu().numpy().copy() for k, v in o.items()
             if v is not None and k not in ("order_key", "order_auto", "order_work")} for o in sw.outputs]


def test_sweep_graph_replays_equal_eager_and_oracle(deg, torch_cuda):
    torch = torch_cuda
    from diffeqgpu_amd import systems

    f, n, np_, _ = systems.SYSTEMS["lorenz"]
    N, sweeps, seed = 3000, 3, 303
    dev = deg.HIPDevice(contract="off", powmode="det", kahan=True, persistent=True)
    p0, tfs0 = host_uneven(seed, N)
    u0 = np.zeros((3, N), F32)
    u0[0] = 1.0
    probs = deg.HIPProbVector.from_arrays(f, u0, np.zeros_like(p0), (0.0, 50.0), dev, tfs=np.ones(N, F32))
    prob = deg.ODEProblem(f, u0[:, 0], (0.0, 50.0), p0[:, 0])
    print("[gpu] sweep capture N=3000 sweeps=3 vern7 kahan persistent", flush=True)
    sw = deg.ParameterSweep(probs, prob, deg.GPUVern7(), sweeps=sweeps, seed=seed, draws=UNEVEN_DRAWS,
                            dt=0.01, abstol=1e-8, reltol=1e-8, save_everystep=False)
    assert sw.num_nodes() >= sweeps * 5
    eager = _snapshot(sw)

    # each sweep against the oracle on host-drawn parameters
    for k in range(sweeps):
        pk, tk = host_uneven(seed + k, N)
        sl = slice(0, 400)
        o = pyoracle.solve("lorenz", "vern7", True, u0[:, sl], pk[:, sl], (0.0, 50.0), 0.01, abstol=1e-8, reltol=1e-8,
                           tfs=tk[sl], kahan=True, detpow=True)
        np.testing.assert_array_equal(eager[k]["ts"][:, sl], o["ts"])
        np.testing.assert_array_equal(eager[k]["us_soa"][:, :, sl], o["us"])
        np.testing.assert_array_equal(eager[k]["naccept"][sl], o["naccept"])
        np.testing.assert_array_equal(eager[k]["retcode"][sl], o["retcode"])

    # replays, with unrelated eager GPU work and synchronisations between them
    for rep in range(3):
        for o in sw.outputs:
            for v in o.values():
                if v is not None and v.dtype == torch.float32 and v.dim() > 1:
                    v.fill_(float("nan"))
        scratch = torch.randn(1 << 20, device="cuda")
        (scratch * 2.0).sum().item()
        print(f"[gpu] replay {rep}", flush=True)
        sw.replay()
        sw.synchronize()
        got = _snapshot(sw)
        for k in range(sweeps):
            for key in eager[k]:
                np.testing.assert_array_equal(got[k][key], eager[k][key], err_msg=f"replay {rep} sweep {k} {key}")
        p_last, tfs_last = host_uneven(seed + sweeps - 1, N)
        np.testing.assert_array_equal(probs.p.cpu().numpy(), p_last)
        np.testing.assert_array_equal(probs.tfs.cpu().numpy(), tfs_last)


def test_graph_capture_rejects_timing_and_null_stream(deg, torch_cuda):
    from diffeqgpu_amd import _lib

    with pytest.raises(_lib.DegError):
        _lib.Graph.begin(0)


def test_order_by_span_and_order_independence(deg, torch_cuda):
    """deg_order_by_span returns a permutation in descending span buckets; solving in that order
    (or any other) gives bit-identical results."""
    torch = torch_cuda
    from diffeqgpu_amd import systems

    f, n, np_, _ = systems.SYSTEMS["lorenz"]
    N = 5000
    p0, tfs = host_uneven(404, N)
    u0 = np.zeros((3, N), F32)
    u0[0] = 1.0
    dev = deg.HIPDevice(kahan=True)
    probs = deg.HIPProbVector.from_arrays(f, u0, p0, (0.0, 50.0), dev, tfs=tfs)
    prob = deg.ODEProblem(f, u0[:, 0], (0.0, 50.0), p0[:, 0])
    print("[gpu] order_by_span N=5000", flush=True)
    order = deg.order_by_span(probs)
    torch.cuda.synchronize()
    o = order.cpu().numpy()
    OB.assert_descending(tfs, o)
    kw = dict(dt=0.01, abstol=1e-8, reltol=1e-8)
    outs = []
    for ordr in (None, order, torch.as_tensor(np.arange(N - 1, -1, -1, dtype=np.int32), device="cuda")):
        r = deg.vectorized_asolve(probs, prob, deg.GPUVern7(), order=ordr, **kw)
        torch.cuda.synchronize()
        outs.append([x.cpu().numpy() for x in (r.ts, r.us_soa, r.retcode, r.naccept, r.nreject)])
    for other in outs[1:]:
        for a, b in zip(outs[0], other):
            np.testing.assert_array_equal(a, b)
    # a sweep with longest-first ordering matches the same sweep without it
    print("[gpu] sweep longest_first", flush=True)
    kw2 = dict(dt=0.01, abstol=1e-8, reltol=1e-8, save_everystep=False)
    res = []
    for lf in (False, True):
        sw = deg.ParameterSweep(probs, prob, deg.GPUVern7(), sweeps=2, seed=77, draws=UNEVEN_DRAWS, longest_first=lf,
                                **kw2)
        sw.replay()
        sw.synchronize()
        res.append(_snapshot(sw))
    for k in range(2):
        for key in res[0][k]:
            np.testing.assert_array_equal(res[0][k][key], res[1][k][key])


def test_order_by_span_layouts_and_ragged_sizes(deg, torch_cuda):
    """deg_order_by_span for each combination of per-trajectory t0s / tfs arrays (present or scalar),
    at sizes around the 8192-trajectory block of its passes (one trajectory, half a block, one short
    of a block, exactly one, one over, three blocks and one), with zero, negative and NaN spans among
    the keys: a permutation in non-increasing bucket order each time."""
    torch = torch_cuda
    from diffeqgpu_amd import _lib

    rs = np.random.default_rng(5)
    for N in (1, 4096, 8191, 8192, 8193, 24577):
        t0s = rs.uniform(-1.0, 1.0, N).astype(F32)
        tfs = (t0s + rs.lognormal(0.0, 2.0, N)).astype(F32)
        if N > 8:
            tfs[3] = t0s[3]            # zero span
            tfs[5] = t0s[5] - 1.0      # negative span
            tfs[7] = np.float32("nan")  # NaN span
        for has_t0, has_tf in ((True, True), (True, False), (False, True), (False, False)):
            t0_d = torch.as_tensor(t0s, device="cuda") if has_t0 else None
            tf_d = torch.as_tensor(tfs, device="cuda") if has_tf else None
            t0, tf = 0.25, 10.0
            keys = (tfs if has_tf else np.full(N, tf, F32)) - (t0s if has_t0 else np.full(N, t0, F32))
            order = torch.full((N,), -1, dtype=torch.int32, device="cuda")
            work = torch.empty(_lib.ORDER_WORK_INTS, dtype=torch.int32, device="cuda")
            print(f"[gpu] order_by_span N={N} t0s={has_t0} tfs={has_tf}", flush=True)
            _lib.order_by_span(t0_d.data_ptr() if has_t0 else 0, tf_d.data_ptr() if has_tf else 0, t0, tf, N,
                               order.data_ptr(), work.data_ptr(), torch.cuda.current_stream().cuda_stream)
            torch.cuda.synchronize()
            OB.assert_descending(keys, order.cpu().numpy())


def test_order_entries_outside_range_are_skipped(deg, torch_cuda):
    """A caller's start order with entries outside [0, N) solves every listed trajectory exactly
    as without an order and leaves the rest alone (the kernel's own bounds check)."""
    torch = torch_cuda
    from diffeqgpu_amd import systems, workloads

    f, n, np_, _ = systems.SYSTEMS["lorenz"]
    cfg = workloads.config("lorenz_adaptive", N=2000)
    prob = deg.ODEProblem(f, cfg["u0"][:, 0], cfg["tspan"], cfg["p"][:, 0])
    kw = dict(dt=0.1, abstol=1e-6, reltol=1e-6, saveat=1.0)
    for persistent in (True, False):
        dev = deg.HIPDevice(persistent=persistent)
        probs = deg.HIPProbVector.from_arrays(f, cfg["u0"], cfg["p"], cfg["tspan"], dev)
        ref = deg.vectorized_asolve(probs, prob, deg.GPUTsit5(), **kw)
        order = np.arange(2000, dtype=np.int32)[::-1].copy()
        order[7], order[1500] = -5, 2003
        skipped = np.array([2000 - 1 - 7, 2000 - 1 - 1500])
        print(f"[gpu] order with out-of-range entries persistent={persistent}", flush=True)
        got = deg.vectorized_asolve(probs, prob, deg.GPUTsit5(), order=torch.as_tensor(order, device="cuda"), **kw)
        torch.cuda.synchronize()
        keep = np.setdiff1d(np.arange(2000), skipped)
        for a, b in ((ref.us_soa, got.us_soa), (ref.ts, got.ts), (ref.naccept, got.naccept)):
            np.testing.assert_array_equal(a.cpu().numpy()[..., keep], b.cpu().numpy()[..., keep])


def test_cost_start_order(deg, torch_cuda):
    """deg_start_order: the key is the span over the initial step estimate h0 = 0.01 d0 / d1, the
    order a permutation in descending key buckets; persistent solves in cost order, index order and
    a reversed order agree bitwise, for ensemble sizes that leave the 8 queues unequal."""
    torch = torch_cuda
    from diffeqgpu_amd import systems

    f, n, np_, _ = systems.SYSTEMS["lorenz"]
    abstol, reltol = 1e-6, 1e-6
    for N in (1, 7, 13, 1001, 5003):
        p0, tfs = host_uneven(900 + N, N)
        tfs = (tfs / 5.0).astype(F32)
        u0 = np.zeros((3, N), F32)
        u0[0] = 1.0
        u0[1] = np.linspace(-1, 1, N, dtype=F32)
        prob = deg.ODEProblem(f, u0[:, 0], (0.0, 10.0), p0[:, 0])
        kw = dict(dt=0.01, abstol=abstol, reltol=reltol, saveat=np.array([0.0, 0.5, 1.0, 2.0], F32))
        outs = []
        print(f"[gpu] cost start order N={N}", flush=True)
        for dev, ordr in ((deg.HIPDevice(), None), (deg.HIPDevice(start_order="cost"), None),
                          (deg.HIPDevice(start_order="cost"),
                           torch.as_tensor(np.arange(N - 1, -1, -1, dtype=np.int32), device="cuda"))):
            probs = deg.HIPProbVector.from_arrays(f, u0, p0, (0.0, 10.0), dev, tfs=tfs)
            r = deg.vectorized_asolve(probs, prob, deg.GPUTsit5(), order=ordr, **kw)
            torch.cuda.synchronize()
            outs.append([x.cpu().numpy() for x in (r.ts, r.us_soa, r.retcode, r.naccept, r.nreject)])
            if dev.start_order == "cost" and ordr is None:
                assert r.order_auto is not None
                key = r.order_key.cpu().numpy()
                o = r.order_auto.cpu().numpy()
                OB.assert_descending(key, o)
                # the key restated in float64 from the inputs
                x = u0.astype(np.float64)
                s_, r_, b_ = p0.astype(np.float64)
                f0 = np.stack([s_ * (x[1] - x[0]), x[0] * (r_ - x[2]) - x[1], x[0] * x[1] - b_ * x[2]])
                sc = abstol + np.abs(x) * reltol
                d0 = np.sqrt(((x / sc) ** 2).mean(axis=0))
                d1 = np.sqrt(((f0 / sc) ** 2).mean(axis=0))
                h0 = np.where((d0 < 1e-5) | (d1 < 1e-5), 1e-6, 0.01 * d0 / d1)
                np.testing.assert_allclose(key, tfs / h0, rtol=1e-4)
            else:
                assert r.order_auto is None
        # save rows after a trajectory's own end time are not written (the reference leaves them too)
        written = np.broadcast_to((kw["saveat"][:, None] <= tfs[None, :])[:, None, :], outs[0][1].shape)
        for other in outs[1:]:
            for j, (a, b) in enumerate(zip(outs[0], other)):
                if j == 1:
                    np.testing.assert_array_equal(a[written], b[written])
                else:
                    np.testing.assert_array_equal(a, b)
        assert (outs[0][3] > 0).all()  # every trajectory was solved
    # fixed-step and plain grids take no start order
    probs = deg.HIPProbVector.from_arrays(f, u0, p0, (0.0, 10.0), deg.HIPDevice(start_order="cost"), tfs=tfs)
    assert deg.vectorized_asolve(probs, prob, deg.GPUTsit5(), dt=0.01, order=None).order_auto is not None
    probs = deg.HIPProbVector.from_arrays(f, u0, p0, (0.0, 10.0), deg.HIPDevice(), tfs=tfs)
    assert deg.vectorized_asolve(probs, prob, deg.GPUTsit5(), dt=0.01).order_auto is None
    probs = deg.HIPProbVector.from_arrays(f, u0, p0, (0.0, 10.0), deg.HIPDevice(persistent=False), tfs=tfs)
    assert deg.vectorized_asolve(probs, prob, deg.GPUTsit5(), dt=0.01).order_auto is None
