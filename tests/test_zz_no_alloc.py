"""GPU: deg_start_order_noalloc, the allocation-free variant of deg_start_order.

Bars (index work, so exact):
  * the order is a permutation of 0..n_traj-1;
  * its sequence of key buckets equals that of deg_start_order's order on the same arguments, entry
    by entry, and every bucket holds the same trajectories (the order within a bucket follows the
    sort's atomics in both, so it is not compared);
  * a solve started in the no-alloc order equals the solve started in deg_start_order's order bit for bit;
  * two calls in a row, the second with fewer trajectories, both meet these bars (the static scratch keeps
    the first call's contents and its counters are cleared by a kernel in every call);
  * more than DEG_NOALLOC_MAX_TRAJ trajectories are rejected before any device work.
"""
import ctypes

import _orderbuckets as OB
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
F32 = np.float32
NOALLOC_MAX_TRAJ = 4194304


@pytest.fixture(scope="module")
def torch_cuda():
    import torch

    if not torch.cuda.is_available():
        pytest.fail("no HIP device visible: the GPU tests need an MI355X")
    return torch


def _buckets(key, order):
    return OB.buckets(key)[order]


def _check_same_order(key, ref, got):
    n = ref.size
    assert np.array_equal(np.sort(got), np.arange(n))
    bref, bgot = _buckets(key, ref), _buckets(key, got)
    np.testing.assert_array_equal(bref, bgot)
    assert np.all(np.diff(bref) <= 0)
    # same trajectories in every bucket: sorting (bucket, id) pairs makes both orders canonical
    canon = lambda o, b: o[np.lexsort((o, -b))]
    np.testing.assert_array_equal(canon(ref, bref), canon(got, bgot))


def test_start_order_noalloc_matches_start_order(deg, torch_cuda):
    torch = torch_cuda
    from diffeqgpu_amd import _lib, rng, systems

    f, n, np_, _ = systems.SYSTEMS["lorenz"]
    N = 20_000
    U = rng.uniform_table(4242, 4, N)
    p0 = np.stack([F32(5.0) + F32(10.0) * U[0], F32(60.0) * U[1], F32(1.0) + F32(3.0) * U[2]]).astype(F32)
    tfs = (F32(2.0) + F32(8.0) * U[3]).astype(F32)  # every save row (up to t = 2) is written
    u0 = np.zeros((3, N), F32)
    u0[0] = 1.0
    u0[1] = np.linspace(-1, 1, N, dtype=F32)
    prob = deg.ODEProblem(f, u0[:, 0], (0.0, 10.0), p0[:, 0])
    kw = dict(dt=0.01, abstol=1e-6, reltol=1e-6, saveat=np.array([0.0, 0.5, 1.0, 2.0], F32))
    dev = deg.HIPDevice(start_order="cost")
    probs = deg.HIPProbVector.from_arrays(f, u0, p0, (0.0, 10.0), dev, tfs=tfs)
    print(f"[gpu] no-alloc start order: reference solve N={N}", flush=True)
    res = deg.vectorized_asolve(probs, prob, deg.GPUTsit5(), **kw)
    torch.cuda.synchronize()
    kern, args = res.kernel, res.args

    for n_call in (N, 7_001):  # the second call is smaller
        print(f"[gpu] no-alloc start order: n_traj={n_call}", flush=True)
        a = _lib.SolveArgs()
        ctypes.pointer(a)[0] = args
        a.n_traj = n_call
        a.stride = N
        key = torch.empty(n_call, dtype=torch.float32, device="cuda")
        ref = torch.empty(n_call, dtype=torch.int32, device="cuda")
        work = torch.empty(_lib.ORDER_WORK_INTS, dtype=torch.int32, device="cuda")
        got = torch.full((n_call,), -1, dtype=torch.int32, device="cuda")
        kern.start_order(a, key.data_ptr(), ref.data_ptr(), work.data_ptr())
        kern.start_order_noalloc(a, got.data_ptr())
        torch.cuda.synchronize()
        _check_same_order(key.cpu().numpy(), ref.cpu().numpy(), got.cpu().numpy())

        # the solver's results do not depend on the start order: both orders give the same bits
        outs = []
        for o in (ref, got):
            pr = deg.HIPProbVector.from_arrays(f, u0[:, :n_call], p0[:, :n_call], (0.0, 10.0), deg.HIPDevice(),
                                               tfs=tfs[:n_call])
            r = deg.vectorized_asolve(pr, prob, deg.GPUTsit5(), order=o, **kw)
            torch.cuda.synchronize()
            outs.append([x.cpu().numpy() for x in (r.ts, r.us_soa, r.retcode, r.naccept, r.nreject)])
        for x, y in zip(*outs):
            np.testing.assert_array_equal(x, y)

    # beyond the static scratch: rejected before any device work
    a = _lib.SolveArgs()
    ctypes.pointer(a)[0] = args
    a.n_traj = NOALLOC_MAX_TRAJ + 1
    a.stride = NOALLOC_MAX_TRAJ + 1
    with pytest.raises(_lib.DegError, match="DEG_NOALLOC_MAX_TRAJ"):
        kern.start_order_noalloc(a, 0)
    assert _lib.lib().deg_last_error().decode().startswith("deg_start_order_noalloc takes at most")
