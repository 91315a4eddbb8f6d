"""GPU: the learned step-size proposer, through the C ABI.

Bars:
  * deg_mlp_grad (summed partial rows) equals the float64 oracle gradient to 2e-5 of its
    largest entry (FP32 accumulation over 3,000 rows); the loss to 1e-5 relative;
  * deg_mlp_adam equals the oracle's Adam step to 1e-6 relative (same gradient);
  * deg_mlp_forward equals the oracle forward pass to 1e-4;
  * deg_mlp_pack + deg_mlp_infer (the solver's matrix-core evaluation) equals the bf16
    emulation to 2e-2 everywhere and 1e-4 in the median (only the FP32 accumulation order
    and rare bf16 rounding ties differ);
  * step logs: one row per attempt with a successor; within one trajectory the rows chain
    (x1 of row j+1 is x0 of row j, x2 + col6 gives the next log2 dt, col7 is the next x0);
  * the learned solve: every trajectory succeeds; trained for a target error ratio of 0.05 the
    proposer removes at least 80 % of the PI controller's rejections on held-out trajectories, at a
    99th-percentile FP64 error no higher than PI's and a median within 10 % (the Float32 floor); the
    throughput target (1.5x PI at equal error) is an expected failure until it is met.
"""
import numpy as np
import pytest

import mlp_oracle as M
import pyoracle

pytestmark = pytest.mark.gpu
F32 = np.float32


@pytest.fixture(scope="module")
def torch_cuda():
    import torch

    if not torch.cuda.is_available():
        pytest.fail("no HIP device visible: the GPU tests need an MI355X")
    return torch


def _rows(n, seed):
    r = np.random.default_rng(seed)
    rows = r.normal(0, 1, (n, 12)).astype(F32)
    rows[:, :6] = rows[:, :6] * F32([3.0, 3.0, 2.0, 1.0, 2.0, 0.5]) + F32([-8.0, -9.0, -6.0, 1.0, 3.0, 0.0])
    rows[:, 7] = rows[:, 7] * 3 - 4
    rows[:, 8] = rows[:, 8] * 0.5
    return rows


def _trainer(deg, torch, rows, seed=0, nblocks=16):
    from diffeqgpu_amd import learned

    t = torch.as_tensor(rows, device="cuda")
    tr = learned.ProposerTrainer(t, expo=0.2, rho=0.5, seed=seed, nblocks=nblocks)
    P = learned.init_params(seed)
    r = np.random.default_rng(seed + 1)
    P[M.OFF_B1:M.OFF_W2] = r.normal(0, 0.3, 32)
    P[M.OFF_B2:M.OFF_W3] = r.normal(0, 0.3, 32)
    tr.params.copy_(torch.as_tensor(P, device="cuda"))
    return tr, P, tr.norm.cpu().numpy()


def test_mlp_grad_and_adam_match_oracle(deg, torch_cuda):
    torch = torch_cuda
    rows = _rows(3000, 7)
    tr, P, norm = _trainer(deg, torch, rows)
    print("[gpu] deg_mlp_grad 3000 rows, 16 blocks", flush=True)
    tr.grad()
    torch.cuda.synchronize()
    part = tr.partial.cpu().numpy().astype(np.float64)
    g_gpu = part[:, :M.NPARAM].sum(axis=0)
    loss_o, g_o = M.loss_and_grad(P, rows, norm, 0.2, -1.0, 6.0)
    assert np.abs(g_gpu - g_o).max() <= 2e-5 * np.abs(g_o).max(), np.abs(g_gpu - g_o).max()
    assert part[:, M.NPARAM + 1].sum() == 3000
    assert part[:, M.NPARAM].sum() / 3000 == pytest.approx(loss_o, rel=1e-5)
    # one Adam step from this gradient
    tr.adam(1e-3)
    torch.cuda.synchronize()
    P1, _, _ = M.adam_step(P.astype(np.float64), np.zeros(M.NPARAM), np.zeros(M.NPARAM), g_gpu, 1, 1e-3)
    np.testing.assert_allclose(tr.params.cpu().numpy(), P1, rtol=1e-6, atol=1e-7)
    assert tr.loss.item() == pytest.approx(loss_o, rel=1e-5)
    # gathered rows through an index array
    idx = torch.as_tensor(np.arange(2999, -1, -3, dtype=np.int32), device="cuda")
    tr.params.copy_(torch.as_tensor(P, device="cuda"))
    tr.grad(idx)
    torch.cuda.synchronize()
    g2 = tr.partial.cpu().numpy().astype(np.float64)[:, :M.NPARAM].sum(axis=0)
    _, g2o = M.loss_and_grad(P, rows[idx.cpu().numpy()], norm, 0.2, -1.0, 6.0)
    assert np.abs(g2 - g2o).max() <= 2e-5 * np.abs(g2o).max()


def test_mlp_forward_and_matrix_core_inference(deg, torch_cuda):
    torch = torch_cuda
    from diffeqgpu_amd import learned

    rows = _rows(5000, 11)
    tr, P, norm = _trainer(deg, torch, rows, seed=3)
    x = torch.as_tensor(rows, device="cuda")
    print("[gpu] deg_mlp_forward / deg_mlp_infer 5000 rows", flush=True)
    z = tr.forward(x).cpu().numpy()
    zo, _ = M.forward(P, rows, norm)
    assert np.abs(z - zo).max() < 1e-4 * max(1.0, np.abs(zo).max())
    packed = tr.pack(-50.0, 50.0)
    for n in (5000, 1, 63, 65, 257):
        zi = learned.infer(packed, x[:n]).cpu().numpy()
        ze = M.infer_bf16(P, rows[:n], norm, -50.0, 50.0)
        d = np.abs(zi - ze)
        assert d.max() < 2e-2 and np.median(d) < 1e-4, (n, d.max(), np.median(d))
    # the clamp
    packed = tr.pack(-0.5, 0.25)
    zi = learned.infer(packed, x).cpu().numpy()
    assert zi.min() >= -0.5 and zi.max() <= 0.25
    np.testing.assert_allclose(zi, np.clip(M.infer_bf16(P, rows, norm, -50, 50), -0.5, 0.25), atol=2e-2)


def _vdp(N, seed):
    from diffeqgpu_amd import workloads

    return workloads.config("vdp", N=N, seed=seed)


def test_steplog_rows_chain(deg, torch_cuda):
    torch = torch_cuda
    from diffeqgpu_amd import learned, systems

    f, n, np_, _ = systems.SYSTEMS["vdp"]
    cfg = _vdp(257, 9)
    prob = deg.ODEProblem(f, cfg["u0"][:, 0], (0.0, 10.0), cfg["p"][:, 0])
    dev = deg.HIPDevice()
    # one trajectory: its rows arrive in order
    probs = deg.HIPProbVector.from_arrays(f, cfg["u0"][:, :1], cfg["p"][:, :1], (0.0, 10.0), dev)
    print("[gpu] steplog vdp N=1", flush=True)
    rows, res = learned.collect_steplog(probs, prob, deg.GPUTsit5(), 100_000, dt=0.01, abstol=1e-6, reltol=1e-6)
    R = rows.cpu().numpy().astype(np.float64)
    att = int(res.naccept.item() + res.nreject.item())
    assert R.shape[0] == att - 1 and att > 50
    assert np.all(np.isfinite(R))
    assert R[0, 1] == pytest.approx(-13.2877123795, abs=1e-6) and R[0, 5] == 0.0
    np.testing.assert_array_equal(R[1:, 1], R[:-1, 0])
    np.testing.assert_array_equal(R[:-1, 7], R[1:, 0])
    np.testing.assert_allclose(R[:-1, 2] + R[:-1, 6], R[1:, 2], atol=1e-5)
    np.testing.assert_allclose(R[1:, 5], R[1:, 2] - R[:-1, 2], atol=1e-5)
    assert R[0, 2] == pytest.approx(np.log2(F32(0.01)), abs=1e-5)
    # z_PI and the accept flag: a rejected attempt is followed by a smaller step, and while the
    # PI controller is in charge the next step is exactly its proposal (up to the tf clamp)
    assert set(np.unique(R[:, 9])) <= {0.0, 1.0} and R[:, 9].sum() > 0
    rej = R[:, 9] == 0.0
    assert np.all(R[rej, 6] < 0.0)
    np.testing.assert_allclose(R[:-1, 6], R[:-1, 8], atol=2e-5)
    # many trajectories: the row count is the number of attempts less one per trajectory
    probs = deg.HIPProbVector.from_arrays(f, cfg["u0"], cfg["p"], (0.0, 10.0), dev)
    rows, res = learned.collect_steplog(probs, prob, deg.GPUTsit5(), 1_000_000, dt=0.01, abstol=1e-6, reltol=1e-6)
    att = (res.naccept.to(torch.int64) + res.nreject.to(torch.int64)).sum().item()
    assert rows.shape[0] == att - 257
    # a small cap: rows beyond it are dropped, nothing else changes
    rows_c, res_c = learned.collect_steplog(probs, prob, deg.GPUTsit5(), 1000, dt=0.01, abstol=1e-6, reltol=1e-6)
    assert rows_c.shape[0] == 1000
    np.testing.assert_array_equal(res_c.naccept.cpu().numpy(), res.naccept.cpu().numpy())


def _train_and_compare(deg, torch, rho, every, n_test=16384):
    """Train on the solver's own step logs (VdP, 8,192 trajectories), then solve held-out
    trajectories with the PI controller and with the proposer (target error ratio rho, evaluated every
    `every` iterations); FP64 errors on all of them against Vern9 at 1e-12."""
    from diffeqgpu_amd import learned, systems

    f, n, np_, _ = systems.SYSTEMS["vdp"]
    train = _vdp(8192, 1001)
    test = _vdp(n_test, 505)
    dev = deg.HIPDevice()
    kw = dict(dt=0.01, abstol=1e-6, reltol=1e-6)
    prob = deg.ODEProblem(f, test["u0"][:, 0], test["tspan"], test["p"][:, 0])
    print(f"[gpu] learned rho={rho}: collect logs", flush=True)
    probs_tr = deg.HIPProbVector.from_arrays(f, train["u0"], train["p"], train["tspan"], dev)
    rows, _ = learned.collect_steplog(probs_tr, prob, deg.GPUTsit5(), 4_000_000, **kw)
    assert rows.shape[0] > 1_000_000
    tr = learned.ProposerTrainer(rows, expo=0.2, rho=rho, seed=0)
    loss0 = tr.evaluate()
    print(f"[gpu] learned rho={rho}: train", flush=True)
    tr.train(600, batch=16384, lr=3e-3)
    loss1 = tr.evaluate()
    packed = tr.pack()
    probs = deg.HIPProbVector.from_arrays(f, test["u0"], test["p"], test["tspan"], dev)
    out = {}
    for name, mlp in (("pi", None), ("learned", packed)):
        res = deg.vectorized_asolve(probs, prob, deg.GPUTsit5(), mlp=mlp, mlp_every=every, **kw)
        torch.cuda.synchronize()
        ms = []
        for _ in range(3):
            r2 = deg.vectorized_asolve(probs, prob, deg.GPUTsit5(), mlp=mlp, mlp_every=every, timing=True, **kw)
            ms.append(r2.kernel_ms)
        rc = res.retcode.cpu().numpy()
        assert np.all(rc == 0), np.bincount(rc)
        out[name] = dict(acc=int(res.naccept.sum().item()), rej=int(res.nreject.sum().item()), ms=min(ms),
                         u=res.us_soa[-1].cpu().numpy().astype(np.float64))
    ref, st, _ = pyoracle.vern9("vdp", test["u0"], test["p"], test["tspan"])
    assert np.all(st == 0)
    for name in out:
        e = np.abs(out[name]["u"] - ref[0]).max(axis=0)
        out[name]["err_median"] = float(np.median(e))
        out[name]["err_p99"] = float(np.percentile(e, 99))
    pi, ml = out["pi"], out["learned"]
    print(f"[learned rho={rho} every={every} vs PI] attempts {ml['acc'] + ml['rej']} vs {pi['acc'] + pi['rej']}, "
          f"rejections {ml['rej']} vs {pi['rej']}, kernel {ml['ms']:.3f} ms vs {pi['ms']:.3f} ms, FP64 error median "
          f"{ml['err_median']:.3g} vs {pi['err_median']:.3g}, p99 {ml['err_p99']:.3g} vs {pi['err_p99']:.3g}; "
          f"training loss {loss0:.3f} -> {loss1:.3f}", flush=True)
    assert loss1 < 0.5 * loss0, (loss0, loss1)
    return pi, ml


def test_learned_removes_rejections(deg, torch_cuda):
    """What the trained proposer gains, stated and asserted: trained for a target error ratio of 0.05
    and evaluated on every attempt, it removes at least 80 % of the PI controller's rejected attempts on
    16,384 held-out Van der Pol trajectories (mu ~ logU(0.1, 30), Tsit5 at 1e-6), with a 99th-percentile
    FP64 error (over all of them) no higher than the PI controller's and a median within 10 % of it. The
    median error at this tolerance sits at the Float32 rounding floor of the state (about 8e-6 for
    either controller, whatever the step sequence), where runs differ by a few per cent: the proposer's
    training sees its step-log rows in the order the solver's atomics wrote them."""
    pi, ml = _train_and_compare(deg, torch_cuda, rho=0.05, every=1)
    assert ml["rej"] <= 0.2 * pi["rej"], (ml["rej"], pi["rej"])
    assert ml["err_p99"] <= pi["err_p99"], (ml["err_p99"], pi["err_p99"])
    assert ml["err_median"] <= 1.10 * pi["err_median"], (ml["err_median"], pi["err_median"])


@pytest.mark.xfail(reason="the throughput target (>= 1.5x the PI kernel's trajectories/s at equal or lower FP64 "
                          "error) is not met: the PI controller rejects about 3 % of attempts here and the proposer's "
                          "evaluation costs more than the attempts it saves", strict=False)
def test_learned_throughput_target(deg, torch_cuda):
    """The throughput target for the proposer: at least 1.5x the PI kernel's trajectories/s (kernel
    time) on held-out Van der Pol trajectories at equal or lower median and 99th-percentile FP64 error
    (the bench's setting: target ratio 0.2, evaluated every 2nd iteration)."""
    pi, ml = _train_and_compare(deg, torch_cuda, rho=0.2, every=2)
    assert ml["err_median"] <= pi["err_median"] and ml["err_p99"] <= pi["err_p99"]
    assert pi["ms"] >= 1.5 * ml["ms"], (ml["ms"], pi["ms"])
