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
  * the learned solve: every trajectory succeeds and the FP64-oracle error stays within a
    stated factor of the PI controller's. The test prints what the trained proposer gains
    over the PI controller (attempts, rejections, kernel time) on held-out trajectories.
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


def test_learned_solve_vs_pi(deg, torch_cuda):
    """Train on the solver's own step logs (VdP, 8,192 trajectories), then solve 16,384 held-out
    trajectories with the PI controller and with the proposer."""
    torch = torch_cuda
    from diffeqgpu_amd import learned, systems

    f, n, np_, _ = systems.SYSTEMS["vdp"]
    train = _vdp(8192, 1001)
    test = _vdp(16384, 505)
    dev = deg.HIPDevice()
    kw = dict(dt=0.01, abstol=1e-6, reltol=1e-6)
    prob = deg.ODEProblem(f, test["u0"][:, 0], test["tspan"], test["p"][:, 0])
    print("[gpu] learned: collect logs", flush=True)
    probs_tr = deg.HIPProbVector.from_arrays(f, train["u0"], train["p"], train["tspan"], dev)
    rows, _ = learned.collect_steplog(probs_tr, prob, deg.GPUTsit5(), 4_000_000, **kw)
    assert rows.shape[0] > 1_000_000
    tr = learned.ProposerTrainer(rows, expo=0.2, rho=0.5, seed=0)
    loss0 = tr.evaluate()
    print("[gpu] learned: train", flush=True)
    tr.train(600, batch=16384, lr=3e-3)
    loss1 = tr.evaluate()
    assert loss1 < 0.5 * loss0, (loss0, loss1)
    packed = tr.pack()

    probs = deg.HIPProbVector.from_arrays(f, test["u0"], test["p"], test["tspan"], dev)
    out = {}
    for name, mlp in (("pi", None), ("learned", packed)):
        res = deg.vectorized_asolve(probs, prob, deg.GPUTsit5(), mlp=mlp, **kw)
        torch.cuda.synchronize()
        ms = []
        for _ in range(3):
            r2 = deg.vectorized_asolve(probs, prob, deg.GPUTsit5(), mlp=mlp, timing=True, **kw)
            ms.append(r2.kernel_ms)
        rc = res.retcode.cpu().numpy()
        assert np.all(rc == 0), np.bincount(rc)
        out[name] = dict(acc=int(res.naccept.sum().item()), rej=int(res.nreject.sum().item()), ms=min(ms),
                         u=res.us_soa[-1, :, :300].cpu().numpy().astype(np.float64))
    ref, st, _ = pyoracle.vern9("vdp", test["u0"][:, :300], test["p"][:, :300], test["tspan"])
    for name in out:
        e = np.abs(out[name]["u"] - ref[0]).max(axis=0)
        out[name]["err_median"] = float(np.median(e))
        out[name]["err_p99"] = float(np.percentile(e, 99))
    pi, ml = out["pi"], out["learned"]
    print(f"[learned vs PI] attempts {ml['acc'] + ml['rej']} vs {pi['acc'] + pi['rej']}, rejections {ml['rej']} vs "
          f"{pi['rej']}, kernel {ml['ms']:.3f} ms vs {pi['ms']:.3f} ms, FP64 error median {ml['err_median']:.3g} vs "
          f"{pi['err_median']:.3g}, p99 {ml['err_p99']:.3g} vs {pi['err_p99']:.3g}; training loss {loss0:.3f} -> "
          f"{loss1:.3f}", flush=True)
    assert ml["err_median"] <= 4.0 * pi["err_median"] + 1e-6
