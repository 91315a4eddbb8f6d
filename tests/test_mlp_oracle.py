"""Pins for the proposer's CPU oracle (oracle/mlp_oracle.py) before it checks the HIP kernels.

  * the analytic gradient equals central finite differences of the loss (float64);
  * the first Adam step equals its closed form P - lr * g / (|g| + eps');
  * bfloat16 rounding equals torch's float32 -> bfloat16 conversion (round to nearest even);
  * target and loss on hand-computed rows, including the clamp;
  * the bf16 inference stays within bf16 rounding of the float64 forward pass.
"""
import numpy as np
import pytest

import mlp_oracle as M


def _setup(n=40, seed=0):
    import __graft_entry__ as g

    g.load_package()
    from diffeqgpu_amd import learned

    r = np.random.default_rng(seed)
    rows = r.normal(0, 1, (n, 12))
    rows[:, 8] *= 0.5
    rows[:, :6] *= np.array([3.0, 3.0, 2.0, 1.0, 2.0, 0.5])
    rows[:, :6] += np.array([-8.0, -9.0, -6.0, 1.0, 3.0, 0.0])
    norm = np.concatenate([rows[:, :6].mean(0), 1.0 / rows[:, :6].std(0)])
    P = learned.init_params(seed).astype(np.float64)
    P[M.OFF_B1:M.OFF_W2] = r.normal(0, 0.3, 32)  # non-zero biases so that every branch is exercised
    P[M.OFF_B2:M.OFF_W3] = r.normal(0, 0.3, 32)
    return rows, norm, P


def test_gradient_matches_finite_differences():
    rows, norm, P = _setup()
    args = (rows, norm, 0.2, -1.0, 6.0)
    _, g = M.loss_and_grad(P, *args)
    r = np.random.default_rng(5)
    idx = np.concatenate([r.choice(np.arange(M.OFF_B1), 12, replace=False), np.arange(M.OFF_B1, M.OFF_B1 + 4),
                          M.OFF_W2 + r.choice(1024, 12, replace=False), np.arange(M.OFF_B2, M.OFF_B2 + 4),
                          np.arange(M.OFF_W3, M.OFF_W3 + 4), [M.OFF_B3]])
    h = 1e-6
    for i in idx:
        e = np.zeros_like(P)
        e[i] = h
        fd = (M.loss_and_grad(P + e, *args)[0] - M.loss_and_grad(P - e, *args)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])), (i, fd, g[i])
    assert np.abs(g).max() > 1e-3


def test_adam_first_step_closed_form():
    rows, norm, P = _setup()
    _, g = M.loss_and_grad(P, rows, norm, 0.2, -1.0, 6.0)
    P1, m, v = M.adam_step(P, np.zeros_like(P), np.zeros_like(P), g, 1, 1e-3)
    # bias-corrected first moments are g and g^2: the step is lr * g / (|g| + eps)
    np.testing.assert_allclose(P1, P - 1e-3 * g / (np.abs(g) + 1e-8), rtol=1e-12, atol=1e-15)


def test_bf16_rounding_matches_torch():
    import torch

    r = np.random.default_rng(1)
    x = np.concatenate([r.normal(0, 10, 5000), r.normal(0, 1e-3, 1000), [0.0, -0.0, 1.0, 1.00390625, 1.01171875],
                        np.float32(1.0) + np.arange(1, 300) * np.float32(2 ** -16)]).astype(np.float32)
    want = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    np.testing.assert_array_equal(M.bf16_rne(x), want)


def test_target_and_loss_by_hand():
    rows = np.zeros((3, 12))
    rows[0, 6], rows[0, 7] = 0.5, -3.0     # y = 0.5 + 0.2 * (-1 - -3) = 0.9
    rows[1, 6], rows[1, 7] = -1.0, 2.0     # y = -1 + 0.2 * (-1 - 2) = -1.6
    rows[2, 6], rows[2, 7] = 4.0, -60.0    # y = 4 + 0.2 * 59 = 15.8 -> clamped to 6
    np.testing.assert_allclose(M.target(rows, 0.2, -1.0, 6.0), [0.9, -1.6, 6.0])
    shifted = rows.copy()
    shifted[:, 8] = [0.25, -0.5, 1.0]      # the target is the correction to the PI factor z_PI
    np.testing.assert_allclose(M.target(shifted, 0.2, -1.0, 6.0), [0.65, -1.1, 5.0])
    P = np.zeros(M.NPARAM)
    P[M.OFF_B3] = 0.25                     # the network outputs 0.25 everywhere
    loss, _ = M.loss_and_grad(P, rows, np.concatenate([np.zeros(6), np.ones(6)]), 0.2, -1.0, 6.0)
    assert loss == pytest.approx(((0.25 - 0.9) ** 2 + (0.25 + 1.6) ** 2 + (0.25 - 6.0) ** 2) / 3)


def test_bf16_inference_near_fp64_forward():
    rows, norm, P = _setup(n=500)
    z64, _ = M.forward(P, rows, norm)
    zb = M.infer_bf16(P, rows, norm, -50.0, 50.0)
    scale = max(1.0, np.abs(z64).max())
    assert np.abs(zb - z64).max() < 0.05 * scale
    assert np.median(np.abs(zb - z64)) < 0.01 * scale
