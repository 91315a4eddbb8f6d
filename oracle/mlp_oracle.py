"""CPU oracle of the learned step-size proposer (numpy, float64).

TEST INFRASTRUCTURE. Only the test suite loads this module, as the checker of the HIP
training kernels (deg_mlp_grad, deg_mlp_adam, deg_mlp_forward) and of the matrix-core
inference (deg_mlp_pack + deg_mlp_infer / the solver's deg_mlp_eval). The product never
calls it.

The proposer has no counterpart in the reference (its only controller is the PI controller,
integrator_utils.jl:1-11). What is restated here is the network and training procedure the
library implements:
  x' = (x - mean) * inv_std                          (6 features)
  h1 = relu(x' W1 + b1), h2 = relu(h1 W2 + b2)        (32 + 32 ReLU units)
  z  = h2 . w3 + b3
  y  = clamp(col6 + expo * (log2_rho - col7), -ycap, ycap) - col8   (target: correction to the PI
                                                               controller's log2 factor col8)
  L  = mean((z - y)^2)
  Adam (Kingma and Ba 2015) with bias correction.
Parameter vector layout (1313 floats): W1[6][32], b1[32], W2[32][32], b2[32], w3[32], b3.

The matrix-core inference rounds weights (W1, b1, W2, w3) and the layer inputs (the normalised
features, the constant 1, the ReLU outputs of both hidden layers) to bfloat16 with
round-to-nearest-even, and accumulates in FP32 (b2 and b3 are added in FP32). ``infer_bf16`` emulates that (in float64, so it differs from the device
only by FP32 accumulation rounding).

Pins (tests/test_mlp_oracle.py): the gradient against central finite differences; the Adam
step against its closed form on the first step; the bfloat16 rounding against torch's
conversion; the target and loss on hand-computed rows.
"""
from __future__ import annotations

import numpy as np

IN, HID = 6, 32
OFF_W1, OFF_B1, OFF_W2, OFF_B2, OFF_W3, OFF_B3 = 0, 192, 224, 1248, 1280, 1312
NPARAM = 1313


def unpack(P):
    P = np.asarray(P, dtype=np.float64)
    W1 = P[OFF_W1:OFF_B1].reshape(IN, HID)
    b1 = P[OFF_B1:OFF_W2]
    W2 = P[OFF_W2:OFF_B2].reshape(HID, HID)
    b2 = P[OFF_B2:OFF_W3]
    w3 = P[OFF_W3:OFF_B3]
    b3 = P[OFF_B3]
    return W1, b1, W2, b2, w3, b3


def pack(W1, b1, W2, b2, w3, b3):
    return np.concatenate([W1.reshape(-1), b1, W2.reshape(-1), b2, w3, [b3]]).astype(np.float64)


def target(rows, expo, log2_rho, ycap):
    rows = np.asarray(rows, dtype=np.float64)
    return np.clip(rows[:, 6] + expo * (log2_rho - rows[:, 7]), -ycap, ycap) - rows[:, 8]


def normalise(X, norm):
    norm = np.asarray(norm, dtype=np.float64)
    return (np.asarray(X, dtype=np.float64)[:, :IN] - norm[:IN]) * norm[IN:2 * IN]


def forward(P, X, norm):
    W1, b1, W2, b2, w3, b3 = unpack(P)
    xn = normalise(X, norm)
    a1 = xn @ W1 + b1
    h1 = np.maximum(a1, 0.0)
    a2 = h1 @ W2 + b2
    h2 = np.maximum(a2, 0.0)
    z = h2 @ w3 + b3
    return z, (xn, a1, h1, a2, h2)


def loss_and_grad(P, rows, norm, expo, log2_rho, ycap):
    """Mean squared error over the rows and its gradient with respect to P (float64)."""
    W1, b1, W2, b2, w3, b3 = unpack(P)
    y = target(rows, expo, log2_rho, ycap)
    z, (xn, a1, h1, a2, h2) = forward(P, rows, norm)
    n = len(y)
    d = z - y
    loss = float(np.mean(d * d))
    dz = 2.0 * d / n
    gw3 = h2.T @ dz
    gb3 = dz.sum()
    dh2 = np.outer(dz, w3) * (a2 > 0)
    gW2 = h1.T @ dh2
    gb2 = dh2.sum(axis=0)
    dh1 = (dh2 @ W2.T) * (a1 > 0)
    gW1 = xn.T @ dh1
    gb1 = dh1.sum(axis=0)
    return loss, pack(gW1, gb1, gW2, gb2, gw3, gb3)


def adam_step(P, m, v, g, step, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    m = beta1 * m + (1.0 - beta1) * g
    v = beta2 * v + (1.0 - beta2) * g * g
    mh = m / (1.0 - beta1 ** step)
    vh = v / (1.0 - beta2 ** step)
    return P - lr * mh / (np.sqrt(vh) + eps), m, v


def bf16_rne(x) -> np.ndarray:
    """Round float32 values to bfloat16 (nearest, ties to even); returned as float32."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = (u + np.uint64(0x7FFF) + ((u >> np.uint64(16)) & np.uint64(1))) >> np.uint64(16)
    return (r.astype(np.uint32) << np.uint32(16)).view(np.float32)


def infer_bf16(P, X, norm, zlo, zhi):
    """The matrix-core evaluation: bf16 weights (W1, b1, W2, w3) and layer inputs (features, h1, h2),
    float accumulation; b2 and b3 are added in float."""
    W1, b1, W2, b2, w3, b3 = unpack(np.asarray(P, dtype=np.float32))
    norm = np.asarray(norm, dtype=np.float32)
    X = np.asarray(X, dtype=np.float32)[:, :IN]
    shift = (-norm[:IN] * norm[IN:2 * IN]).astype(np.float32)
    # one fused multiply-add per feature, as on the device (the f64 product is exact)
    xn = (X.astype(np.float64) * norm[IN:2 * IN].astype(np.float64) + shift.astype(np.float64)).astype(np.float32)
    xb = bf16_rne(xn).astype(np.float64)
    W1b = bf16_rne(W1).astype(np.float64)
    b1b = bf16_rne(b1).astype(np.float64)
    W2b = bf16_rne(W2).astype(np.float64)
    a1 = (xb @ W1b + b1b).astype(np.float32)
    h1 = bf16_rne(np.maximum(a1, 0.0)).astype(np.float64)
    a2 = (h1 @ W2b + b2.astype(np.float64)).astype(np.float32)
    h2 = bf16_rne(np.maximum(a2, 0.0)).astype(np.float64)
    z = (h2 @ bf16_rne(w3).astype(np.float64)).astype(np.float32).astype(np.float64) + float(b3)
    return np.clip(z, zlo, zhi)
