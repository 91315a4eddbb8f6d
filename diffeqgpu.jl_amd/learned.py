"""The learned step-size proposer: step-log collection, GPU training and the learned solve.

Pipeline (all device work is the library's own HIP kernels; torch only holds the buffers):
  1. ``collect_steplog`` runs an adaptive solve compiled with DEG_F_STEPLOG. The solver writes
     one row per attempt that has a successor: six features at attempt n, then
     log2(dt_{n+1}/dt_n), log2 EEst_{n+1}, the PI controller's log2 factor z_PI at n, and
     whether attempt n was accepted.
  2. ``ProposerTrainer`` fits the 6 -> 32 -> 32 -> 1 MLP with ``deg_mlp_grad`` and
     ``deg_mlp_adam``. The target is the correction to z_PI that would have hit a target error
     ratio rho: y = log2(dt_{n+1}/dt_n) + expo * (log2 rho - log2 EEst_{n+1}) - z_PI.
     ``expo`` is the inverse order of the error estimate's leading term: 1/5 for Tsit5, 1/7 for Vern7.
     A network that outputs zero reproduces the PI controller.
  3. ``ProposerTrainer.pack`` converts the FP32 master weights into the BF16 matrix-core
     layout. ``vectorized_asolve(..., mlp=packed)`` then solves with DEG_F_LEARNED: after every
     attempt the next step is dt * 2^clamp(z_PI + z, log2 qmin, log2 qmax), and the embedded
     error estimate still accepts or rejects every step. The plain PI factor is the fallback
     for non-finite proposals, and after a rejection the proposal is capped at 0.9 dt.
The reference has no learned controller (its only one is the PI controller of
integrator_utils.jl:1-11 and perform_step/gpu_*_perform_step.jl).
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from . import _lib
from .algorithms import GPURosenbrock23, GPUTsit5, GPUVern7

ERR_EXPONENT = {GPUTsit5: 1.0 / 5.0, GPUVern7: 1.0 / 7.0}
ZLO, ZHI = -1.5, 1.5   # range of the correction to the PI controller's log2 factor


def _vp(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _stream(stream):
    import torch

    return ctypes.c_void_p((stream or torch.cuda.current_stream()).cuda_stream)


def collect_steplog(probs, prob, alg, cap: int, **solve_kw):
    """Solve once with step logging; returns (rows (count, 8) device tensor, solve result)."""
    import torch

    from .lowerlevel import vectorized_asolve

    if isinstance(alg, GPURosenbrock23):
        raise NotImplementedError("step logs are written by the GPUTsit5 and GPUVern7 kernels")
    dev = probs.u0.device
    buf = torch.empty((cap, _lib.STEPLOG_COLS), dtype=torch.float32, device=dev)
    cnt = torch.zeros(1, dtype=torch.int32, device=dev)
    res = vectorized_asolve(probs, prob, alg, steplog=(buf, cnt), **solve_kw)
    torch.cuda.synchronize(dev)
    n = min(int(cnt.item()), cap)
    return buf[:n], res


def init_params(seed: int) -> np.ndarray:
    """Initial FP32 master parameters: He-normal hidden layers, a small output layer, zero
    biases. Layout W1[6][32], b1[32], W2[32][32], b2[32], w3[32], b3."""
    r = np.random.default_rng(seed)
    W1 = r.normal(0.0, np.sqrt(2.0 / 6), (6, 32))
    W2 = r.normal(0.0, np.sqrt(2.0 / 32), (32, 32))
    w3 = r.normal(0.0, 0.1 / np.sqrt(32), 32)
    return np.concatenate([W1.reshape(-1), np.zeros(32), W2.reshape(-1), np.zeros(32), w3, [0.0]]).astype(np.float32)


def feature_norm(rows) -> np.ndarray:
    """[mean[6], 1/std[6]] of the feature columns (float32), computed on the host from rows."""
    x = rows[:, :6].double()
    mean = x.mean(dim=0)
    std = x.std(dim=0).clamp_min(1e-3)
    return np.concatenate([mean.cpu().numpy(), (1.0 / std).cpu().numpy()]).astype(np.float32)


class ProposerTrainer:
    """Trains the proposer on device-resident step-log rows with the library's HIP kernels."""

    def __init__(self, rows, expo: float, rho: float = 0.5, ycap: float = 6.0, seed: int = 0, nblocks: int = 256,
                 params=None, norm=None):
        import torch

        self.rows = rows.contiguous()
        self.device = self.rows.device
        self.n = int(self.rows.shape[0])
        if self.n < 1:
            raise ValueError("no step-log rows")
        self.expo, self.log2_rho, self.ycap = float(expo), float(math.log2(rho)), float(ycap)
        self.nblocks = int(nblocks)
        L = _lib.lib()
        self.norm = torch.as_tensor(norm if norm is not None else feature_norm(self.rows), dtype=torch.float32,
                                    device=self.device)
        p0 = params if params is not None else init_params(seed)
        self.params = torch.as_tensor(np.asarray(p0, np.float32), device=self.device).contiguous()
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self.partial = torch.empty((self.nblocks, L.deg_mlp_partial_cols()), dtype=torch.float32, device=self.device)
        self.loss = torch.zeros(1, dtype=torch.float32, device=self.device)
        self.step_count = 0
        self.gen = torch.Generator(device=self.device)
        self.gen.manual_seed(seed)

    def grad(self, idx=None, nrows=None, stream=None):
        """Partial gradients of the batch (rows gathered through idx, or the first nrows)."""
        n = int(idx.numel()) if idx is not None else int(nrows or self.n)
        _lib.check(_lib.lib().deg_mlp_grad(_vp(self.rows), _vp(idx), n, _vp(self.params), _vp(self.norm), self.expo,
                                           self.log2_rho, self.ycap, _vp(self.partial), self.nblocks,
                                           _stream(stream)), "deg_mlp_grad")

    def adam(self, lr: float, beta1=0.9, beta2=0.999, eps=1e-8, stream=None):
        self.step_count += 1
        _lib.check(_lib.lib().deg_mlp_adam(_vp(self.params), _vp(self.m), _vp(self.v), _vp(self.partial),
                                           self.nblocks, self.step_count, lr, beta1, beta2, eps, _vp(self.loss),
                                           _stream(stream)), "deg_mlp_adam")

    def train(self, steps: int, batch: int = 16384, lr: float = 3e-3, lr_final: float = 3e-4, report_every=0):
        """Minibatch Adam with a cosine learning-rate decay; returns the loss history."""
        import torch

        hist = []
        for s in range(steps):
            idx = torch.randint(0, self.n, (batch,), generator=self.gen, device=self.device, dtype=torch.int32)
            lr_s = lr_final + 0.5 * (lr - lr_final) * (1.0 + math.cos(math.pi * s / max(1, steps - 1)))
            self.grad(idx)
            self.adam(lr_s)
            if report_every and (s % report_every == 0 or s == steps - 1):
                hist.append((s, float(self.loss.item())))
        return hist

    def evaluate(self, rows=None) -> float:
        """Mean squared error over all rows (or the given rows) with the FP32 forward pass."""
        import torch

        r = (rows if rows is not None else self.rows).contiguous()
        z = self.forward(r)
        y = (r[:, 6] + self.expo * (self.log2_rho - r[:, 7])).clamp(-self.ycap, self.ycap) - r[:, 8]
        return float(((z - y) ** 2).mean().item())

    def forward(self, x):
        import torch

        x = x.contiguous()
        out = torch.empty(x.shape[0], dtype=torch.float32, device=self.device)
        _lib.check(_lib.lib().deg_mlp_forward(_vp(self.params), _vp(self.norm), _vp(x), int(x.shape[1]),
                                              int(x.shape[0]), _vp(out), _stream(None)), "deg_mlp_forward")
        return out

    def pack(self, zlo: float = ZLO, zhi: float = ZHI):
        """BF16 matrix-core weights for the solver (a uint8 device tensor)."""
        import torch

        packed = torch.zeros(_lib.lib().deg_mlp_packed_size(), dtype=torch.uint8, device=self.device)
        _lib.check(_lib.lib().deg_mlp_pack(_vp(self.params), _vp(self.norm), zlo, zhi, _vp(packed), _stream(None)),
                   "deg_mlp_pack")
        return packed


def infer(packed, x):
    """The solver's matrix-core evaluation over feature rows x (n, >= 6)."""
    import torch

    x = x.contiguous()
    out = torch.empty(x.shape[0], dtype=torch.float32, device=x.device)
    _lib.check(_lib.lib().deg_mlp_infer(_vp(packed), _vp(x), int(x.shape[1]), int(x.shape[0]), _vp(out),
                                        _stream(None)), "deg_mlp_infer")
    return out
