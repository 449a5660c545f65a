"""Learned Newton start: sample collection, normalization and on-GPU training.

The integrator logs, for the first accepted step of every global step, the
state at the step start, log h and the converged Radau stages (Z1, Z2, Z3)
(``accelerInt_log_first_steps``).  A network N + 1 -> 128 -> 128 -> 3N is
trained on them with the library's BF16-MFMA kernels (``accelerInt_mlp_train``)
and then predicts the Newton start of each global step's first step inside the
integrator (``accelerInt_set_learned_start``); the reference's zero start
(radau2a.c:600-603) remains the fallback after any rejection of that step.

Device buffers are torch tensors (plumbing only); the arithmetic is the
library's.
"""
from __future__ import annotations

import numpy as np

from . import (PRECISION_FP64, escalated, get_counters, get_status, integrate_device, log_first_steps,
               mlp_init, mlp_set_normalization, mlp_train, reset_counters, set_learned_start, set_precision,
               synchronize)


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("the learned start needs a GPU")
    return torch


def collect(y0, pr, steps: int, dt: float, precision: int = PRECISION_FP64):
    """Integrates ``steps`` global steps of dt from y0 (column-major N x NUM)
    with sample logging on; returns (X, Z): X[i] = (y, log h), Z[i] = (Z1, Z2, Z3)."""
    torch = _torch()
    n, num = y0.shape
    cap = num * steps
    dev = torch.device("cuda:0")
    xb = torch.zeros((cap, n + 1), dtype=torch.float32, device=dev)
    zb = torch.zeros((cap, 3 * n), dtype=torch.float32, device=dev)
    count = torch.zeros(1, dtype=torch.int32, device=dev)
    y_dev = torch.from_numpy(np.ascontiguousarray(y0.T)).to(dev)
    pr_dev = torch.from_numpy(np.ascontiguousarray(pr)).to(dev)
    codes = torch.zeros(num, dtype=torch.int32, device=dev)
    torch.cuda.synchronize()
    set_precision(precision)
    log_first_steps(xb.data_ptr(), zb.data_ptr(), cap, count.data_ptr())
    try:
        integrate_device(num, 0.0, steps * dt, dt, y_dev.data_ptr(), pr_dev.data_ptr(), codes.data_ptr(), 0)
        synchronize()
    finally:
        log_first_steps(0, 0, 0, 0)
        set_precision(PRECISION_FP64)
    m = min(int(count.item()), cap)
    return xb[:m].cpu().numpy().astype(np.float64), zb[:m].cpu().numpy().astype(np.float64)


def normalization(X, Z):
    """Per-feature mean / scale and per-output scale (constant features get scale 1)."""
    mu = X.mean(0)
    sd = X.std(0)
    sd = np.where(sd > 1e-12 * np.maximum(np.abs(mu), 1.0), sd, 1.0)
    zs = np.sqrt((Z ** 2).mean(0))
    zs = np.where(zs > 0, zs, 1.0)
    return mu, sd, zs


def train(X, Z, epochs: int = 40, batch: int = 512, lr: float = 2e-3, seed: int = 0):
    """Fresh network for N = Z.shape[1] / 3, trained on (X, Z); returns the
    per-epoch mean squared error on normalized targets."""
    torch = _torch()
    n = Z.shape[1] // 3
    mu, sd, zs = normalization(X, Z)
    mlp_init(n, seed)
    mlp_set_normalization(mu, sd, zs)
    dev = torch.device("cuda:0")
    xn = torch.from_numpy(((X - mu) / sd).astype(np.float32)).to(dev).contiguous()
    tn = torch.from_numpy((Z / zs).astype(np.float32)).to(dev).contiguous()
    torch.cuda.synchronize()
    m = (X.shape[0] // batch) * batch
    return mlp_train(m, xn.data_ptr(), tn.data_ptr(), epochs, batch, lr, seed)


def integrate_ab(y0, pr, steps: int, dt: float, precision: int, learned: bool):
    """One integration from y0 with the learned start on or off: (y, codes, counters)."""
    torch = _torch()
    n, num = y0.shape
    dev = torch.device("cuda:0")
    y_dev = torch.from_numpy(np.ascontiguousarray(y0.T)).to(dev)
    pr_dev = torch.from_numpy(np.ascontiguousarray(pr)).to(dev)
    codes = torch.zeros(num, dtype=torch.int32, device=dev)
    torch.cuda.synchronize()
    set_precision(precision)
    set_learned_start(learned)
    reset_counters()
    try:
        integrate_device(num, 0.0, steps * dt, dt, y_dev.data_ptr(), pr_dev.data_ptr(), codes.data_ptr(), 0)
        synchronize()
    finally:
        set_learned_start(False)
        set_precision(PRECISION_FP64)
    return y_dev.cpu().numpy().T.copy(), codes.cpu().numpy(), get_counters(num).astype(np.int64)
