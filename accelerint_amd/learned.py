"""Learned Newton start: sample collection, normalization and on-GPU training.

The integrator logs, for the first accepted step of every global step, the
features of the step start (1000/T, log of each mass fraction, log h, log p)
and the converged Radau stages (Z1, Z2, Z3) (``accelerInt_start_log`` /
``accelerInt_read_log``).  A network N + 2 -> 128 -> 128 -> 3N is trained on
them with the library's BF16-MFMA kernels (``accelerInt_mlp_train_host``) and
then predicts the Newton start of each global step's first step inside the
integrator (``accelerInt_set_learned_start``); the reference's zero start
(radau2a.c:600-603) takes over when the prediction's first Newton correction
is not smaller than the corrected stages, when Newton from it fails, and
after any rejection of that step.
"""
from __future__ import annotations

import numpy as np

from . import (PRECISION_FP64, get_counters, get_status, integrate, mlp_init, mlp_set_normalization,
               mlp_train_host, read_log, reset_counters, set_learned_start, set_precision, start_log)


def collect(y0, pr, steps: int, dt: float, precision: int = PRECISION_FP64):
    """Integrates ``steps`` global steps of dt from y0 (column-major N x NUM)
    with sample logging on; returns (X, Z) with X[i] = the features and
    Z[i] = (Z1, Z2, Z3) as float64 arrays."""
    n, num = y0.shape
    cap = num * steps
    y = np.ascontiguousarray(y0).copy()
    set_precision(precision)
    start_log(cap)
    try:
        integrate(num, 0.0, steps * dt, dt, y, np.ascontiguousarray(pr))
    finally:
        X, Z = read_log(cap, n)
        set_precision(PRECISION_FP64)
    # a mixed-precision pass that is later re-run with FP64 factors may have
    # logged samples before it failed; only finite samples are kept
    keep = np.isfinite(X).all(1) & np.isfinite(Z).all(1)
    return X[keep].astype(np.float64), Z[keep].astype(np.float64)


def normalization(X, Z):
    """Per-feature mean / scale and per-output RMS scale (constant features get scale 1)."""
    mu = X.mean(0)
    sd = X.std(0)
    sd = np.where(sd > 1e-12 * np.maximum(np.abs(mu), 1.0), sd, 1.0)
    zs = np.sqrt((Z ** 2).mean(0))
    zs = np.where(zs > 0, zs, 1.0)
    return mu, sd, zs


CLIP = 16.0  # clamp of normalized features, the same as the kernels' kMlpClip


def train(X, Z, epochs: int = 40, batch: int = 512, lr: float = 2e-3, seed: int = 0):
    """Fresh network for N = Z.shape[1] / 3, trained on (X, Z); returns the
    per-epoch mean squared error on normalized targets."""
    n = Z.shape[1] // 3
    mu, sd, zs = normalization(X, Z)
    mlp_init(n, seed)
    mlp_set_normalization(mu, sd, zs)
    m = (X.shape[0] // batch) * batch
    xn = np.clip((X[:m] - mu) / sd, -CLIP, CLIP).astype(np.float32)
    return mlp_train_host(xn, (Z[:m] / zs).astype(np.float32), epochs, batch, lr, seed)


def integrate_ab(y0, pr, steps: int, dt: float, precision: int, learned: bool):
    """One integration from y0 with the learned start on or off: (y, codes, counters)."""
    num = y0.shape[1]
    y = np.ascontiguousarray(y0).copy()
    set_precision(precision)
    set_learned_start(learned)
    reset_counters()
    try:
        integrate(num, 0.0, steps * dt, dt, y, np.ascontiguousarray(pr))
    finally:
        set_learned_start(False)
        set_precision(PRECISION_FP64)
    return y, get_status(num), get_counters(num).astype(np.int64)
