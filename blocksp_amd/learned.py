"""Training-data generation and training driver for the learned ILU0 correction.

The samples are the Krylov basis vectors q_j of FP64 ILU0-pGMRES solves
(recorded on the device by bsp_ctx_record_basis); the loss pairs them with
r_j = A q_j.  Everything runs through the C ABI on the device.
"""
from __future__ import annotations

import numpy as np

from . import (BSR, ILU0, ILU0Ex, MPNN, MPNNGraph, PREC_F64, Context, pGMRES, record_basis, recorded)


def record_krylov(ctx: Context, A: BSR, M, b, count: int, restart: int = 30, k: int = 10):
    """Run FP64 pGMRES(restart, k) with M on A x = b and keep its first `count` basis vectors.

    Returns a device array of shape (K, n*bs) and K (<= count)."""
    length = A.n_rows * A.bs
    buf = ctx.empty((count, length))
    b_dev = b if not isinstance(b, np.ndarray) else ctx.to_device(b)
    record_basis(ctx, buf, length, count, 1)
    try:
        pGMRES(ctx, A, b_dev, 1e-14, M, maxIters=count, restart=restart, k=k)
    finally:
        K = recorded(ctx)
        record_basis(ctx, None, 0, 0)
    return buf, K


def training_set(ctx: Context, problems, count: int = 250, colour_fn=None, krylov_colour=False):
    """(A, M, graph, Z, K) for each problem: M is the ILU0 whose factor the network corrects
    (elimination order from colour_fn(problem), natural when None); Z holds the recorded
    Krylov vectors of an FP64 ILU0-pGMRES solve -- with the natural-order ILU0, or with M itself
    when krylov_colour (the directions the solver with the coloured factor explores)."""
    out = []
    for p in problems:
        A = BSR.from_problem(ctx, p)
        col = colour_fn(p) if colour_fn else None
        M0 = ILU0(ctx, A) if col is None or not krylov_colour else None
        M = M0 if col is None else ILU0Ex(ctx, A, colour=col, precision=PREC_F64)
        Z, K = record_krylov(ctx, A, M if krylov_colour else M0, p.b, count)
        g = MPNNGraph(ctx, A, M)
        out.append((A, M, g, Z, K))
    return out


def train(ctx: Context, data, epochs: int = 40, batch: int = 16, lr: float = 1e-3, seed: int = 0, bs: int = 8):
    """Train a fresh network on `data` (from training_set); returns (net, per-epoch losses)."""
    net = MPNN(ctx, bs=bs, seed=seed)
    losses = net.train([d[2] for d in data], [d[3] for d in data], [d[4] for d in data], epochs, batch, lr)
    return net, losses
