"""Allocation-free Gram-Schmidt step (bsp_orthogonalize_noalloc) against bsp_orthogonalize.

The variant runs the same kernels on the library's static device scratch; its coefficients,
norm and new basis vector must equal the allocating entry point's bit for bit, on two calls in
a row (the second with a smaller basis and shorter vectors, over the first call's leftover
scratch), also when the second pass is taken, and m above 64 is rejected.
"""
import numpy as np
import pytest

import blocksp_amd as bsp

pytestmark = pytest.mark.gpu


def _orthonormal(n, m, seed):
    q, _ = np.linalg.qr(np.random.default_rng(seed).standard_normal((n, m)))
    return q.T.copy()


def _both(ctx, Qh, w):
    n, m = w.size, Qh.shape[0]
    Q = [ctx.to_device(q) for q in Qh]
    w1, q1 = ctx.to_device(w), ctx.zeros(n)
    h1 = bsp.orthogonalize(ctx, Q, w1, q1)
    w2, q2 = ctx.to_device(w), ctx.zeros(n)
    hd = ctx.zeros(m + 2)
    bsp.orthogonalize_noalloc(ctx, Q, w2, hd, q2)
    ctx.sync()
    h2 = hd.numpy()
    assert np.array_equal(h2[:m + 1], h1), (h2[:m + 1] - h1)
    assert np.array_equal(q2.numpy(), q1.numpy())
    assert np.array_equal(w2.numpy(), w1.numpy())
    return h2


def test_noalloc_matches_allocating_entry_point_twice(ctx):
    # first call: 40 vectors of 300000 entries; second: 7 vectors of 5000 entries
    Qa = _orthonormal(300000, 40, 1)
    wa = np.random.default_rng(2).standard_normal(300000)
    _both(ctx, Qa, wa)
    Qb = _orthonormal(5000, 7, 3)
    wb = np.random.default_rng(4).standard_normal(5000)
    h = _both(ctx, Qb, wb)
    assert h[8] == 0.0                      # no second pass for a random w
    np.testing.assert_allclose(h[:7], Qb @ wb, rtol=0, atol=1e-12 * np.linalg.norm(wb))


def test_noalloc_second_pass_matches(ctx):
    # w almost in span(Q): the second pass runs (h[m + 1] = 1) in both entry points
    n, m = 20000, 12
    Qh = _orthonormal(n, m, 5)
    rng = np.random.default_rng(6)
    w = Qh.T @ rng.standard_normal(m) + 1e-9 * rng.standard_normal(n)
    h = _both(ctx, Qh, w)
    assert h[m + 1] == 1.0


def test_noalloc_rejects_too_many_vectors(ctx):
    n = 256
    Q = [ctx.zeros(n) for _ in range(65)]
    with pytest.raises(bsp.BspError) as ei:
        bsp.orthogonalize_noalloc(ctx, Q, ctx.zeros(n), ctx.zeros(67), ctx.zeros(n))
    assert ei.value.code == bsp.E_INVALID and "64" in str(ei.value)
