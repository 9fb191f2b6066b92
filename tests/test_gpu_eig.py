"""Ritz-value diagnostics: device Arnoldi against numpy restatements."""
import numpy as np
import pytest

import blocksp_amd as bsp
from blocksp_amd import problems as pr
import pyoracle as po

pytestmark = pytest.mark.gpu


def _nearest_dist(a, b):
    return max(np.min(np.abs(b - x)) for x in a)


def _arnoldi_numpy(Aop, n, iters):
    Q = [np.eye(n)[0]]
    H = np.zeros((iters + 1, iters))
    for j in range(iters):
        v = Aop(Q[j])
        for i in range(j + 1):
            H[i, j] = Q[i] @ v
            v = v - H[i, j] * Q[i]
        H[j + 1, j] = np.linalg.norm(v)
        Q.append(v / H[j + 1, j])
    return H


def _harmonic(H):
    n = H.shape[1]
    Hn = H[:n, :n].copy()
    w = np.linalg.solve(Hn.T, np.eye(n)[n - 1])
    Hn[:, n - 1] += H[n, n - 1] ** 2 * w
    return np.linalg.eigvals(Hn)


def test_extreme_ritz_value_converges_and_matches_numpy_arnoldi(ctx):
    # (a full-space run is not a test of anything: modified Gram-Schmidt loses orthogonality
    # over 128 steps and the Ritz values then include spurious copies)
    p = pr.stencil2d_b2(8, 8, seed=0)
    A = bsp.BSR.from_problem(ctx, p)
    Ad = pr.to_dense(p)
    ev = bsp.arnoldi_eigenvalues(ctx, A, iters=30)
    lam = np.linalg.eigvals(Ad)
    assert abs(np.abs(ev).max() - np.abs(lam).max()) <= 1e-4 * np.abs(lam).max()   # converging
    ref = np.linalg.eigvals(_arnoldi_numpy(lambda v: Ad @ v, p.n * p.bs, 30)[:30, :30])
    scale = np.abs(ref).max()
    assert _nearest_dist(ref, ev) <= 1e-7 * scale and _nearest_dist(ev, ref) <= 1e-7 * scale


@pytest.mark.parametrize("harmonic", [False, True])
def test_partial_arnoldi_matches_numpy_with_preconditioner(ctx, harmonic):
    # block Jacobi: (ILU0 would give M^{-1} A e_0 = e_0 -- ILU0 is exact on the first column --
    # and Arnoldi from e_0 stops after one step, in the reference as here)
    p = pr.stencil2d_b2(10, 10, seed=3)
    A = bsp.BSR.from_problem(ctx, p)
    M = bsp.Diag(ctx, A)
    Ad = pr.to_dense(p)
    bs = p.bs
    Dinv = np.zeros_like(Ad)
    for i in range(p.n):
        sl = slice(i * bs, (i + 1) * bs)
        Dinv[sl, sl] = np.linalg.inv(Ad[sl, sl])
    op = lambda v: Dinv @ (Ad @ v)
    H = _arnoldi_numpy(op, p.n * p.bs, 15)
    ref = _harmonic(H) if harmonic else np.linalg.eigvals(H[:15, :15])
    ev = bsp.arnoldi_eigenvalues(ctx, A, iters=15, M=M, harmonic=harmonic)
    assert len(ev) == 15
    scale = np.abs(ref).max()
    assert _nearest_dist(ref, ev) <= 1e-8 * scale and _nearest_dist(ev, ref) <= 1e-8 * scale


def test_pgmres_ritz_values_after_solves(ctx):
    p = pr.stencil2d_b2(12, 12, seed=5)
    A = bsp.BSR.from_problem(ctx, p)
    M = bsp.ILU0(ctx, A)
    b = ctx.to_device(p.b)
    lu = po.ilu0_factor(p)
    Ad = pr.to_dense(p)
    MA = np.stack([po.ilu0_apply(p, *lu, Ad[:, j]) for j in range(Ad.shape[1])], axis=1)
    lam = np.linalg.eigvals(MA)
    # unrestarted: Ritz values of M^{-1}A; the extreme one converges first
    bsp.pGMRES(ctx, A, b, 1e-12, M, maxIters=500)
    ev, harm = bsp.pgmres_ritz(ctx)
    assert not harm and len(ev) > 5
    assert abs(np.abs(ev).max() - np.abs(lam).max()) <= 1e-2 * np.abs(lam).max()
    # restarted with deflation: harmonic Ritz values, which approximate the smallest eigenvalues
    bsp.pGMRES(ctx, A, b, 1e-12, M, maxIters=500, restart=10, k=4)
    ev, harm = bsp.pgmres_ritz(ctx)
    assert harm
    assert np.min(np.abs(ev)) <= 1.5 * np.min(np.abs(lam))
