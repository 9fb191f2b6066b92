"""Error paths and degenerate systems on the device: iteration limits (return code 1, the
iteration count the oracle reports for the same limit), a non-finite right-hand side
(BSP_E_NUMERIC, no hang), exact Krylov breakdown on A = I (one iteration, x = b), singular
diagonal blocks for the Diag and SOR preconditioners (BSP_E_NUMERIC at creation), and a zero
right-hand side (x = 0 at once)."""
import numpy as np
import pytest

import blocksp_amd as bsp
import pyoracle as po
import zoracle as zo
from blocksp_amd import problems as pr

pytestmark = pytest.mark.gpu


def identity(n, bs):
    rp = np.arange(n + 1, dtype=np.int32)
    ci = np.arange(n, dtype=np.int32)
    v = np.repeat(np.eye(bs)[None], n, axis=0)
    return rp, ci, v


@pytest.mark.parametrize("restart,k", [(-1, -1), (10, 4)])
def test_pgmres_iteration_limit(ctx, restart, k):
    print(f"pgmres iteration limit restart {restart}", flush=True)
    p = pr.dg2d_veclap(12, 10, seed=1)
    A = bsp.BSR.from_problem(ctx, p)
    M = bsp.ILU0(ctx, A)
    xd, st, rc = bsp.pGMRES(ctx, A, ctx.to_device(p.b), 1e-14, M, maxIters=17, restart=restart, k=k)
    xo, sto = po.pgmres(p, p.b, 1e-14, po.PRE_ILU0, maxIters=17, restart=restart, k=k)
    assert rc == bsp.NOT_CONVERGED and st.converged == 0
    assert st.iters == sto.iters            # the reference stops once totIters > maxIters
    assert np.linalg.norm(xd.numpy() - xo) <= 1e-8 * np.linalg.norm(xo)


def test_pcg_iteration_limit(ctx):
    p = pr.dg2d_veclap(12, 10, seed=1)
    A = bsp.BSR.from_problem(ctx, p)
    M = bsp.IChol0(ctx, A)
    _, st, rc = bsp.pCG(ctx, A, ctx.to_device(p.b), 1e-14, M, maxIters=9)
    _, sto = po.pcg(p, p.b, 1e-14, po.PRE_ICHOL0, maxIters=9)
    assert rc == bsp.NOT_CONVERGED and st.converged == 0
    assert st.iters == sto.iters


@pytest.mark.parametrize("solver", ["pgmres", "pcg", "zpgmres"])
def test_non_finite_rhs_is_an_error(ctx, solver):
    print(f"non-finite rhs {solver}", flush=True)
    p = pr.dg2d_veclap(6, 5, seed=1)
    b = p.b.copy()
    b[7] = np.nan
    if solver == "zpgmres":
        rp, ci, v = zo.complexify(p)
        A = bsp.BSR.from_complex(ctx, p.n, p.n, p.bs, rp, ci, v)
        bd = ctx.to_device(b.astype(np.complex128))
    else:
        A = bsp.BSR.from_problem(ctx, p)
        bd = ctx.to_device(b)
    M = bsp.ILU0(ctx, A)
    with pytest.raises(bsp.BspError) as ei:
        if solver == "pcg":
            bsp.pCG(ctx, A, bd, 1e-10, M)
        else:
            bsp.pGMRES(ctx, A, bd, 1e-10, M, restart=10, k=3)
    assert ei.value.code == bsp.E_NUMERIC


@pytest.mark.parametrize("bs", [1, 3, 8])
def test_identity_breaks_down_exactly(ctx, bs):
    """A = I: the first Arnoldi vector spans the solution, h_21 = 0 (the reference's
    zero_thresh breakdown, solvers.hpp:133) -- one iteration, x = b; pCG likewise."""
    n = 40
    rp, ci, v = identity(n, bs)
    A = bsp.BSR(ctx, n, n, bs, rp, ci, v)
    b = np.random.default_rng(bs).standard_normal(n * bs)
    xd, st, rc = bsp.pGMRES(ctx, A, ctx.to_device(b), 1e-10, bsp.NoPre(ctx, A), maxIters=50)
    q = pr.Problem(n, bs, bs, 1, rp, ci, v, b)
    _, sto = po.pgmres(q, b, 1e-10, po.PRE_NONE, maxIters=50)
    assert rc == 0 and st.iters == sto.iters == 1
    np.testing.assert_allclose(xd.numpy(), b, rtol=1e-15, atol=0)
    xc, stc, rcc = bsp.pCG(ctx, A, ctx.to_device(b), 1e-10, bsp.NoPre(ctx, A), maxIters=50)
    assert rcc == 0 and stc.iters <= 1
    np.testing.assert_allclose(xc.numpy(), b, rtol=1e-15, atol=0)


def test_zero_rhs(ctx):
    p = pr.dg2d_veclap(6, 5, seed=1)
    A = bsp.BSR.from_problem(ctx, p)
    z = ctx.zeros(p.n * p.bs)
    x, st, rc = bsp.pGMRES(ctx, A, z, 1e-10, bsp.ILU0(ctx, A), restart=10, k=3)
    assert rc == 0 and st.iters == 0 and not np.any(x.numpy())
    x, st, rc = bsp.pCG(ctx, A, z, 1e-10, bsp.IChol0(ctx, A))
    assert rc == 0 and not np.any(x.numpy())


@pytest.mark.parametrize("kind", ["diag", "sor"])
def test_singular_diagonal_block(ctx, kind):
    p = pr.dg2d_veclap(4, 4, seed=1)
    vals = p.vals.copy()
    for q in range(p.rowptr[3], p.rowptr[4]):
        if p.colidx[q] == 3:
            vals[q][:, 2] = 0.0                   # a zero column: the block is singular
    A = bsp.BSR(ctx, p.n, p.n, p.bs, p.rowptr, p.colidx, vals)
    with pytest.raises(bsp.BspError) as ei:
        if kind == "diag":
            bsp.Diag(ctx, A)
        else:
            bsp.SOR(ctx, A, 1.0, 1, 0)
    assert ei.value.code == bsp.E_NUMERIC
