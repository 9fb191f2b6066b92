"""Coloured / FP32 ILU0 and the mixed-precision FGMRES on the GPU.

Pins:
  * coloured ILU0 = the oracle's natural-order ILU0 of P A P^T (P: colour-major
    permutation), mapped back: relative 2-norm difference <= 1e-10 (FP64);
  * FP32 storage: within 1e-5 relative of the FP64 coloured apply;
  * parity colouring of a 2D grid gives 3-level sweeps (a closed form);
  * FGMRES_mixed reaches the true FP64 relative residual it is asked for.
"""
import numpy as np
import pytest

import blocksp_amd as bsp
import pyoracle as po
from blocksp_amd import problems as pr

pytestmark = pytest.mark.gpu


def permuted(p, perm):
    """Problem with rows/columns renumbered: new row t = old row perm[t], each row's
    blocks in ascending new column order."""
    inv = np.empty_like(perm)
    inv[perm] = np.arange(len(perm))
    rp, ci, vals = [0], [], []
    for old in perm:
        cols = p.colidx[p.rowptr[old]:p.rowptr[old + 1]]
        v = p.vals[p.rowptr[old]:p.rowptr[old + 1]]
        nc = inv[cols]
        o = np.argsort(nc, kind="stable")
        ci.extend(nc[o])
        vals.append(v[o])
        rp.append(len(ci))
    b = p.b.reshape(p.n, p.bs)[perm].reshape(-1)
    return pr.Problem(p.n, p.bs, p.B, p.D, np.array(rp, np.int32), np.array(ci, np.int32), np.concatenate(vals), b)


def test_ilu0_ex_natural_equals_ilu0(ctx):
    p = pr.dg2d_veclap(10, 9, seed=1)
    A = bsp.BSR.from_problem(ctx, p)
    r = np.random.default_rng(0).standard_normal(p.n * p.bs)
    a, b = ctx.to_device(r), ctx.to_device(r)
    bsp.ILU0(ctx, A)(a)
    bsp.ILU0Ex(ctx, A, None, bsp.PREC_F64)(b)
    assert np.array_equal(a.numpy(), b.numpy())


@pytest.mark.parametrize("case", ["dg2d", "dg3d"])
def test_coloured_ilu0_matches_oracle_on_permuted_matrix(ctx, case):
    if case == "dg2d":
        dims = (12, 10)
        p = pr.dg2d_veclap(*dims, seed=1)
    else:
        dims = (5, 4, 6)
        p = pr.dg3d_lap(*dims, seed=2)
    colour = pr.parity_colours(dims)
    A = bsp.BSR.from_problem(ctx, p)
    M = bsp.ILU0Ex(ctx, A, colour, bsp.PREC_F64)
    nl_lo, _ = M.levels(True)
    nl_up, _ = M.levels(False)
    assert nl_lo == len(dims) + 1 and nl_up == len(dims) + 1
    perm = np.argsort(colour, kind="stable")
    q = permuted(p, perm)
    lu, dlu, dpiv = po.ilu0_factor(q)
    r = np.random.default_rng(1).standard_normal(p.n * p.bs)
    zq = po.ilu0_apply(q, lu, dlu, dpiv, r.reshape(p.n, p.bs)[perm].reshape(-1))
    zref = np.empty_like(zq.reshape(p.n, p.bs))
    zref[perm] = zq.reshape(p.n, p.bs)
    rd = ctx.to_device(r)
    M(rd)
    assert np.linalg.norm(rd.numpy() - zref.reshape(-1)) <= 1e-10 * np.linalg.norm(zref)
    # FP32 storage and arithmetic
    M32 = bsp.ILU0Ex(ctx, A, colour, bsp.PREC_F32)
    rd32 = ctx.to_device(r)
    M32(rd32)
    assert np.linalg.norm(rd32.numpy() - zref.reshape(-1)) <= 1e-5 * np.linalg.norm(zref)


@pytest.mark.parametrize("spmv", [bsp.PREC_F32, bsp.PREC_F64])
@pytest.mark.parametrize("prec", [bsp.PREC_F32, bsp.PREC_F64])
def test_fgmres_mixed_reaches_fp64_residual(ctx, spmv, prec):
    print(f"fgmres spmv={spmv} precond={prec}", flush=True)
    dims = (32, 24)
    p = pr.dg2d_veclap(*dims, seed=1)
    A = bsp.BSR.from_problem(ctx, p)
    M = bsp.ILU0Ex(ctx, A, pr.parity_colours(dims), prec)
    xd, st, rc = bsp.FGMRES_mixed(ctx, A, ctx.to_device(p.b), 1e-10, M, maxIters=5000, restart=30, k=10,
                                  spmv_precision=spmv)
    x = xd.numpy()
    relres = np.linalg.norm(p.b - po.bsmv(p, x)) / np.linalg.norm(p.b)
    assert rc == 0 and st.converged == 1 and relres <= 1e-10, (rc, relres, st.iters)
    assert abs(st.true_relres - relres) <= 1e-2 * relres + 1e-16
    # forward error against the FP64 oracle solution
    xo, sto = po.pgmres(p, p.b, 1e-12, po.PRE_ILU0, maxIters=5000, restart=30, k=10)
    assert np.linalg.norm(x - xo) <= 1e-7 * np.linalg.norm(xo)


def test_fgmres_mixed_natural_fp32(ctx):
    p = pr.stencil2d_b2(24, 24, seed=0)
    A = bsp.BSR.from_problem(ctx, p)
    M = bsp.ILU0Ex(ctx, A, None, bsp.PREC_F32)
    xd, st, rc = bsp.FGMRES_mixed(ctx, A, ctx.to_device(p.b), 1e-10, M, maxIters=5000, restart=30, k=10)
    relres = np.linalg.norm(p.b - po.bsmv(p, xd.numpy())) / np.linalg.norm(p.b)
    assert rc == 0 and relres <= 1e-10


def _orthonormal(n, m, seed):
    q, _ = np.linalg.qr(np.random.default_rng(seed).standard_normal((n, m)))
    return q.T.copy()


@pytest.mark.parametrize("m", [1, 7, 8, 9, 31])
def test_orthogonalize_matches_numpy(ctx, m):
    n = 5000
    Qh = _orthonormal(n, m, m)
    w = np.random.default_rng(100 + m).standard_normal(n)
    Q = [ctx.to_device(q) for q in Qh]
    wd, qn = ctx.to_device(w), ctx.zeros(n)
    h = bsp.orthogonalize(ctx, Q, wd, qn)
    href = Qh @ w
    r = w - Qh.T @ href
    np.testing.assert_allclose(h[:m], href, rtol=0, atol=1e-12 * np.linalg.norm(w))
    assert abs(h[m] - np.linalg.norm(r)) <= 1e-12 * np.linalg.norm(w)
    np.testing.assert_allclose(qn.numpy(), r / np.linalg.norm(r), rtol=0, atol=1e-12)


def test_orthogonalize_reorthogonalises_under_cancellation(ctx):
    # w almost in span(Q): one classical pass leaves O(eps ||w|| / ||r||) components along Q;
    # the second pass (taken because ||r|| << ||w||) removes them
    n, m = 20000, 12
    Qh = _orthonormal(n, m, 3)
    rng = np.random.default_rng(4)
    w = Qh.T @ rng.standard_normal(m) + 1e-9 * rng.standard_normal(n)
    Q = [ctx.to_device(q) for q in Qh]
    wd, qn = ctx.to_device(w), ctx.zeros(n)
    h = bsp.orthogonalize(ctx, Q, wd, qn)
    q = qn.numpy()
    assert np.abs(Qh @ q).max() <= 1e-12
    assert abs(np.linalg.norm(q) - 1.0) <= 1e-12
    np.testing.assert_allclose(h[:m], Qh @ w, rtol=0, atol=1e-12 * np.linalg.norm(w))


@pytest.mark.parametrize("ortho", [0, 1])
def test_fgmres_mixed_with_both_orthogonalisations(ctx, ortho):
    p = pr.dg2d_veclap(48, 48, seed=1)
    A = bsp.BSR.from_problem(ctx, p)
    b = ctx.to_device(p.b)
    M = bsp.ILU0Ex(ctx, A, colour=pr.parity_colours((48, 48)), precision=bsp.PREC_F32)
    ctx.set_option("fgmres_ortho", ortho)
    try:
        x, st, rc = bsp.FGMRES_mixed(ctx, A, b, 1e-10, M)
    finally:
        ctx.set_option("fgmres_ortho", 0)
    assert rc == 0 and st.true_relres <= 1e-10
    assert np.linalg.norm(p.b - p.bsmv(x.numpy())) <= 1e-10 * np.linalg.norm(p.b)


@pytest.mark.parametrize("prec", [bsp.PREC_F32, bsp.PREC_F64])
def test_fgmres_mixed_fp32_basis_reaches_fp64_residual(ctx, prec):
    # FP32 storage of the Arnoldi basis (FP64 Gram-Schmidt arithmetic); the FP64 outer
    # refinement still drives the true residual to the FP64 tolerance
    dims = (48, 48)
    p = pr.dg2d_veclap(*dims, seed=1)
    A = bsp.BSR.from_problem(ctx, p)
    b = ctx.to_device(p.b)
    M = bsp.ILU0Ex(ctx, A, colour=pr.parity_colours(dims), precision=prec)
    x64, st64, rc64 = bsp.FGMRES_mixed(ctx, A, b, 1e-10, M)
    ctx.set_option("fgmres_basis_f32", 1)
    try:
        x, st, rc = bsp.FGMRES_mixed(ctx, A, b, 1e-10, M)
    finally:
        ctx.set_option("fgmres_basis_f32", 0)
    relres = np.linalg.norm(p.b - p.bsmv(x.numpy())) / np.linalg.norm(p.b)
    assert rc == 0 and st.converged == 1 and relres <= 1e-10, (rc, relres, st.iters)
    assert abs(st.true_relres - relres) <= 1e-2 * relres + 1e-16
    assert rc64 == 0 and st.iters <= 1.3 * st64.iters + 30, (st.iters, st64.iters)
    xo = x64.numpy()
    assert np.linalg.norm(x.numpy() - xo) <= 1e-7 * np.linalg.norm(xo)


def test_fgmres_basis_option_is_checked(ctx):
    with pytest.raises(Exception):
        ctx.set_option("fgmres_basis_f32", 2)


def _bf16(a):
    """FP64 -> FP32 (nearest) -> BF16 (nearest even), back to FP64: the library's rounding."""
    u = np.asarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def test_bf16_ilu0_apply_matches_the_rounded_factor(ctx):
    # natural order (the reference's ILU0, many levels): the device apply in FP32 arithmetic
    # equals forward/backward block substitution with the oracle's factor rounded to BF16
    # (strict L and U blocks, and the inverses of U's diagonal blocks)
    p = pr.dg2d_veclap(9, 7, seed=4)
    A = bsp.BSR.from_problem(ctx, p)
    lu, _, _ = po.ilu0_factor(p)
    bs, n = p.bs, p.n
    r = np.random.default_rng(5).standard_normal(n * bs)
    y = r.reshape(n, bs).copy()
    for i in range(n):
        for q in range(p.rowptr[i], p.rowptr[i + 1]):
            if p.colidx[q] < i:
                y[i] -= _bf16(lu[q]) @ y[p.colidx[q]]
    x = y.copy()
    for i in reversed(range(n)):
        d = None
        for q in range(p.rowptr[i], p.rowptr[i + 1]):
            c = p.colidx[q]
            if c > i:
                x[i] -= _bf16(lu[q]) @ x[c]
            elif c == i:
                d = _bf16(np.linalg.inv(lu[q]))
        x[i] = d @ x[i]
    M = bsp.ILU0Ex(ctx, A, None, bsp.PREC_BF16)
    rd = ctx.to_device(r)
    M(rd)
    got = rd.numpy()
    assert np.linalg.norm(got - x.reshape(-1)) <= 1e-5 * np.linalg.norm(x)
    # and it is a different operator from the FP32 one (the rounding is really BF16)
    r32 = ctx.to_device(r)
    bsp.ILU0Ex(ctx, A, None, bsp.PREC_F32)(r32)
    assert np.linalg.norm(got - r32.numpy()) >= 1e-4 * np.linalg.norm(x)


def test_fgmres_mixed_bf16_preconditioner_reaches_fp64_residual(ctx):
    dims = (48, 48)
    p = pr.dg2d_veclap(*dims, seed=1)
    A = bsp.BSR.from_problem(ctx, p)
    b = ctx.to_device(p.b)
    col = pr.parity_colours(dims)
    _, st32, rc32 = bsp.FGMRES_mixed(ctx, A, b, 1e-10, bsp.ILU0Ex(ctx, A, col, bsp.PREC_F32))
    x, st, rc = bsp.FGMRES_mixed(ctx, A, b, 1e-10, bsp.ILU0Ex(ctx, A, col, bsp.PREC_BF16))
    relres = np.linalg.norm(p.b - p.bsmv(x.numpy())) / np.linalg.norm(p.b)
    assert rc == 0 and st.converged == 1 and relres <= 1e-10, (rc, relres, st.iters)
    assert rc32 == 0 and st.iters <= 1.15 * st32.iters + 20, (st.iters, st32.iters)


def test_ilu0_ex_rejects_unknown_precision(ctx):
    p = pr.dg2d_veclap(6, 5, seed=1)
    A = bsp.BSR.from_problem(ctx, p)
    with pytest.raises(Exception):
        bsp.ILU0Ex(ctx, A, None, 7)


@pytest.mark.parametrize("spmv", [bsp.PREC_F64, bsp.PREC_F32])
def test_fgmres_mixed_second_gram_schmidt_pass_on_cancellation(ctx, spmv):
    # diagonal A whose entries form three tight clusters: from the fourth step on, A z lies almost
    # in the span of the basis, one classical pass cancels to ~1e-4 of ||w|| and the solver runs the
    # second pass (counted in stats.reorth); the solution is b / diag(A)
    n, bs = 3000, 8
    rng = np.random.default_rng(11)
    d = rng.choice([1.0, 2.0, 5.0], size=n * bs) * (1.0 + 1e-4 * rng.standard_normal(n * bs))
    rp = np.arange(n + 1, dtype=np.int32)
    ci = np.arange(n, dtype=np.int32)
    vals = np.zeros((n, bs, bs))
    for k in range(bs):
        vals[:, k, k] = d.reshape(n, bs)[:, k]
    A = bsp.BSR(ctx, n, n, bs, rp, ci, vals)
    bh = rng.standard_normal(n * bs)
    x, st, rc = bsp.FGMRES_mixed(ctx, A, ctx.to_device(bh), 1e-10, bsp.NoPre(ctx, A), spmv_precision=spmv)
    assert rc == 0 and st.true_relres <= 1e-10, (rc, st.true_relres, st.iters)
    assert st.reorth >= 1, st.reorth
    np.testing.assert_allclose(x.numpy(), bh / d, rtol=0, atol=1e-9 * np.abs(bh / d).max())


@pytest.mark.parametrize("restart,k", [(0, 0), (-1, 10), (63, 10), (30, 30), (30, -1)])
def test_fgmres_mixed_rejects_bad_restart(ctx, restart, k):
    p = pr.dg2d_veclap(6, 5, seed=1)
    A = bsp.BSR.from_problem(ctx, p)
    M = bsp.ILU0Ex(ctx, A, None, bsp.PREC_F32)
    with pytest.raises(bsp.BspError):
        bsp.FGMRES_mixed(ctx, A, ctx.to_device(p.b), 1e-10, M, restart=restart, k=k)


@pytest.mark.parametrize("prec", [bsp.PREC_F32, bsp.PREC_BF16])
def test_fgmres_mixed_ell_spmv_gives_the_chunk_major_iterates(ctx, prec):
    """The mixed solver's FP32 product over the ELL copy (bsmv_layout 2, the default) gives bitwise
    the iterates of the product over the chunk-major copy (bsmv_layout 1)."""
    p = pr.dg2d_veclap(24, 20, seed=2)
    col = pr.parity_colours((24, 20))
    res = {}
    try:
        for layout in (1, 2):
            ctx.set_option("bsmv_layout", layout)
            A = bsp.BSR.from_problem(ctx, p)      # a new matrix: its FP32 copy follows the option
            M = bsp.ILU0Ex(ctx, A, col, prec)
            xd, st, rc = bsp.FGMRES_mixed(ctx, A, ctx.to_device(p.b), 1e-10, M, maxIters=5000)
            assert rc == 0
            res[layout] = (st.iters, xd.numpy())
    finally:
        ctx.set_option("bsmv_layout", 2)
    assert res[1][0] == res[2][0]
    np.testing.assert_array_equal(res[1][1], res[2][1])


@pytest.mark.parametrize("pyth", [1, 0])
def test_fgmres_mixed_pythagorean_norm(ctx, pyth):
    """Classical Gram-Schmidt with the new norm taken as ||w||^2 - ||h||^2 and the normalisation in
    the update pass (cgs_pythagoras = 1) or summed after the update (0): both reach the FP64 residual
    and the oracle's solution, in iteration counts within a few percent of each other."""
    dims = (40, 36)
    p = pr.dg2d_veclap(*dims, seed=3)
    xo, sto = po.pgmres(p, p.b, 1e-12, po.PRE_ILU0, maxIters=5000, restart=30, k=10)
    its = {}
    try:
        for mode in (pyth, 1 - pyth):
            ctx.set_option("cgs_pythagoras", mode)
            A = bsp.BSR.from_problem(ctx, p)
            M = bsp.ILU0Ex(ctx, A, pr.parity_colours(dims), bsp.PREC_BF16)
            xd, st, rc = bsp.FGMRES_mixed(ctx, A, ctx.to_device(p.b), 1e-10, M, maxIters=5000, restart=30, k=10)
            x = xd.numpy()
            relres = np.linalg.norm(p.b - po.bsmv(p, x)) / np.linalg.norm(p.b)
            assert rc == 0 and relres <= 1e-10, (mode, rc, relres, st.iters)
            assert np.linalg.norm(x - xo) <= 1e-7 * np.linalg.norm(xo)
            its[mode] = st.iters
    finally:
        ctx.set_option("cgs_pythagoras", 1)
    assert abs(its[0] - its[1]) <= max(3, 0.05 * its[0]), its
