"""GPU parity of the HIP path against the CPU oracle (through the C ABI).

Tolerances:
  * bsmv FP64: |y_gpu - y_ref| <= 1e-13 * (|A||x|)_i per entry (cancellation-safe
    form of "1e-13 relative per entry"), and plain relative 1e-13 wherever
    |y_i| >= 1e-3 (|A||x|)_i.
  * structure / level sets: exact integer equality.
  * preconditioner applies: relative 2-norm difference <= 1e-10 (the device
    forms explicit diagonal-block inverses; the oracle uses LU solves).
  * solves: both reach the requested true relative residual.
"""
import ctypes as C

import numpy as np
import pytest

import blocksp_amd as bsp
import pyoracle as po
from blocksp_amd import problems as pr

pytestmark = pytest.mark.gpu


def random_bsr(n, bs, density_blocks=4, seed=0, empty_rows=True, square=True, ncols=None):
    rng = np.random.default_rng(seed)
    ncols = n if ncols is None else ncols
    rows = []
    for i in range(n):
        if empty_rows and rng.random() < 0.1:
            rows.append([])
            continue
        k = int(rng.integers(1, density_blocks + 1))
        cols = list(rng.choice(ncols, size=min(k, ncols), replace=False))
        if square and i < ncols and i not in cols:
            cols.append(i)
        rng.shuffle(cols)   # unsorted insertion order
        rows.append(cols)
    rowptr = np.zeros(n + 1, np.int32)
    rowptr[1:] = np.cumsum([len(r) for r in rows])
    colidx = np.array([c for r in rows for c in r], np.int32)
    vals = rng.standard_normal((colidx.shape[0], bs, bs))
    return rowptr, colidx, vals


def check_bsmv(y, yref, absax, tol=1e-13):
    err = np.abs(y - yref)
    assert np.all(err <= tol * absax + 1e-300), f"max scaled err {np.max(err / (absax + 1e-300)):.3e}"
    # entries without significant cancellation: plain relative error
    big = np.abs(yref) >= 0.5 * absax
    assert np.all(err[big] <= tol * np.abs(yref[big]))


@pytest.mark.parametrize("bs", [1, 2, 3, 4, 5, 6, 7, 8, 12, 16, 32])
def test_bsmv_random_structures(ctx, bs):
    print(f"bsmv bs={bs}", flush=True)
    n = 300
    rp, ci, v = random_bsr(n, bs, seed=bs)
    A = bsp.BSR(ctx, n, n, bs, rp, ci, v)
    rng = np.random.default_rng(100 + bs)
    x = rng.standard_normal(n * bs)
    y0 = rng.standard_normal(n * bs)
    xd = ctx.to_device(x)
    for acc in (0, 1):
        yd = ctx.to_device(y0)
        bsp.bsmv(ctx, A, xd, yd, accumulate=acc)
        y = yd.numpy()
        yref = po.bsmv_arrays(n, bs, rp, ci, v, x, y0 if acc else None, accumulate=acc)
        absax = po.bsmv_arrays(n, bs, rp, ci, np.abs(v), np.abs(x), np.abs(y0) if acc else None, accumulate=acc)
        check_bsmv(y, yref, absax)


def test_bsmv_rectangular_and_empty(ctx):
    bs = 4
    rp, ci, v = random_bsr(50, bs, seed=3, square=False, ncols=70)
    A = bsp.BSR(ctx, 50, 70, bs, rp, ci, v)
    x = np.random.default_rng(1).standard_normal(70 * bs)
    yd = ctx.zeros(50 * bs)
    bsp.bsmv(ctx, A, ctx.to_device(x), yd)
    yref = po.bsmv_arrays(50, bs, rp, ci, v, x)
    absax = po.bsmv_arrays(50, bs, rp, ci, np.abs(v), np.abs(x))
    check_bsmv(yd.numpy(), yref, absax)
    # a matrix with no blocks at all
    A0 = bsp.BSR(ctx, 5, 5, bs, np.zeros(6, np.int32), np.zeros(0, np.int32), np.zeros((0, bs, bs)))
    yd = ctx.to_device(np.ones(5 * bs))
    bsp.bsmv(ctx, A0, ctx.to_device(np.ones(5 * bs)), yd)
    assert np.all(yd.numpy() == 0.0)


@pytest.mark.parametrize("B,D", [(2, 2), (4, 2), (3, 1)])
def test_bsmv_dim(ctx, B, D):
    n = 120
    rp, ci, v = random_bsr(n, B, seed=B * 10 + D)
    A = bsp.BSR(ctx, n, n, B, rp, ci, v)
    x = np.random.default_rng(7).standard_normal(n * B * D)
    for dim in [-1] + list(range(D)):
        yd = ctx.zeros(n * B * D)
        bsp.bsmv_dim(ctx, A, D, dim, ctx.to_device(x), yd)
        yref = po.bsmv_dim(n, B, D, dim, rp, ci, v, x)
        absax = po.bsmv_dim(n, B, D, dim, rp, ci, np.abs(v), np.abs(x))
        check_bsmv(yd.numpy(), yref, absax)


@pytest.mark.parametrize("bs,k", [(8, 5), (8, 32), (8, 70), (8, 16), (8, 48), (8, 96), (1, 3), (3, 33), (4, 16),
                                  (16, 9), (32, 2), (3, 20), (5, 12), (7, 30), (6, 10), (12, 4)])
def test_bsmm_rhs(ctx, bs, k):
    """Multi-RHS product on the RHS tile kernels -- columns split over lane groups (bs a
    multiple of the group count) or rows split (bs 3, 5, 7 and 6 at 16-wide tiles), tiles of
    8-64 right-hand sides, several per row when nrhs > 64 -- and the per-output kernel (bs 12):
    each column within bsmv's bound."""
    n = 200 if bs <= 16 else 40
    rp, ci, v = random_bsr(n, bs, seed=11 + bs)
    A = bsp.BSR(ctx, n, n, bs, rp, ci, v)
    X = np.random.default_rng(2).standard_normal((n * bs, k))
    Yd = ctx.zeros(n * bs * k)
    bsp.bsmm_rhs(ctx, A, k, ctx.to_device(X.reshape(-1)), Yd)
    Y = Yd.numpy().reshape(n * bs, k)
    for j in range(k):
        yref = po.bsmv_arrays(n, bs, rp, ci, v, X[:, j])
        absax = po.bsmv_arrays(n, bs, rp, ci, np.abs(v), np.abs(X[:, j]))
        check_bsmv(Y[:, j], yref, absax)


def test_blas1(ctx):
    rng = np.random.default_rng(5)
    for n in (1, 7, 1000, 300001):
        x = rng.standard_normal(n)
        y = rng.standard_normal(n)
        xd, yd = ctx.to_device(x), ctx.to_device(y)
        o = C.c_double()
        ctx.check(ctx.L.bsp_dot(ctx.h, n, xd.ptr, yd.ptr, C.byref(o)))
        assert abs(o.value - x @ y) <= 1e-12 * (np.abs(x) @ np.abs(y))
        ctx.check(ctx.L.bsp_nrm2(ctx.h, n, xd.ptr, C.byref(o)))
        assert abs(o.value - np.linalg.norm(x)) <= 1e-13 * np.linalg.norm(x)
        ctx.check(ctx.L.bsp_infnorm(ctx.h, n, xd.ptr, C.byref(o)))
        assert o.value == np.abs(x).max()
        ctx.check(ctx.L.bsp_axpy(ctx.h, n, 0.5, xd.ptr, yd.ptr))
        assert np.allclose(yd.numpy(), y + 0.5 * x, rtol=0, atol=1e-15 * 4)


def _precond_case(name):
    if name == "stencil_b2":
        return pr.stencil2d_b2(16, 12, seed=0)
    if name == "dg2d":
        return pr.dg2d_veclap(12, 10, seed=1)
    if name == "dg3d":
        return pr.dg3d_lap(5, 4, 6, seed=2)
    raise KeyError(name)


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("case", ["stencil_b2", "dg2d", "dg3d"])
def test_ilu0_parity(ctx, case, mode):
    print(f"ilu0 {case} trsv={mode}", flush=True)
    ctx.set_option("trsv", mode)
    p = _precond_case(case)
    A = bsp.BSR.from_problem(ctx, p)
    M = bsp.ILU0(ctx, A)
    # level sets bit-exact
    for lower in (True, False):
        nl, lev = M.levels(lower)
        rp, ci, _ = M.factor()
        onl, olev = po.levels(p.n, rp, ci, lower)
        assert nl == onl and np.array_equal(lev, olev)
    # factor values
    lu, dlu, dpiv = po.ilu0_factor(p)
    frp, fci, fv = M.factor()
    assert np.array_equal(frp, p.rowptr) and np.array_equal(fci, p.colidx)
    assert np.max(np.abs(fv - lu)) <= 1e-11 * np.max(np.abs(lu))
    # apply
    r = np.random.default_rng(3).standard_normal(p.n * p.bs)
    rd = ctx.to_device(r)
    M(rd)
    z = rd.numpy()
    zref = po.ilu0_apply(p, lu, dlu, dpiv, r)
    assert np.linalg.norm(z - zref) <= 1e-10 * np.linalg.norm(zref)
    ctx.set_option("trsv", 0)


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("case", ["dg2d", "dg3d"])
def test_ichol0_parity(ctx, case, mode):
    print(f"ichol0 {case} trsv={mode}", flush=True)
    ctx.set_option("trsv", mode)
    p = _precond_case(case)
    A = bsp.BSR.from_problem(ctx, p)
    M = bsp.IChol0(ctx, A)
    lrp, lci, lv = po.ichol0_factor(p)
    frp, fci, fv = M.factor()
    assert np.array_equal(frp, lrp) and np.array_equal(fci, lci)
    assert np.max(np.abs(fv - lv)) <= 1e-11 * np.max(np.abs(lv))
    r = np.random.default_rng(4).standard_normal(p.n * p.bs)
    rd = ctx.to_device(r)
    M(rd)
    zref = po.ichol0_apply(p.n, p.bs, lrp, lci, lv, r)
    assert np.linalg.norm(rd.numpy() - zref) <= 1e-10 * np.linalg.norm(zref)
    ctx.set_option("trsv", 0)


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("typ", [-1, 0, 1])
def test_sor_parity(ctx, typ, mode):
    print(f"sor type={typ} trsv={mode}", flush=True)
    ctx.set_option("trsv", mode)
    p = _precond_case("stencil_b2")
    A = bsp.BSR.from_problem(ctx, p)
    for w, nu in ((1.2, 3), (0.6, 1), (1.0, 2)):
        M = bsp.SOR(ctx, A, w, nu, typ)
        r = np.random.default_rng(6).standard_normal(p.n * p.bs)
        rd = ctx.to_device(r)
        M(rd)
        zref = po.sor_apply(p, r, w, nu, typ)
        assert np.linalg.norm(rd.numpy() - zref) <= 1e-10 * np.linalg.norm(zref)
    ctx.set_option("trsv", 0)


def test_diag_parity(ctx):
    p = _precond_case("dg2d")
    A = bsp.BSR.from_problem(ctx, p)
    for chol in (False, True):
        M = bsp.Diag(ctx, A, cholesky=chol)
        r = np.random.default_rng(8).standard_normal(p.n * p.bs)
        rd = ctx.to_device(r)
        M(rd)
        zref = po.diag_apply(p, r, int(chol))
        assert np.linalg.norm(rd.numpy() - zref) <= 1e-12 * np.linalg.norm(zref)


def _true_relres(p, x):
    r = p.b - po.bsmv(p, x)
    return np.linalg.norm(r) / np.linalg.norm(p.b)


@pytest.mark.parametrize("restart,k", [(-1, -1), (30, 10), (30, 0), (12, 4)])
def test_pgmres_ilu0_parity(ctx, restart, k):
    print(f"pgmres ilu0 restart={restart} k={k}", flush=True)
    p = pr.dg2d_veclap(24, 20, seed=1)
    A = bsp.BSR.from_problem(ctx, p)
    M = bsp.ILU0(ctx, A)
    bd = ctx.to_device(p.b)
    xd, st, rc = bsp.pGMRES(ctx, A, bd, 1e-10, M, maxIters=2000, restart=restart, k=k)
    x = xd.numpy()
    xo, sto = po.pgmres(p, p.b, 1e-10, po.PRE_ILU0, maxIters=2000, restart=restart, k=k)
    tr = _true_relres(p, x)
    assert rc == 0 and st.converged == 1
    assert abs(st.true_relres - tr) <= 1e-3 * tr + 1e-16
    # same stopping rule, same attained accuracy (within a factor 10 of the oracle's)
    assert tr <= max(10 * sto.true_relres, 1e-9)
    assert abs(st.iters - sto.iters) <= max(2, 0.1 * sto.iters), (st.iters, sto.iters)
    assert np.linalg.norm(x - xo) <= 1e-6 * np.linalg.norm(xo)


def test_pgmres_sor_parity(ctx):
    p = pr.stencil2d_b2(32, 32, seed=0)
    A = bsp.BSR.from_problem(ctx, p)
    M = bsp.SOR(ctx, A, 1.2, 3, 0)
    xd, st, rc = bsp.pGMRES(ctx, A, ctx.to_device(p.b), 1e-10, M, maxIters=2000, restart=30, k=10)
    xo, sto = po.pgmres(p, p.b, 1e-10, po.PRE_SOR, maxIters=2000, restart=30, k=10, w=1.2, nu=3)
    assert rc == 0
    assert _true_relres(p, xd.numpy()) <= max(10 * sto.true_relres, 1e-9)
    assert abs(st.iters - sto.iters) <= max(2, 0.1 * sto.iters)


def test_pcg_ichol0_parity(ctx):
    p = pr.dg2d_veclap(24, 20, seed=1)
    A = bsp.BSR.from_problem(ctx, p)
    M = bsp.IChol0(ctx, A)
    xd, st, rc = bsp.pCG(ctx, A, ctx.to_device(p.b), 1e-10, M, maxIters=2000)
    xo, sto = po.pcg(p, p.b, 1e-10, po.PRE_ICHOL0, maxIters=2000)
    assert rc == 0
    assert abs(st.iters - sto.iters) <= 2
    assert _true_relres(p, xd.numpy()) <= max(10 * sto.true_relres, 1e-9)


def test_true_residual_mode(ctx):
    p = pr.dg2d_veclap(24, 20, seed=1)
    A = bsp.BSR.from_problem(ctx, p)
    M = bsp.ILU0(ctx, A)
    xd, st, rc = bsp.pGMRES(ctx, A, ctx.to_device(p.b), 1e-10, M, maxIters=3000, restart=30, k=10,
                            tol_mode=bsp.TOL_TRUE_RESIDUAL)
    assert rc == 0 and _true_relres(p, xd.numpy()) <= 1e-10
    xd, st, rc = bsp.pCG(ctx, A, ctx.to_device(p.b), 1e-10, bsp.IChol0(ctx, A), maxIters=3000,
                         tol_mode=bsp.TOL_TRUE_RESIDUAL)
    assert rc == 0 and _true_relres(p, xd.numpy()) <= 1e-10


def test_ldg_fixture_known_answer(ctx):
    """The reference's own fixture: ||x - x_sol||_inf ~ 0.00029 (blocksp_tests.hpp:318)."""
    import os
    from blocksp_amd import mtx
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "ldg_2d_veclap.npz"))
    for bs in (2, 4, 8):
        rp, ci, v = mtx.entries_to_bsr(int(z["M"]), int(z["N"]), z["i"], z["j"], z["v"], bs, symmetric=True)
        n = rp.shape[0] - 1
        A = bsp.BSR(ctx, n, n, bs, rp, ci, v)
        bd = ctx.to_device(z["b"])
        for M in (bsp.ILU0(ctx, A), bsp.SOR(ctx, A, 1.6, 3, 0)):
            xd, st, rc = bsp.pGMRES(ctx, A, bd, 1e-8, M, maxIters=5000, restart=30, k=10)
            err = np.abs(xd.numpy() - z["x_sol"]).max()
            assert rc == 0 and abs(err - 2.946e-4) < 2e-6, err
        xd, st, rc = bsp.pCG(ctx, A, bd, 1e-8, bsp.IChol0(ctx, A), maxIters=5000)
        err = np.abs(xd.numpy() - z["x_sol"]).max()
        assert rc == 0 and abs(err - 2.946e-4) < 2e-6, err


def test_full_size_bsmv_properties(ctx):
    """dg2d_veclap at full size (262144 block rows, bs=8): linearity and A*1 row sums."""
    p = pr.dg2d_veclap()
    A = bsp.BSR.from_problem(ctx, p)
    n = p.n * p.bs
    rng = np.random.default_rng(0)
    x1, x2 = rng.standard_normal(n), rng.standard_normal(n)
    y1, y2, y3 = ctx.zeros(n), ctx.zeros(n), ctx.zeros(n)
    bsp.bsmv(ctx, A, ctx.to_device(x1), y1)
    bsp.bsmv(ctx, A, ctx.to_device(x2), y2)
    bsp.bsmv(ctx, A, ctx.to_device(2.0 * x1 - 3.0 * x2), y3)
    lhs, rhs = y3.numpy(), 2.0 * y1.numpy() - 3.0 * y2.numpy()
    assert np.max(np.abs(lhs - rhs)) <= 1e-12 * np.max(np.abs(rhs))
    # against the numpy product (independent order)
    assert np.max(np.abs(y1.numpy() - p.bsmv(x1))) <= 1e-12 * np.max(np.abs(p.bsmv(np.abs(x1))))


@pytest.mark.parametrize("layout", [0, 1, 2])
@pytest.mark.parametrize("bs", [2, 3, 4, 8])
def test_bsmv_layouts(ctx, bs, layout):
    """The solvers' bsmv on the values as stored (0), on their chunk-major copy (1) and on an
    ELL copy of that (2, the default for even bs <= 8; these random rows pad too much for an
    ELL copy, so 2 reads the chunk-major copy; bs 3 always reads the stored values)."""
    ctx.set_option("bsmv_layout", layout)
    try:
        n = 257
        rp, ci, v = random_bsr(n, bs, density_blocks=9, seed=40 + bs)
        A = bsp.BSR(ctx, n, n, bs, rp, ci, v)
        x = np.random.default_rng(3).standard_normal(n * bs)
        yd = ctx.zeros(n * bs)
        bsp.bsmv(ctx, A, ctx.to_device(x), yd)
        yref = po.bsmv_arrays(n, bs, rp, ci, v, x)
        absax = po.bsmv_arrays(n, bs, rp, ci, np.abs(v), np.abs(x))
        check_bsmv(yd.numpy(), yref, absax)
    finally:
        ctx.set_option("bsmv_layout", 2)   # the default


def uniform_bsr(n, bs, seed):
    """Rows of 4 or 5 blocks in shuffled insertion order, one empty row: an ELL copy of width 5
    pads about a tenth of the blocks."""
    rng = np.random.default_rng(seed)
    rows = []
    for i in range(n):
        if i == 3:
            rows.append([])
            continue
        cols = list(rng.choice(n, size=int(rng.integers(4, 6)), replace=False))
        rng.shuffle(cols)
        rows.append(cols)
    rowptr = np.zeros(n + 1, np.int32)
    rowptr[1:] = np.cumsum([len(r) for r in rows])
    colidx = np.array([c for r in rows for c in r], np.int32)
    return rowptr, colidx, rng.standard_normal((colidx.shape[0], bs, bs))


@pytest.mark.parametrize("bs", [2, 4, 6, 8])
def test_bsmv_ell_copy_matches_oracle_and_chunk_major_bitwise(ctx, bs):
    """bsmv over the ELL copy (layout 2): within the oracle's bound and bitwise the product over
    the chunk-major copy (layout 1) -- the same blocks added in the same order."""
    n = 301
    rp, ci, v = uniform_bsr(n, bs, seed=60 + bs)
    x = np.random.default_rng(5).standard_normal(n * bs)
    out = {}
    try:
        for layout in (1, 2):
            ctx.set_option("bsmv_layout", layout)
            A = bsp.BSR(ctx, n, n, bs, rp, ci, v)
            yd = ctx.zeros(n * bs)
            bsp.bsmv(ctx, A, ctx.to_device(x), yd)
            out[layout] = yd.numpy()
    finally:
        ctx.set_option("bsmv_layout", 2)
    yref = po.bsmv_arrays(n, bs, rp, ci, v, x)
    absax = po.bsmv_arrays(n, bs, rp, ci, np.abs(v), np.abs(x))
    check_bsmv(out[2], yref, absax)
    np.testing.assert_array_equal(out[2], out[1])


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("typ", [-1, 0, 1])
def test_multicolour_sor_is_the_sor_of_the_permuted_matrix(ctx, typ, mode):
    # rows swept colour by colour (coordinate parity, natural order within a colour): equal to
    # the reference's SOR of P A P^T (oracle) on the permuted vector; 3 levels per sweep
    from test_gpu_mixed import permuted
    print(f"multicolour sor type={typ} trsv={mode}", flush=True)
    ctx.set_option("trsv", mode)
    try:
        dims = (11, 9)
        p = pr.dg2d_veclap(*dims, seed=2)
        col = pr.parity_colours(dims)
        perm = np.argsort(col, kind="stable")
        q = permuted(p, perm)
        A = bsp.BSR.from_problem(ctx, p)
        for w, nu in ((1.2, 3), (0.7, 1)):
            M = bsp.SOR(ctx, A, w, nu, typ, colour=col)
            r = np.random.default_rng(7).standard_normal(p.n * p.bs)
            rd = ctx.to_device(r)
            M(rd)
            zq = po.sor_apply(q, r.reshape(p.n, p.bs)[perm].reshape(-1), w, nu, typ)
            z = rd.numpy().reshape(p.n, p.bs)[perm].reshape(-1)
            assert np.linalg.norm(z - zq) <= 1e-10 * np.linalg.norm(zq)
        nl_lo, _ = M.levels(True)
        assert nl_lo == 3
    finally:
        ctx.set_option("trsv", 0)
