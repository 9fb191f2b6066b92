"""Device block ILU(0) / IChol(0) factorisation against the oracle and the host loop.

The device factor (level-scheduled rows, one lane per block entry) restates the reference's
right-looking factorisations (blocksp_preconditioners.hpp:274-291 ILU0, 374-429 IChol0) row by
row.  Pins:
  * structure of the factor bit-exact and values within 1e-11 (relative to the largest entry)
    of the oracle's factor, for insertion-ordered (unsorted) rows and block sizes 1-8;
  * the same against the library's host loop (option factor_device = 0), natural and
    colour-ordered elimination;
  * the preconditioner applies agree with the oracle's apply to 1e-10;
  * a singular pivot block (ILU0) and a pivot that is not positive definite (IChol0) are
    reported as BSP_E_NUMERIC.
"""
import numpy as np
import pytest

import blocksp_amd as bsp
import pyoracle as po
from blocksp_amd import problems as pr

pytestmark = pytest.mark.gpu


def shuffled(p, seed):
    """The same matrix with every row's blocks in a random (insertion) order."""
    rng = np.random.default_rng(seed)
    ci, vals = p.colidx.copy(), p.vals.copy()
    for i in range(p.n):
        a, b = p.rowptr[i], p.rowptr[i + 1]
        o = rng.permutation(b - a)
        ci[a:b] = p.colidx[a:b][o]
        vals[a:b] = p.vals[a:b][o]
    return pr.Problem(p.n, p.bs, p.B, p.D, p.rowptr.copy(), ci, vals, p.b.copy())


def random_block(nx, ny, bs, seed, spd):
    """5-point grid of bs x bs blocks, structurally symmetric; SPD (diagonally dominant
    symmetric) when asked, otherwise nonsymmetric with a dominant diagonal."""
    rng = np.random.default_rng(seed)
    n = nx * ny
    rows = []
    for iy in range(ny):
        for ix in range(nx):
            i = iy * nx + ix
            nb = [i]
            if ix > 0:
                nb.append(i - 1)
            if ix < nx - 1:
                nb.append(i + 1)
            if iy > 0:
                nb.append(i - nx)
            if iy < ny - 1:
                nb.append(i + nx)
            rows.append(sorted(nb))
    rp = np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype(np.int32)
    ci = np.concatenate(rows).astype(np.int32)
    vals = np.zeros((len(ci), bs, bs))
    off = {}
    for i in range(n):
        for q in range(rp[i], rp[i + 1]):
            j = ci[q]
            if j == i:
                continue
            if spd and (j, i) in off:
                vals[q] = off[(j, i)].T
            else:
                vals[q] = -0.2 * np.eye(bs) + 0.1 * rng.standard_normal((bs, bs))
                off[(i, j)] = vals[q]
    for i in range(n):
        for q in range(rp[i], rp[i + 1]):
            if ci[q] == i:
                s = rng.standard_normal((bs, bs))
                vals[q] = 2.0 * bs * np.eye(bs) + 0.2 * (s + s.T if spd else s)
    b = rng.standard_normal(n * bs)
    return pr.Problem(n, bs, bs, 1, rp, ci, vals, b)


def _rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.mark.parametrize("bs", [1, 2, 3, 4, 5, 8])
def test_device_ilu0_matches_oracle_on_unsorted_rows(ctx, bs):
    print(f"device ilu0 bs={bs}", flush=True)
    p = shuffled(random_block(9, 7, bs, seed=bs, spd=False), seed=bs + 10)
    A = bsp.BSR.from_problem(ctx, p)
    M = bsp.ILU0(ctx, A)
    lu, dlu, dpiv = po.ilu0_factor(p)
    frp, fci, fv = M.factor()
    assert np.array_equal(frp, p.rowptr) and np.array_equal(fci, p.colidx)
    assert _rel(fv, lu) <= 1e-11
    r = np.random.default_rng(1).standard_normal(p.n * p.bs)
    rd = ctx.to_device(r)
    M(rd)
    zref = po.ilu0_apply(p, lu, dlu, dpiv, r)
    assert np.linalg.norm(rd.numpy() - zref) <= 1e-10 * np.linalg.norm(zref)


@pytest.mark.parametrize("bs", [1, 2, 4, 6, 8])
def test_device_ichol0_matches_oracle_on_unsorted_rows(ctx, bs):
    print(f"device ichol0 bs={bs}", flush=True)
    p = shuffled(random_block(8, 9, bs, seed=20 + bs, spd=True), seed=bs + 30)
    A = bsp.BSR.from_problem(ctx, p)
    M = bsp.IChol0(ctx, A)
    lrp, lci, lv = po.ichol0_factor(p)
    frp, fci, fv = M.factor()
    assert np.array_equal(frp, lrp) and np.array_equal(fci, lci)
    assert _rel(fv, lv) <= 1e-11
    r = np.random.default_rng(2).standard_normal(p.n * p.bs)
    rd = ctx.to_device(r)
    M(rd)
    zref = po.ichol0_apply(p.n, p.bs, lrp, lci, lv, r)
    assert np.linalg.norm(rd.numpy() - zref) <= 1e-10 * np.linalg.norm(zref)


def _factor_with(ctx, make, device):
    ctx.set_option("factor_device", 1 if device else 0)
    try:
        M = make()
        return M, M.factor()
    finally:
        ctx.set_option("factor_device", 1)


@pytest.mark.parametrize("case", ["dg2d", "dg3d", "stencil_b2"])
@pytest.mark.parametrize("coloured", [False, True])
def test_device_factor_equals_host_loop(ctx, case, coloured):
    print(f"device vs host factor {case} coloured={coloured}", flush=True)
    if case == "dg2d":
        dims = (20, 18)
        p = pr.dg2d_veclap(*dims, seed=1)
    elif case == "dg3d":
        dims = (6, 5, 7)
        p = pr.dg3d_lap(*dims, seed=2)
    else:
        dims = (24, 20)
        p = pr.stencil2d_b2(*dims, seed=0)
    col = pr.parity_colours(dims) if coloured else None
    A = bsp.BSR.from_problem(ctx, p)
    r = np.random.default_rng(3).standard_normal(p.n * p.bs)
    for kind in ("ilu0", "ichol0"):
        if kind == "ichol0" and case == "stencil_b2":
            continue          # nonsymmetric matrix
        make = (lambda: bsp.ILU0Ex(ctx, A, col, bsp.PREC_F64)) if kind == "ilu0" else \
            (lambda: bsp.IChol0(ctx, A, colour=col))
        Md, (drp, dci, dv) = _factor_with(ctx, make, True)
        Mh, (hrp, hci, hv) = _factor_with(ctx, make, False)
        assert np.array_equal(drp, hrp) and np.array_equal(dci, hci)
        assert _rel(dv, hv) <= 1e-12, kind
        for lower in (True, False):
            assert Md.levels(lower)[0] == Mh.levels(lower)[0]
            assert np.array_equal(Md.levels(lower)[1], Mh.levels(lower)[1])
        a, b = ctx.to_device(r), ctx.to_device(r)
        Md(a)
        Mh(b)
        assert np.linalg.norm(a.numpy() - b.numpy()) <= 1e-12 * np.linalg.norm(b.numpy()), kind


def test_device_factor_large_natural_dg2d(ctx):
    # 96 x 96 DG vector Laplacian (9216 block rows, bs 8, 191 levels) against the oracle
    p = pr.dg2d_veclap(96, 96, seed=1)
    A = bsp.BSR.from_problem(ctx, p)
    M = bsp.ILU0(ctx, A)
    lu, _, _ = po.ilu0_factor(p)
    assert _rel(M.factor()[2], lu) <= 1e-11
    Mc = bsp.IChol0(ctx, A)
    _, _, lv = po.ichol0_factor(p)
    assert _rel(Mc.factor()[2], lv) <= 1e-11


@pytest.mark.parametrize("bs", [2, 8])
def test_singular_pivot_is_an_error(ctx, bs):
    p = random_block(5, 4, bs, seed=3, spd=False)
    vals = p.vals.copy()
    # the diagonal block of row 0 is singular: its ILU0 pivot is A_00 itself
    d0 = [q for q in range(p.rowptr[0], p.rowptr[1]) if p.colidx[q] == 0][0]
    vals[d0] = 0.0
    A = bsp.BSR(ctx, p.n, p.n, bs, p.rowptr, p.colidx, vals)
    with pytest.raises(bsp.BspError) as ei:
        bsp.ILU0(ctx, A)
    assert ei.value.code == bsp.E_NUMERIC


def test_not_positive_definite_pivot_is_an_error(ctx):
    p = random_block(5, 4, 4, seed=4, spd=True)
    vals = p.vals.copy()
    d = [q for q in range(p.rowptr[7], p.rowptr[8]) if p.colidx[q] == 7][0]
    vals[d] = -vals[d]
    A = bsp.BSR(ctx, p.n, p.n, 4, p.rowptr, p.colidx, vals)
    with pytest.raises(bsp.BspError) as ei:
        bsp.IChol0(ctx, A)
    assert ei.value.code == bsp.E_NUMERIC
