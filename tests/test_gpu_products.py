"""Transposed block products and block SpGEMM (bsmm) against the oracle."""
import numpy as np
import pytest

import blocksp_amd as bsp
from blocksp_amd import problems as pr
import pyoracle as po

pytestmark = pytest.mark.gpu


def _rand_bsr(nr, nc, bs, seed, dens=0.3):
    rng = np.random.default_rng(seed)
    rows = []
    for i in range(nr):
        k = int(rng.integers(0, max(1, int(dens * nc)) + 1)) if i != 2 else 0   # row 2 empty
        rows.append(rng.permutation(rng.choice(nc, size=min(k, nc), replace=False)))   # unsorted columns
    rp = np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype(np.int32)
    ci = np.concatenate(rows).astype(np.int32) if rp[-1] else np.zeros(0, np.int32)
    return rp, ci, rng.standard_normal((len(ci), bs, bs))


def _abs_transpose_bound(nr, nc, bs, rp, ci, v, x):
    return po.bsmv_transpose(nr, nc, bs, rp, ci, np.abs(v), np.abs(x))


@pytest.mark.parametrize("bs", [1, 3, 8])
def test_bsmv_transpose_matches_oracle(ctx, bs):
    nr, nc = 23, 17
    rp, ci, v = _rand_bsr(nr, nc, bs, seed=bs)
    A = bsp.BSR(ctx, nr, nc, bs, rp, ci, v)
    x = np.random.default_rng(1).standard_normal(nr * bs)
    y0 = np.random.default_rng(2).standard_normal(nc * bs)
    bound = _abs_transpose_bound(nr, nc, bs, rp, ci, v, x)
    for acc in (False, True):
        xd, yd = ctx.to_device(x), ctx.to_device(y0)
        bsp.bsmv_transpose(ctx, A, xd, yd, accumulate=acc)
        ref = po.bsmv_transpose(nr, nc, bs, rp, ci, v, x, y=y0 if acc else None, accumulate=acc)
        assert np.all(np.abs(yd.numpy() - ref) <= 1e-13 * (bound + (np.abs(y0) if acc else 0)) + 1e-300)


def test_bsmv_dim_transpose_is_the_per_dimension_adjoint(ctx):
    # <A^T x, y>_dim == <x, A y>_dim for every dimension, and dim = -1 covers all
    n, B, D = 19, 4, 3
    rp, ci, v = _rand_bsr(n, n, B, seed=7)
    A = bsp.BSR(ctx, n, n, B, rp, ci, v)
    rng = np.random.default_rng(3)
    x, y = rng.standard_normal(n * B * D), rng.standard_normal(n * B * D)
    xd, yd = ctx.to_device(x), ctx.to_device(y)
    Aty, Ax = ctx.zeros(n * B * D), ctx.zeros(n * B * D)
    bsp.bsmv_dim_transpose(ctx, A, D, -1, yd, Aty)
    bsp.bsmv_dim(ctx, A, D, -1, xd, Ax)
    at = Aty.numpy().reshape(n, D, B)
    ax = Ax.numpy().reshape(n, D, B)
    for d in range(D):
        lhs = np.sum(x.reshape(n, D, B)[:, d] * at[:, d])
        rhs = np.sum(ax[:, d] * y.reshape(n, D, B)[:, d])
        assert abs(lhs - rhs) <= 1e-12 * (abs(lhs) + 1)
    # one dimension only: the others stay zero
    out = ctx.zeros(n * B * D)
    bsp.bsmv_dim_transpose(ctx, A, D, 1, yd, out)
    o = out.numpy().reshape(n, D, B)
    np.testing.assert_allclose(o[:, 1], at[:, 1], rtol=1e-14, atol=1e-14)
    assert not o[:, 0].any() and not o[:, 2].any()


@pytest.mark.parametrize("bs", [2, 8])
def test_bsmm_structure_bit_exact_values_within_tolerance(ctx, bs):
    n, m, k = 21, 15, 18
    arp, aci, av = _rand_bsr(n, m, bs, seed=10 + bs)
    brp, bci, bv = _rand_bsr(m, k, bs, seed=20 + bs)
    A = bsp.BSR(ctx, n, m, bs, arp, aci, av)
    Bm = bsp.BSR(ctx, m, k, bs, brp, bci, bv)
    Cm = bsp.bsmm(ctx, A, Bm)
    crp, cci, cv = Cm.download()
    orp, oci, ov = po.bsmm(n, bs, arp, aci, av, brp, bci, bv)
    np.testing.assert_array_equal(crp, orp)
    np.testing.assert_array_equal(cci, oci)
    _, _, bound = po.bsmm(n, bs, arp, aci, np.abs(av), brp, bci, np.abs(bv))
    assert np.all(np.abs(cv - ov) <= 1e-13 * bound + 1e-300)
    # new values, same structures: the numeric pass only
    av2, bv2 = -2.0 * av + 1.0, 0.5 * bv
    av2_d, bv2_d = ctx.to_device(av2), ctx.to_device(bv2)      # kept alive across the calls
    ctx.check(ctx.L.bsp_bsr_set_values_dev(ctx.h, A.h, av2_d.ptr), "set_values")
    ctx.check(ctx.L.bsp_bsr_set_values_dev(ctx.h, Bm.h, bv2_d.ptr), "set_values")
    bsp.bsmm_numeric(ctx, A, Bm, Cm)
    _, _, ov2 = po.bsmm(n, bs, arp, aci, av2, brp, bci, bv2)
    _, _, bound2 = po.bsmm(n, bs, arp, aci, np.abs(av2), brp, bci, np.abs(bv2))
    assert np.all(np.abs(Cm.download()[2] - ov2) <= 1e-13 * bound2 + 1e-300)


def test_bsmm_square_of_dg_matrix_feeds_a_preconditioner(ctx):
    p = pr.dg2d_veclap(12, 12, seed=1)
    A = bsp.BSR.from_problem(ctx, p)
    A2 = bsp.bsmm(ctx, A, A)
    orp, oci, ov = po.bsmm(p.n, p.bs, p.rowptr, p.colidx, p.vals, p.rowptr, p.colidx, p.vals)
    rp, ci, v = A2.download()
    np.testing.assert_array_equal(rp, orp)
    np.testing.assert_array_equal(ci, oci)
    assert np.abs(v - ov).max() <= 1e-12 * np.abs(ov).max()
    # the device-computed values reach the host factorisation (ILU0 of A^2)
    M = bsp.ILU0(ctx, A2)
    _, _, lu = M.factor()
    q = pr.Problem(p.n, p.bs, p.B, p.D, orp, oci, ov, p.b)
    lu_ref = po.ilu0_factor(q)[0]
    assert np.abs(lu - lu_ref).max() <= 1e-9 * np.abs(lu_ref).max()
    # A^T on a symmetric matrix is A
    x = ctx.to_device(p.b)
    y1, y2 = ctx.zeros(p.n * p.bs), ctx.zeros(p.n * p.bs)
    bsp.bsmv(ctx, A, x, y1)
    bsp.bsmv_transpose(ctx, A, x, y2)
    np.testing.assert_allclose(y2.numpy(), y1.numpy(), rtol=1e-12, atol=1e-12 * np.abs(y1.numpy()).max())


@pytest.mark.parametrize("bs", [3, 8])
def test_bsmm_transA_structure_bit_exact_values_within_tolerance(ctx, bs):
    n, m, k = 19, 13, 16
    arp, aci, av = _rand_bsr(n, m, bs, seed=50 + bs)
    brp, bci, bv = _rand_bsr(n, k, bs, seed=60 + bs)
    A = bsp.BSR(ctx, n, m, bs, arp, aci, av)
    Bm = bsp.BSR(ctx, n, k, bs, brp, bci, bv)
    Cm = bsp.bsmm(ctx, A, Bm, Atran=True)
    assert (Cm.n_rows, Cm.n_cols) == (m, k)
    crp, cci, cv = Cm.download()
    orp, oci, ov = po.bsmm_transA(n, m, bs, arp, aci, av, brp, bci, bv)
    np.testing.assert_array_equal(crp, orp)
    np.testing.assert_array_equal(cci, oci)
    _, _, bound = po.bsmm_transA(n, m, bs, arp, aci, np.abs(av), brp, bci, np.abs(bv))
    assert np.all(np.abs(cv - ov) <= 1e-13 * bound + 1e-300)
    # new values of A: the numeric pass rebuilds A^T from them
    av2 = 3.0 * av - 0.5
    av2_d = ctx.to_device(av2)
    ctx.check(ctx.L.bsp_bsr_set_values_dev(ctx.h, A.h, av2_d.ptr), "set_values")
    bsp.bsmm_numeric(ctx, A, Bm, Cm)
    _, _, ov2 = po.bsmm_transA(n, m, bs, arp, aci, av2, brp, bci, bv)
    _, _, bound2 = po.bsmm_transA(n, m, bs, arp, aci, np.abs(av2), brp, bci, np.abs(bv))
    assert np.all(np.abs(Cm.download()[2] - ov2) <= 1e-13 * bound2 + 1e-300)


def test_bsmm_transA_of_dg_matrix_is_its_square_and_gives_frobenius_norm(ctx):
    # the DG matrix is symmetric, so A^T A = A A; trace(A^T A) = ||A||_F^2 (utility.hpp:383-388)
    p = pr.dg2d_veclap(10, 10, seed=2)
    A = bsp.BSR.from_problem(ctx, p)
    AtA = bsp.bsmm(ctx, A, A, Atran=True)
    rp, ci, v = AtA.download()
    orp, oci, ov = po.bsmm_transA(p.n, p.n, p.bs, p.rowptr, p.colidx, p.vals, p.rowptr, p.colidx, p.vals)
    np.testing.assert_array_equal(rp, orp)
    np.testing.assert_array_equal(ci, oci)
    assert np.abs(v - ov).max() <= 1e-12 * np.abs(ov).max()
    tr = sum(np.trace(v[q]) for i in range(p.n) for q in range(rp[i], rp[i + 1]) if ci[q] == i)
    fro2 = float(np.sum(p.vals * p.vals))
    assert abs(tr - fro2) <= 1e-12 * fro2


def test_bsmm_transA_rejects_mismatched_rows(ctx):
    rp, ci, v = _rand_bsr(5, 4, 2, seed=1)
    rp2, ci2, v2 = _rand_bsr(6, 4, 2, seed=2)
    A = bsp.BSR(ctx, 5, 4, 2, rp, ci, v)
    Bm = bsp.BSR(ctx, 6, 4, 2, rp2, ci2, v2)
    with pytest.raises(Exception):
        bsp.bsmm(ctx, A, Bm, Atran=True)


@pytest.mark.parametrize("bs,width", [(1, 12), (4, 9), (8, 7), (16, 6), (5, 11)])
def test_bsmm_wide_rows_and_block_sizes(ctx, bs, width):
    """Rows of C with more blocks than one register pass holds (16), empty rows, and the block
    sizes of the row kernel: structure bit-exact (first-touch order), values to 1e-13 of |A||B|."""
    rng = np.random.default_rng(bs * 100 + width)
    n, m, k = 40, 37, 45

    def rand(nr, nc, w):
        rows = []
        for i in range(nr):
            cnt = 0 if i % 13 == 5 else int(rng.integers(1, w + 1))
            rows.append(rng.choice(nc, size=cnt, replace=False))   # unsorted: insertion order
        rp = np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype(np.int32)
        ci = np.concatenate(rows).astype(np.int32)
        return rp, ci, rng.standard_normal((len(ci), bs, bs))

    arp, aci, av = rand(n, m, width)
    brp, bci, bv = rand(m, k, width)
    Cm = bsp.bsmm(ctx, bsp.BSR(ctx, n, m, bs, arp, aci, av), bsp.BSR(ctx, m, k, bs, brp, bci, bv))
    crp, cci, cv = Cm.download()
    orp, oci, ov = po.bsmm(n, bs, arp, aci, av, brp, bci, bv)
    np.testing.assert_array_equal(crp, orp)
    np.testing.assert_array_equal(cci, oci)
    assert np.diff(orp).max() > 16        # some rows need more than one pass
    _, _, bound = po.bsmm(n, bs, arp, aci, np.abs(av), brp, bci, np.abs(bv))
    assert np.all(np.abs(cv - ov) <= 1e-13 * bound + 1e-300)


@pytest.mark.parametrize("mfma", [0, 1])
def test_bsmm_8x8_fma_and_matrix_core_paths(ctx, mfma):
    """The 8 x 8 numeric pass on the matrix cores (v_mfma_f64_16x16x4f64, two product streams
    per 16 x 16 tile) and the FMA row kernel both meet the oracle's bound, on the DG operator,
    on random insertion-ordered rows wider than one register pass and on dense block rows."""
    ctx.set_option("bsmm_mfma", mfma)
    try:
        p = pr.dg3d_lap(5, 4, 6, seed=2)
        A = bsp.BSR.from_problem(ctx, p)
        rp, ci, v = bsp.bsmm(ctx, A, A).download()
        orp, oci, ov = po.bsmm(p.n, 8, p.rowptr, p.colidx, p.vals, p.rowptr, p.colidx, p.vals)
        _, _, bound = po.bsmm(p.n, 8, p.rowptr, p.colidx, np.abs(p.vals), p.rowptr, p.colidx, np.abs(p.vals))
        np.testing.assert_array_equal(rp, orp)
        np.testing.assert_array_equal(ci, oci)
        assert np.all(np.abs(v - ov) <= 1e-13 * bound + 1e-300)
        n, m, k = 33, 29, 31
        arp, aci, av = _rand_bsr(n, m, 8, seed=91, dens=0.4)
        brp, bci, bv = _rand_bsr(m, k, 8, seed=92, dens=0.4)
        rp, ci, v = bsp.bsmm(ctx, bsp.BSR(ctx, n, m, 8, arp, aci, av), bsp.BSR(ctx, m, k, 8, brp, bci, bv)).download()
        orp, oci, ov = po.bsmm(n, 8, arp, aci, av, brp, bci, bv)
        _, _, bound = po.bsmm(n, 8, arp, aci, np.abs(av), brp, bci, np.abs(bv))
        np.testing.assert_array_equal(ci, oci)
        assert np.all(np.abs(v - ov) <= 1e-13 * bound + 1e-300)
        # dense block rows: 24 blocks of 24 products each per row of C (many steps per stream)
        n = 24
        rng = np.random.default_rng(93)
        rp = (np.arange(n + 1) * n).astype(np.int32)
        ci = np.concatenate([rng.permutation(n) for _ in range(n)]).astype(np.int32)
        dv = rng.standard_normal((n * n, 8, 8))
        D = bsp.BSR(ctx, n, n, 8, rp, ci, dv)
        rp2, ci2, v2 = bsp.bsmm(ctx, D, D).download()
        orp, oci, ov = po.bsmm(n, 8, rp, ci, dv, rp, ci, dv)
        _, _, bound = po.bsmm(n, 8, rp, ci, np.abs(dv), rp, ci, np.abs(dv))
        np.testing.assert_array_equal(ci2, oci)
        assert np.all(np.abs(v2 - ov) <= 1e-13 * bound + 1e-300)
    finally:
        ctx.set_option("bsmm_mfma", 1)
