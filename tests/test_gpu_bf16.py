"""BF16 block SpMM on the matrix cores against a numpy statement of the same arithmetic."""
import numpy as np
import pytest

import blocksp_amd as bsp
from blocksp_amd import problems as pr

pytestmark = pytest.mark.gpu


def bf16_bits(x):
    """float64 -> float32 -> BF16 bits, round to nearest even (finite inputs)."""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return u.astype(np.uint16)


def bf16_value(bits):
    return (bits.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def _random_bsr(n, seed, even=False):
    """Random block pattern; row 5 empty.  even=False: 1-7 blocks per row and row 7 with 11 (more
    than one round of slots); even=True: 5 or 6 blocks per row (rows padded to an ELL copy)."""
    rng = np.random.default_rng(seed)
    rows = []
    for i in range(n):
        if even:
            k = 0 if i == 5 else rng.integers(5, 7)
        else:
            k = 0 if i == 5 else (11 if i == 7 else rng.integers(1, 8))
        rows.append(np.sort(rng.choice(n, size=k, replace=False)))
    rp = np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype(np.int32)
    ci = np.concatenate(rows).astype(np.int32)
    vals = rng.standard_normal((len(ci), 8, 8))
    return rp, ci, vals


def _ref(rp, ci, Ab, Xb, n, nrhs):
    # Xb: (n, 8, nrhs) values; Y[i] = sum_p A_p X[col p]
    Y = np.zeros((n, 8, nrhs))
    for i in range(n):
        for p in range(rp[i], rp[i + 1]):
            Y[i] += Ab[p] @ Xb[ci[p]]
    return Y


@pytest.fixture(params=[1, 0])
def kernel(ctx, request):
    """bf16_ell = 1: ELL copy and persistent kernel where rows allow it; 0: CSR copy."""
    ctx.set_option("bf16_ell", request.param)
    yield request.param
    ctx.set_option("bf16_ell", 1)


@pytest.mark.parametrize("even", [False, True])
@pytest.mark.parametrize("n,nrhs", [(37, 32), (36, 16), (11, 48)])
def test_bf16_spmm_matches_numpy(ctx, kernel, n, nrhs, even):
    _check_spmm(ctx, n, nrhs, even)


def _check_spmm(ctx, n, nrhs, even=False):
    rp, ci, vals = _random_bsr(n, seed=n + nrhs, even=even)
    A = bsp.BSR(ctx, n, n, 8, rp, ci, vals)
    X = np.random.default_rng(1).standard_normal((n * 8, nrhs))     # RHS-minor
    Xd = ctx.to_device(X)
    Xb = ctx.empty((n * nrhs * 8,), np.uint16)
    bsp.rhs_pack_bf16(ctx, n, 8, nrhs, Xd, Xb)
    # packing is the round-to-nearest-even BF16 of each entry, in block-major order
    want = bf16_bits(X.reshape(n, 8, nrhs).transpose(0, 2, 1).reshape(-1))
    np.testing.assert_array_equal(Xb.numpy(), want)
    Yb = ctx.empty((n * nrhs * 8,), np.float32)
    bsp.bsmm_rhs_bf16(ctx, A, nrhs, Xb, Yb)
    Y = ctx.empty((n * 8, nrhs))
    bsp.rhs_unpack_f32(ctx, n, 8, nrhs, Yb, Y)
    Ab = bf16_value(bf16_bits(vals))
    Xv = bf16_value(want).reshape(n, nrhs, 8).transpose(0, 2, 1)
    ref = _ref(rp, ci, Ab, Xv, n, nrhs).reshape(n * 8, nrhs)
    bound = _ref(rp, ci, np.abs(Ab), np.abs(Xv), n, nrhs).reshape(n * 8, nrhs)
    # FP32 accumulation of at most 7 x 8 products
    assert np.all(np.abs(Y.numpy() - ref) <= 8e-6 * bound + 1e-30)
    assert np.all(Y.numpy()[5 * 8:6 * 8] == 0.0)


def test_bf16_spmm_close_to_fp64_on_dg_matrix(ctx, kernel):
    p = pr.dg3d_lap(6, 6, 6, seed=4)
    A = bsp.BSR.from_problem(ctx, p)
    nrhs = 32
    X = np.random.default_rng(2).standard_normal((p.n * 8, nrhs))
    Xd = ctx.to_device(X)
    Y64 = ctx.empty(X.shape)
    bsp.bsmm_rhs(ctx, A, nrhs, Xd, Y64)
    Xb = ctx.empty((p.n * nrhs * 8,), np.uint16)
    bsp.rhs_pack_bf16(ctx, p.n, 8, nrhs, Xd, Xb)
    Yb = ctx.empty((p.n * nrhs * 8,), np.float32)
    bsp.bsmm_rhs_bf16(ctx, A, nrhs, Xb, Yb)
    Y = ctx.empty(X.shape)
    bsp.rhs_unpack_f32(ctx, p.n, 8, nrhs, Yb, Y)
    a, b = Y.numpy(), Y64.numpy()
    # BF16 operands: relative error of order 2^-8 per entry
    assert np.linalg.norm(a - b) <= 1e-2 * np.linalg.norm(b)


def test_bf16_spmm_rejects_bad_shapes(ctx):
    p = pr.stencil2d_b2(8, 8, seed=0)
    A = bsp.BSR.from_problem(ctx, p)
    Xb = ctx.empty((p.n * 16 * 2,), np.uint16)
    Yb = ctx.empty((p.n * 16 * 2,), np.float32)
    with pytest.raises(bsp.BspError):
        bsp.bsmm_rhs_bf16(ctx, A, 16, Xb, Yb)     # 2 x 2 blocks


def test_bf16_copy_follows_option_and_new_values(ctx):
    """The BF16 copy of the values is rebuilt when the bf16_ell option changes and after new
    values; both copies give the same products (same FP32 sums)."""
    rp, ci, vals = _random_bsr(40, seed=7, even=True)
    n, nrhs = 40, 32
    A = bsp.BSR(ctx, n, n, 8, rp, ci, vals)
    X = np.random.default_rng(4).standard_normal((n * 8, nrhs))
    Xb = ctx.empty((n * nrhs * 8,), np.uint16)
    bsp.rhs_pack_bf16(ctx, n, 8, nrhs, ctx.to_device(X), Xb)
    Yb = ctx.empty((n * nrhs * 8,), np.float32)
    outs = []
    try:
        for ell in (1, 0, 1):
            ctx.set_option("bf16_ell", ell)
            bsp.bsmm_rhs_bf16(ctx, A, nrhs, Xb, Yb)
            outs.append(Yb.numpy().copy())
        np.testing.assert_array_equal(outs[0], outs[1])
        np.testing.assert_array_equal(outs[0], outs[2])
        v2 = ctx.to_device(np.ascontiguousarray(2.0 * vals))
        ctx.check(ctx.L.bsp_bsr_set_values_dev(ctx.h, A.h, v2.ptr), "set_values")
        bsp.bsmm_rhs_bf16(ctx, A, nrhs, Xb, Yb)
        np.testing.assert_array_equal(Yb.numpy(), 2.0 * outs[0])
    finally:
        ctx.set_option("bf16_ell", 1)
