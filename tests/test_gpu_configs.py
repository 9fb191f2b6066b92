"""The named workloads at oracle-checkable sizes, on the GPU against the CPU oracle.

* 3D DG Laplacian (p=1 hex, bs 8, 7 blocks per row): mixed-precision FGMRES with the
  parity-coloured ILU0 stored in FP32 and in BF16 reaches the true FP64 relative residual
  1e-10 and the FP64 oracle's solution to a forward error of 1e-7.
* The same operator's BF16 MFMA block SpMM with 32 right-hand sides against a numpy statement
  of the BF16 arithmetic (bit-exact packing, FP32 accumulation bound).
* The reference's CPU-sized case at full size: 64 x 64 block stencil (B=2, 4096 block rows),
  SOR(w 1.2, nu 3)-pGMRES(30, k 10) to 1e-10 against the oracle's pGMRES with the same
  preconditioner (blocksp_solvers.hpp:156-295, blocksp_preconditioners.hpp:116-240).
"""
import numpy as np
import pytest

import blocksp_amd as bsp
import pyoracle as po
from blocksp_amd import problems as pr

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dg3d_12():
    p = pr.dg3d_lap(12, 12, 12, seed=4)
    xo, sto = po.pgmres(p, p.b, 1e-13, po.PRE_ILU0, maxIters=5000, restart=30, k=10)
    assert sto.true_relres <= 1e-12
    return p, xo


@pytest.mark.parametrize("prec", [bsp.PREC_F32, bsp.PREC_BF16])
def test_dg3d_mixed_fgmres_matches_fp64_oracle(ctx, dg3d_12, prec):
    p, xo = dg3d_12
    print(f"dg3d 12^3 mixed FGMRES factor precision {prec}", flush=True)
    A = bsp.BSR.from_problem(ctx, p)
    M = bsp.ILU0Ex(ctx, A, pr.parity_colours((12, 12, 12)), prec)
    assert M.levels(True)[0] == 4 and M.levels(False)[0] == 4
    xd, st, rc = bsp.FGMRES_mixed(ctx, A, ctx.to_device(p.b), 1e-10, M, maxIters=5000, restart=30, k=10,
                                  spmv_precision=bsp.PREC_F32)
    x = xd.numpy()
    relres = np.linalg.norm(p.b - po.bsmv(p, x)) / np.linalg.norm(p.b)
    assert rc == 0 and st.converged == 1 and relres <= 1e-10, (rc, relres, st.iters)
    assert abs(st.true_relres - relres) <= 1e-2 * relres + 1e-16
    assert np.linalg.norm(x - xo) <= 1e-7 * np.linalg.norm(xo)


def _bf16_bits(x):
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return u.astype(np.uint16)


def _bf16_value(bits):
    return (bits.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


@pytest.mark.parametrize("ell", [1, 0])
def test_dg3d_bf16_spmm_32rhs_matches_numpy_bf16(ctx, ell):
    ctx.set_option("bf16_ell", ell)
    p = pr.dg3d_lap(8, 8, 8, seed=4)
    n, nrhs = p.n, 32
    A = bsp.BSR.from_problem(ctx, p)
    X = np.random.default_rng(3).standard_normal((n * 8, nrhs))
    Xd = ctx.to_device(X)
    Xb = ctx.empty((n * nrhs * 8,), np.uint16)
    bsp.rhs_pack_bf16(ctx, n, 8, nrhs, Xd, Xb)
    want = _bf16_bits(X.reshape(n, 8, nrhs).transpose(0, 2, 1).reshape(-1))
    np.testing.assert_array_equal(Xb.numpy(), want)
    Yb = ctx.empty((n * nrhs * 8,), np.float32)
    bsp.bsmm_rhs_bf16(ctx, A, nrhs, Xb, Yb)
    Y = ctx.empty((n * 8, nrhs))
    bsp.rhs_unpack_f32(ctx, n, 8, nrhs, Yb, Y)
    Ab = _bf16_value(_bf16_bits(p.vals))
    Xv = _bf16_value(want).reshape(n, nrhs, 8).transpose(0, 2, 1)     # (n, 8, nrhs)
    rows = np.repeat(np.arange(n), np.diff(p.rowptr))
    ref = np.zeros((n, 8, nrhs))
    bound = np.zeros((n, 8, nrhs))
    np.add.at(ref, rows, np.einsum("prc,pck->prk", Ab, Xv[p.colidx]))
    np.add.at(bound, rows, np.einsum("prc,pck->prk", np.abs(Ab), np.abs(Xv[p.colidx])))
    got = Y.numpy().reshape(n, 8, nrhs)
    ctx.set_option("bf16_ell", 1)
    # FP32 accumulation of 7 blocks x 8 products of exactly representable BF16 operands
    assert np.all(np.abs(got - ref) <= 8e-6 * bound + 1e-30)


def test_c1_stencil64_sor_pgmres_matches_oracle(ctx):
    p = pr.stencil2d_b2(64, 64, seed=0)
    xo, sto = po.pgmres(p, p.b, 1e-10, po.PRE_SOR, maxIters=5000, restart=30, k=10, w=1.2, nu=3)
    A = bsp.BSR.from_problem(ctx, p)
    M = bsp.SOR(ctx, A, 1.2, 3, 0)
    xd, st, rc = bsp.pGMRES(ctx, A, ctx.to_device(p.b), 1e-10, M, maxIters=5000, restart=30, k=10)
    x = xd.numpy()
    assert rc == 0
    assert abs(st.iters - sto.iters) <= max(2, 0.1 * sto.iters), (st.iters, sto.iters)
    assert np.linalg.norm(x - xo) <= 1e-6 * np.linalg.norm(xo)
    relres = np.linalg.norm(p.b - po.bsmv(p, x)) / np.linalg.norm(p.b)
    assert relres <= 10 * max(sto.true_relres, 1e-12), (relres, sto.true_relres)
