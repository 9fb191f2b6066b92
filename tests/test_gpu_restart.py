"""GMRES least squares on the device and the two routes of the deflated restart's eigenproblem.

The Givens least squares and the restart's solution update, residual vector and basis rotation
run on the device in every solve; the harmonic-Ritz eigenproblem (deflated_restart.hpp:66-245)
runs on the host by default and in one workgroup on the device with the option deflate_device.
Pins: the oracle's pGMRES (blocksp_solvers.hpp:156-295) on the same systems -- iteration counts
within max(2, 10%), solutions within 1e-6 -- and the mixed FGMRES reaching the FP64 residual
and the FP64 oracle solution, for both routes.
"""
import numpy as np
import pytest

import blocksp_amd as bsp
import pyoracle as po
from blocksp_amd import problems as pr

pytestmark = pytest.mark.gpu


@pytest.fixture(params=[0, 1], ids=["host_eig", "device_eig"])
def route(ctx, request):
    ctx.set_option("deflate_device", request.param)
    yield request.param
    ctx.set_option("deflate_device", 0)


@pytest.mark.parametrize("restart,k", [(30, 10), (20, 4), (12, 1)])
def test_pgmres_restarts_match_oracle(ctx, route, restart, k):
    print(f"pgmres dg2d restart {restart} k {k} device eig {route}", flush=True)
    p = pr.dg2d_veclap(24, 20, seed=1)
    A = bsp.BSR.from_problem(ctx, p)
    M = bsp.ILU0(ctx, A)
    xd, st, rc = bsp.pGMRES(ctx, A, ctx.to_device(p.b), 1e-10, M, maxIters=3000, restart=restart, k=k)
    xo, sto = po.pgmres(p, p.b, 1e-10, po.PRE_ILU0, maxIters=3000, restart=restart, k=k)
    assert rc == 0 and st.restarts >= 1, (rc, st.restarts)
    assert abs(st.iters - sto.iters) <= max(2, 0.1 * sto.iters), (st.iters, sto.iters)
    assert np.linalg.norm(xd.numpy() - xo) <= 1e-6 * np.linalg.norm(xo)


def test_c1_sor_pgmres_device_eig(ctx, route):
    p = pr.stencil2d_b2(64, 64, seed=0)
    xo, sto = po.pgmres(p, p.b, 1e-10, po.PRE_SOR, maxIters=5000, restart=30, k=10, w=1.2, nu=3)
    A = bsp.BSR.from_problem(ctx, p)
    M = bsp.SOR(ctx, A, 1.2, 3, 0)
    xd, st, rc = bsp.pGMRES(ctx, A, ctx.to_device(p.b), 1e-10, M, maxIters=5000, restart=30, k=10)
    assert rc == 0
    assert abs(st.iters - sto.iters) <= max(2, 0.1 * sto.iters), (st.iters, sto.iters)
    assert np.linalg.norm(xd.numpy() - xo) <= 1e-6 * np.linalg.norm(xo)


def test_fgmres_mixed_restarts_both_routes(ctx, route):
    dims = (48, 40)
    p = pr.dg2d_veclap(*dims, seed=1)
    A = bsp.BSR.from_problem(ctx, p)
    M = bsp.ILU0Ex(ctx, A, pr.parity_colours(dims), bsp.PREC_BF16)
    xd, st, rc = bsp.FGMRES_mixed(ctx, A, ctx.to_device(p.b), 1e-10, M, maxIters=5000, restart=30, k=10)
    x = xd.numpy()
    relres = np.linalg.norm(p.b - po.bsmv(p, x)) / np.linalg.norm(p.b)
    assert rc == 0 and st.restarts >= 3 and relres <= 1e-10, (rc, st.restarts, relres)
    xo, _ = po.pgmres(p, p.b, 1e-12, po.PRE_ILU0, maxIters=5000, restart=30, k=10)
    assert np.linalg.norm(x - xo) <= 1e-7 * np.linalg.norm(xo)


def test_pure_restart_k0(ctx):
    p = pr.dg2d_veclap(16, 14, seed=3)
    A = bsp.BSR.from_problem(ctx, p)
    M = bsp.ILU0(ctx, A)
    xd, st, rc = bsp.pGMRES(ctx, A, ctx.to_device(p.b), 1e-10, M, maxIters=3000, restart=10, k=0)
    xo, sto = po.pgmres(p, p.b, 1e-10, po.PRE_ILU0, maxIters=3000, restart=10, k=0)
    assert rc == 0 and abs(st.iters - sto.iters) <= max(2, 0.1 * sto.iters), (st.iters, sto.iters)
    assert np.linalg.norm(xd.numpy() - xo) <= 1e-6 * np.linalg.norm(xo)


def test_option_is_checked(ctx):
    with pytest.raises(bsp.BspError):
        ctx.set_option("deflate_device", 2)
