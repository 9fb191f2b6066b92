"""The modified Gram-Schmidt of an Arnoldi step on short vectors (at most 8192 entries: at most 4
block partials) in one single-workgroup kernel against one launch per step (context option
mgs_onewg): each thread group does what one block of the per-step kernels does and the partials
are summed in the same order, so whole pGMRES solves must agree bitwise.  The per-step kernels
are checked against the oracle in test_gpu_parity.py."""
import numpy as np
import pytest

import blocksp_amd as bsp
from blocksp_amd import problems as pr

pytestmark = pytest.mark.gpu


def _blocklap(B, n, seed):
    return pr.batch_blocklap(nsys=1, lo=n, hi=n, seed=seed, B=B, nrhs=1)[0]


CASES = {   # len = n * bs: 8192 and 2048 (whole tiles: w kept in registers), 2880 (2 partials, last
            # tile partial), 2997 (odd: misaligned basis vectors, scalar path)
    "stencil_64x64_b2": lambda: pr.stencil2d_b2(64, 64, seed=0),
    "stencil_32x32_b2": lambda: pr.stencil2d_b2(32, 32, seed=2),
    "stencil_40x36_b2": lambda: pr.stencil2d_b2(40, 36, seed=1),
    "blocklap_b3_odd": lambda: _blocklap(3, 1000, 5),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_mgs_onewg_pgmres_bitwise(ctx, case):
    print(f"mgs onewg {case}", flush=True)
    p = CASES[case]()
    A = bsp.BSR.from_problem(ctx, p)
    b = ctx.to_device(p.b)
    out = {}
    for one in (1, 0):
        ctx.set_option("mgs_onewg", one)
        try:
            M = bsp.SOR(ctx, A, 1.2, 2, 0)
            x, st, rc = bsp.pGMRES(ctx, A, b, 1e-10, M, maxIters=3000, restart=30, k=10)
            assert rc == 0
            out[one] = (st.iters, x.numpy())
        finally:
            ctx.set_option("mgs_onewg", 1)
    assert out[1][0] == out[0][0]
    np.testing.assert_array_equal(out[1][1], out[0][1])
