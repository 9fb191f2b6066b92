"""Triangular sweeps over the slot-ordered ELL copy of a factor against the same sweeps over
the CSR factor: the ELL copy keeps each row's blocks in CSR order and skips its padding, so
applies and whole solves must agree bitwise (context option sweep_ell = 0 / 1).  The CSR
sweeps themselves are checked against the oracle in test_gpu_parity.py."""
import numpy as np
import pytest

import blocksp_amd as bsp
from blocksp_amd import problems as pr

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _level_launches(ctx):
    """These systems are small enough for one-workgroup sweeps, which read CSR: launch the
    sweeps level by level so that the ELL copy is the one read."""
    ctx.set_option("sweep_onewg_max", 0)
    yield
    ctx.set_option("sweep_onewg_max", 1 << 22)


def _apply(ctx, M, r, ell):
    ctx.set_option("sweep_ell", ell)
    try:
        rd = ctx.to_device(r)
        M(rd)
        return rd.numpy()
    finally:
        ctx.set_option("sweep_ell", 1)


_COL = pr.parity_colours((40, 36))

CASES = {
    "dg2d_ilu0": lambda ctx, A: bsp.ILU0(ctx, A),
    "dg2d_ilu0_f32": lambda ctx, A: bsp.ILU0Ex(ctx, A, precision=bsp.PREC_F32),
    "dg2d_ilu0_bf16": lambda ctx, A: bsp.ILU0Ex(ctx, A, precision=bsp.PREC_BF16),
    "dg2d_ichol0": lambda ctx, A: bsp.IChol0(ctx, A),
    "dg3d_ichol0": lambda ctx, A: bsp.IChol0(ctx, A),
    "stencil_ilu0": lambda ctx, A: bsp.ILU0(ctx, A),
}


def _problem(case):
    if case.startswith("dg2d"):
        return pr.dg2d_veclap(40, 36, seed=1)
    if case.startswith("dg3d"):
        return pr.dg3d_lap(12, 10, 11, seed=2)
    return pr.stencil2d_b2(48, 40, seed=0)


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("case", sorted(CASES))
def test_ell_apply_equals_csr_apply_bitwise(ctx, case, mode):
    print(f"sweep ell {case} trsv={mode}", flush=True)
    p = _problem(case)
    A = bsp.BSR.from_problem(ctx, p)
    ctx.set_option("trsv", mode)
    try:
        M = CASES[case](ctx, A)
        assert M.ell_width(True) > 0 and M.ell_width(False) > 0
        r = np.random.default_rng(5).standard_normal(p.n * p.bs)
        z_ell = _apply(ctx, M, r, 1)
        z_csr = _apply(ctx, M, r, 0)
        np.testing.assert_array_equal(z_ell, z_csr)
        assert np.isfinite(z_ell).all() and np.linalg.norm(z_ell) > 0
    finally:
        ctx.set_option("trsv", 0)


def test_ell_pgmres_and_pcg_iterates_equal_csr(ctx):
    p = pr.dg2d_veclap(40, 36, seed=3)
    A = bsp.BSR.from_problem(ctx, p)
    b = ctx.to_device(p.b)
    out = {}
    for ell in (1, 0):
        ctx.set_option("sweep_ell", ell)
        try:
            M = bsp.ILU0(ctx, A)
            x, st, _ = bsp.pGMRES(ctx, A, b, 1e-10, M, maxIters=4000, restart=30, k=10)
            Mc = bsp.IChol0(ctx, A)
            xc, stc, _ = bsp.pCG(ctx, A, b, 1e-10, Mc, maxIters=4000)
            out[ell] = (st.iters, x.numpy(), stc.iters, xc.numpy())
        finally:
            ctx.set_option("sweep_ell", 1)
    assert out[1][0] == out[0][0] and out[1][2] == out[0][2]
    np.testing.assert_array_equal(out[1][1], out[0][1])
    np.testing.assert_array_equal(out[1][3], out[0][3])


@pytest.mark.parametrize("kind", ["ilu0", "ilu0_bf16", "ichol0"])
def test_few_level_sweeps_keep_csr(ctx, kind):
    """Parity-coloured factors sweep in 3 levels (bandwidth-bound): no ELL copy."""
    p = pr.dg2d_veclap(40, 36, seed=1)
    A = bsp.BSR.from_problem(ctx, p)
    M = {"ilu0": lambda: bsp.ILU0Ex(ctx, A, colour=_COL),
         "ilu0_bf16": lambda: bsp.ILU0Ex(ctx, A, colour=_COL, precision=bsp.PREC_BF16),
         "ichol0": lambda: bsp.IChol0(ctx, A, colour=_COL)}[kind]()
    assert M.ell_width(True) == 0 and M.ell_width(False) == 0


def test_padded_or_wide_sweeps_keep_csr(ctx):
    """Rows 0-9 form a chain (10 levels); rows 10.. all depend on row 9, and row 10 also on rows
    6-8: a level whose widest row has 4 strict blocks while the others have 1 (padding far above
    a quarter), and an upper sweep whose row 9 has n - 10 blocks.  Neither sweep gets an ELL
    copy; the apply still runs."""
    n, bs = 40, 2
    nb = {i: set() for i in range(n)}

    def couple(i, j):
        nb[i].add(j)
        nb[j].add(i)

    for i in range(1, 10):
        couple(i, i - 1)
    for i in range(10, n):
        couple(i, 9)
    for j in (6, 7, 8):
        couple(10, j)
    rowptr, colidx, vals = [0], [], []
    for i in range(n):
        for j in sorted(nb[i] | {i}):
            colidx.append(j)
            vals.append((len(nb[i]) + 1.0) * np.eye(bs) if j == i else -0.5 * np.eye(bs))
        rowptr.append(len(colidx))
    p = pr.Problem(n, bs, bs, 1, np.array(rowptr, np.int32), np.array(colidx, np.int32), np.array(vals),
                   np.ones(n * bs), name="padded")
    A = bsp.BSR.from_problem(ctx, p)
    M = bsp.ILU0(ctx, A)
    assert M.levels(True)[0] > 8
    assert M.ell_width(True) == 0 and M.ell_width(False) == 0
    z = _apply(ctx, M, np.ones(n * bs), 1)
    assert np.isfinite(z).all()
