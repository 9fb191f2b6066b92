"""SOR sweeps over the slot-ordered ELL copy of their off-diagonal blocks against the same
sweeps over CSR: the copy keeps each row's blocks in CSR order, tags each column with the side
it is read from (this sweep's new values or the previous iterate) and reads the values from the
matrix itself, so applies and whole solves must agree bitwise (context option sweep_ell = 0 / 1),
in every launch mode and after the matrix values change.  The CSR sweeps themselves are checked
against the oracle in test_gpu_parity.py."""
import numpy as np
import pytest

import blocksp_amd as bsp
from blocksp_amd import problems as pr

pytestmark = pytest.mark.gpu

MODES = {   # launch mode -> context options
    "onewg": {},
    "levels": {"sweep_onewg_max": 0},
    "syncfree": {"trsv": 1},
}
_DEFAULTS = {"sweep_onewg_max": 1 << 22, "trsv": 0, "sweep_ell": 1}


def _blocklap(B, n, seed):
    return pr.batch_blocklap(nsys=1, lo=n, hi=n, seed=seed, B=B, nrhs=1)[0]


PROBLEMS = {
    "stencil_b2": lambda: pr.stencil2d_b2(48, 40, seed=0),
    "blocklap_b1": lambda: _blocklap(1, 3000, 7),
    "blocklap_b3": lambda: _blocklap(3, 1500, 8),
    "blocklap_b4": lambda: _blocklap(4, 2000, 9),
}


def _set(ctx, opts):
    for k, v in opts.items():
        ctx.set_option(k, v)


def _apply(ctx, M, r, ell):
    ctx.set_option("sweep_ell", ell)
    try:
        rd = ctx.to_device(r)
        M(rd)
        return rd.numpy()
    finally:
        ctx.set_option("sweep_ell", 1)


@pytest.mark.parametrize("mode", sorted(MODES))
@pytest.mark.parametrize("case", sorted(PROBLEMS))
def test_sor_ell_apply_equals_csr_bitwise(ctx, case, mode):
    p = PROBLEMS[case]()
    A = bsp.BSR.from_problem(ctx, p)
    r = np.random.default_rng(3).standard_normal(p.n * p.bs)
    _set(ctx, MODES[mode])
    try:
        for typ, nu in ((0, 3), (1, 1), (-1, 2)):
            print(f"sor ell {case} mode={mode} type={typ} nu={nu}", flush=True)
            M = bsp.SOR(ctx, A, 1.2, nu, typ)
            assert M.ell_width(True) > 0 and M.ell_width(False) > 0
            z_ell = _apply(ctx, M, r, 1)
            z_csr = _apply(ctx, M, r, 0)
            np.testing.assert_array_equal(z_ell, z_csr)
            assert np.isfinite(z_ell).all() and np.linalg.norm(z_ell) > 0
    finally:
        _set(ctx, _DEFAULTS)


@pytest.mark.parametrize("mode", sorted(MODES))
def test_sor_ell_in_an_elimination_order(ctx, mode):
    """Colours that reverse the natural order: the forward sweep runs from the last row to the
    first, so the side each column is read from comes from the elimination steps, not from the
    column numbers."""
    p = pr.stencil2d_b2(40, 36, seed=2)
    A = bsp.BSR.from_problem(ctx, p)
    colour = np.arange(p.n, dtype=np.int32)[::-1].copy()
    r = np.random.default_rng(4).standard_normal(p.n * p.bs)
    _set(ctx, MODES[mode])
    try:
        M = bsp.SOR(ctx, A, 1.1, 2, 0, colour=colour)
        assert M.ell_width(True) > 0 and M.ell_width(False) > 0
        z_ell = _apply(ctx, M, r, 1)
        z_csr = _apply(ctx, M, r, 0)
        np.testing.assert_array_equal(z_ell, z_csr)
        # against natural-order SOR it must differ: the order matters
        z_nat = _apply(ctx, bsp.SOR(ctx, A, 1.1, 2, 0), r, 1)
        assert not np.array_equal(z_ell, z_nat)
    finally:
        _set(ctx, _DEFAULTS)


def test_sor_ell_reads_the_matrix_values_live(ctx):
    """Values set on the matrix after the preconditioner was made are used by the sweeps over the
    copy exactly as by the CSR sweeps (the diagonal inverses stay those computed at creation in
    both)."""
    p = pr.stencil2d_b2(48, 40, seed=5)
    A = bsp.BSR.from_problem(ctx, p)
    M = bsp.SOR(ctx, A, 1.2, 3, 0)
    assert M.ell_width(True) > 0
    r = np.random.default_rng(6).standard_normal(p.n * p.bs)
    before = _apply(ctx, M, r, 1)
    v2 = 0.5 * p.vals + 0.01
    v2_d = ctx.to_device(np.ascontiguousarray(v2))     # kept alive across the calls
    ctx.check(ctx.L.bsp_bsr_set_values_dev(ctx.h, A.h, v2_d.ptr), "set_values")
    z_ell = _apply(ctx, M, r, 1)
    z_csr = _apply(ctx, M, r, 0)
    np.testing.assert_array_equal(z_ell, z_csr)
    assert not np.array_equal(z_ell, before)


def test_sor_ell_pgmres_iterates_equal_csr(ctx):
    """The 64 x 64 block stencil (B=2) with SOR(1.2, nu 3)-pGMRES(30, k 10) to 1e-10: the same
    iteration count and bitwise the same solution with and without the copy."""
    p = pr.stencil2d_b2(64, 64, seed=0)
    A = bsp.BSR.from_problem(ctx, p)
    b = ctx.to_device(p.b)
    out = {}
    for ell in (1, 0):
        ctx.set_option("sweep_ell", ell)
        try:
            M = bsp.SOR(ctx, A, 1.2, 3, 0)
            x, st, rc = bsp.pGMRES(ctx, A, b, 1e-10, M, maxIters=5000, restart=30, k=10)
            assert rc == 0
            out[ell] = (st.iters, x.numpy())
        finally:
            ctx.set_option("sweep_ell", 1)
    assert out[1][0] == out[0][0]
    np.testing.assert_array_equal(out[1][1], out[0][1])


def test_sor_large_blocks_and_few_levels_keep_csr(ctx):
    """Blocks above 4 x 4 (the DG matrix, bs 8) and parity-coloured SOR (2 levels) have no copy."""
    p = pr.dg2d_veclap(16, 12, seed=1)
    A = bsp.BSR.from_problem(ctx, p)
    M = bsp.SOR(ctx, A, 1.2, 1, 0)
    assert M.ell_width(True) == 0 and M.ell_width(False) == 0
    q = pr.stencil2d_b2(32, 30, seed=1)
    Aq = bsp.BSR.from_problem(ctx, q)
    Mq = bsp.SOR(ctx, Aq, 1.2, 1, 0, colour=pr.parity_colours((32, 30)))
    assert Mq.ell_width(True) == 0 and Mq.ell_width(False) == 0
