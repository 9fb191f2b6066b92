"""Sweeps of small systems in one workgroup (context option sweep_onewg_max) against the same
sweeps launched level by level: every row of a level is independent of the others, so the
applies and whole solves must agree bitwise.  Level launches are checked against the oracle in
test_gpu_parity.py."""
import numpy as np
import pytest

import blocksp_amd as bsp
from blocksp_amd import problems as pr

pytestmark = pytest.mark.gpu


def _with(ctx, onewg, fn):
    ctx.set_option("sweep_onewg_max", onewg)
    try:
        return fn()
    finally:
        ctx.set_option("sweep_onewg_max", 1 << 22)


PRECONDS = {
    "ilu0": lambda ctx, A, p: bsp.ILU0(ctx, A),
    "ilu0_f32": lambda ctx, A, p: bsp.ILU0Ex(ctx, A, precision=bsp.PREC_F32),
    "ilu0_bf16_coloured": lambda ctx, A, p: bsp.ILU0Ex(ctx, A, colour=pr.parity_colours(p.meta["dims"]),
                                                      precision=bsp.PREC_BF16),
    "ichol0": lambda ctx, A, p: bsp.IChol0(ctx, A),
    "sor_sym": lambda ctx, A, p: bsp.SOR(ctx, A, 1.2, 3, 0),
    "sor_fwd": lambda ctx, A, p: bsp.SOR(ctx, A, 0.9, 2, 1),
    "sor_multicolour": lambda ctx, A, p: bsp.SOR(ctx, A, 1.2, 3, 0, colour=pr.parity_colours(p.meta["dims"])),
}


def _problem(kind):
    if kind.startswith("ichol0") or kind.startswith("ilu0"):
        return pr.dg2d_veclap(20, 18, seed=2)
    return pr.stencil2d_b2(40, 36, seed=1)


@pytest.mark.parametrize("kind", sorted(PRECONDS))
def test_onewg_apply_equals_level_launches_bitwise(ctx, kind):
    print(f"onewg {kind}", flush=True)
    p = _problem(kind)
    A = bsp.BSR.from_problem(ctx, p)
    M = PRECONDS[kind](ctx, A, p)
    r = np.random.default_rng(9).standard_normal(p.n * p.bs)
    out = []
    for onewg in (0, 1 << 20):
        def run():
            rd = ctx.to_device(r)
            M(rd)
            return rd.numpy()
        out.append(_with(ctx, onewg, run))
    np.testing.assert_array_equal(out[0], out[1])
    assert np.isfinite(out[0]).all() and np.linalg.norm(out[0]) > 0


def test_onewg_solves_equal_level_launches_bitwise(ctx):
    p = pr.stencil2d_b2(48, 40, seed=3)
    A = bsp.BSR.from_problem(ctx, p)
    b = ctx.to_device(p.b)
    res = []
    for onewg in (0, 1 << 20):
        def run():
            M = bsp.SOR(ctx, A, 1.2, 3, 0)
            x, st, rc = bsp.pGMRES(ctx, A, b, 1e-10, M, maxIters=3000, restart=30, k=10)
            Mi = bsp.ILU0(ctx, A)
            xi, sti, rci = bsp.pGMRES(ctx, A, b, 1e-10, Mi, maxIters=3000, restart=30, k=10)
            return st.iters, x.numpy(), sti.iters, xi.numpy()
        res.append(_with(ctx, onewg, run))
    assert res[0][0] == res[1][0] and res[0][2] == res[1][2]
    np.testing.assert_array_equal(res[0][1], res[1][1])
    np.testing.assert_array_equal(res[0][3], res[1][3])


def _blocklap(B, n, seed):
    return pr.batch_blocklap(nsys=1, lo=n, hi=n, seed=seed, B=B, nrhs=1)[0]


AHEAD_PRECONDS = {
    "ilu0": lambda ctx, A: bsp.ILU0(ctx, A),
    "ilu0_f32": lambda ctx, A: bsp.ILU0Ex(ctx, A, precision=bsp.PREC_F32),
    "ilu0_bf16": lambda ctx, A: bsp.ILU0Ex(ctx, A, precision=bsp.PREC_BF16),
    "ichol0": lambda ctx, A: bsp.IChol0(ctx, A),
    "sor_sym": lambda ctx, A: bsp.SOR(ctx, A, 1.2, 3, 0),
    "sor_fwd": lambda ctx, A: bsp.SOR(ctx, A, 0.9, 1, 1),
    "sor_bwd": lambda ctx, A: bsp.SOR(ctx, A, 1.1, 2, -1),
}
AHEAD_PROBLEMS = {
    "b1": lambda: _blocklap(1, 3000, 11),
    "b2": lambda: pr.stencil2d_b2(44, 40, seed=4),
    "b3": lambda: _blocklap(3, 1800, 12),
    "b4": lambda: _blocklap(4, 2200, 13),
}


@pytest.mark.parametrize("size", sorted(AHEAD_PROBLEMS))
@pytest.mark.parametrize("kind", sorted(AHEAD_PRECONDS))
def test_onewg_loads_ahead_bitwise(ctx, kind, size):
    """One-workgroup sweeps over an ELL copy that load each level's operands before the previous
    level's barrier (sweep_ahead = 1), with the solution vector in LDS where it fits
    (sweep_lds = 1) or not, against the same sweeps loading the operands after the barrier."""
    print(f"onewg ahead {kind} {size}", flush=True)
    p = AHEAD_PROBLEMS[size]()
    A = bsp.BSR.from_problem(ctx, p)
    M = AHEAD_PRECONDS[kind](ctx, A)
    assert M.ell_width(True) > 0 and M.ell_width(False) > 0
    r = np.random.default_rng(10).standard_normal(p.n * p.bs)
    out = []
    for ahead, lds in ((1, 1), (1, 0), (0, 0)):
        ctx.set_option("sweep_ahead", ahead)
        ctx.set_option("sweep_lds", lds)
        try:
            rd = ctx.to_device(r)
            M(rd)
            out.append(rd.numpy())
        finally:
            ctx.set_option("sweep_ahead", 1)
            ctx.set_option("sweep_lds", 1)
    np.testing.assert_array_equal(out[0], out[2])
    np.testing.assert_array_equal(out[1], out[2])
    assert np.isfinite(out[0]).all() and np.linalg.norm(out[0]) > 0


@pytest.mark.parametrize("lds", [1, 0])
def test_onewg_partial_level_ranges(ctx, lds):
    """Without sweep graphs, ILU0 applies launch its sweeps over level ranges (the forward
    sweep's first level and the shared boundary level are split off), so one-workgroup sweeps
    start past level 0 and read rows of earlier launches: against level launches, bitwise."""
    p = pr.stencil2d_b2(40, 36, seed=6)
    A = bsp.BSR.from_problem(ctx, p)
    r = np.random.default_rng(12).standard_normal(p.n * p.bs)
    ctx.set_option("sweep_graphs", 0)
    ctx.set_option("sweep_lds", lds)
    try:
        M = bsp.ILU0(ctx, A)
        out = []
        for onewg in (0, 1 << 22):
            def run():
                rd = ctx.to_device(r)
                M(rd)
                return rd.numpy()
            out.append(_with(ctx, onewg, run))
        np.testing.assert_array_equal(out[0], out[1])
    finally:
        ctx.set_option("sweep_graphs", 1)
        ctx.set_option("sweep_lds", 1)


def test_onewg_level_wider_than_one_pass(ctx):
    """bs 3 on a 340 x 340 grid: levels of up to 340 rows (within 1024 / 3) pad to 357 slots,
    more than one pass of the workgroup holds (16 wavefronts x 21 rows), so the sweep takes the
    multi-pass one-workgroup kernel; it must still agree with level launches."""
    p = _blocklap(3, 340 * 340, 14)
    assert p.meta["dims"] == (340, 340)
    A = bsp.BSR.from_problem(ctx, p)
    r = np.random.default_rng(11).standard_normal(p.n * p.bs)
    for make in (lambda: bsp.ILU0(ctx, A), lambda: bsp.SOR(ctx, A, 1.2, 1, 0)):
        M = make()
        out = []
        for onewg in (0, 1 << 22):
            def run():
                rd = ctx.to_device(r)
                M(rd)
                return rd.numpy()
            out.append(_with(ctx, onewg, run))
        np.testing.assert_array_equal(out[0], out[1])
