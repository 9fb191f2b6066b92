"""Batched IChol0-pCG over independent systems in one persistent launch."""
import numpy as np
import pytest

import blocksp_amd as bsp
from blocksp_amd import problems as pr
import pyoracle as po

pytestmark = pytest.mark.gpu


def _batch(ctx, nsys=12, lo=40, hi=1500, seed=2, nrhs=4):
    probs = pr.batch_blocklap(nsys=nsys, lo=lo, hi=hi, seed=seed, nrhs=nrhs)
    rp, ci, vals, sys = bsp.block_diagonal(probs)
    n = int(sys[-1])
    A = bsp.BSR(ctx, n, n, 4, rp, ci, vals)
    Bh = np.stack([np.concatenate([p.meta["rhs"][k] for p in probs]) for k in range(nrhs)])
    return probs, A, sys, Bh


def _solution_of(X, probs, sys, s, k):
    return X[k, sys[s] * 4:sys[s + 1] * 4]


def test_reference_rule_matches_oracle_pcg(ctx):
    probs, A, sys, Bh = _batch(ctx)
    bat = bsp.Batch(ctx, A, sys, 4)
    B = ctx.to_device(Bh)
    X = ctx.zeros(Bh.shape)
    its, rel, ms, rc = bat.pcg(B, X, 1e-10, maxIters=2000, tol_mode=bsp.TOL_REFERENCE)
    assert rc == 0
    Xh = X.numpy()
    for s, p in enumerate(probs):
        for k in range(4):
            b = p.meta["rhs"][k]
            xo, st = po.pcg(p, b, 1e-10, po.PRE_ICHOL0, maxIters=2000)
            x = _solution_of(Xh, probs, sys, s, k)
            assert abs(int(its[s, k]) - st.iters) <= 1, (s, k, its[s, k], st.iters)
            assert np.linalg.norm(x - xo) <= 1e-8 * np.linalg.norm(xo)
            assert np.linalg.norm(b - p.bsmv(x)) <= 1e-8 * np.linalg.norm(b)


def test_true_residual_rule_and_device_pcg_agree(ctx):
    probs, A, sys, Bh = _batch(ctx, nsys=6, seed=5)
    bat = bsp.Batch(ctx, A, sys, 4)
    B = ctx.to_device(Bh)
    X = ctx.zeros(Bh.shape)
    its, rel, ms, rc = bat.pcg(B, X, 1e-10, maxIters=2000, tol_mode=bsp.TOL_TRUE_RESIDUAL)
    assert rc == 0 and (rel <= 1e-10).all()
    Xh = X.numpy()
    for s, p in enumerate(probs):
        Ai = bsp.BSR.from_problem(ctx, p)
        M = bsp.IChol0(ctx, Ai)
        for k in range(4):
            b = p.meta["rhs"][k]
            x = _solution_of(Xh, probs, sys, s, k)
            assert np.linalg.norm(b - p.bsmv(x)) <= 2e-10 * np.linalg.norm(b)
            xd, st, _ = bsp.pCG(ctx, Ai, ctx.to_device(b), 1e-10, M, maxIters=2000)
            assert abs(int(its[s, k]) - st.iters) <= 2


@pytest.mark.parametrize("wg,r", [(64, 1), (256, 2), (512, 4), (1024, 4), (128, 4)])
def test_tiles_give_the_same_solutions(ctx, wg, r):
    probs, A, sys, Bh = _batch(ctx, nsys=8, seed=7)
    bat = bsp.Batch(ctx, A, sys, 4)
    B = ctx.to_device(Bh)
    X0 = ctx.zeros(Bh.shape)
    its0, _, _, _ = bat.pcg(B, X0, 1e-10, maxIters=2000)
    bat.set_tiles(wg, r)
    X1 = ctx.zeros(Bh.shape)
    its1, _, _, rc = bat.pcg(B, X1, 1e-10, maxIters=2000)
    assert rc == 0
    assert np.abs(its1 - its0).max() <= 1
    a, b = X0.numpy(), X1.numpy()
    assert np.linalg.norm(a - b) <= 1e-9 * np.linalg.norm(a)


def test_graph_replay_is_bitwise_repeatable_and_follows_new_rhs(ctx):
    probs, A, sys, Bh = _batch(ctx, nsys=5, seed=9)
    bat = bsp.Batch(ctx, A, sys, 4)
    B = ctx.to_device(Bh)
    X = ctx.zeros(Bh.shape)
    bat.pcg(B, X, 1e-10)
    x1 = X.numpy()
    X.zero()
    bat.pcg(B, X, 1e-10)              # replay of the captured graph
    np.testing.assert_array_equal(X.numpy(), x1)
    B.upload(-2.0 * Bh)               # same buffers, new values: the replay reads them
    bat.pcg(B, X, 1e-10)
    np.testing.assert_allclose(X.numpy(), -2.0 * x1, rtol=0, atol=1e-9 * np.abs(x1).max())


def test_autotune_picks_a_valid_tile_and_caches_it(ctx):
    probs, A, sys, Bh = _batch(ctx, nsys=10, seed=11)
    bat = bsp.Batch(ctx, A, sys, 4)
    B = ctx.to_device(Bh)
    X = ctx.zeros(Bh.shape)
    wg, r = bat.autotune(B, X, probe_iters=4)
    assert wg in (128, 256, 512, 1024) and r in (1, 2, 4)
    bat2 = bsp.Batch(ctx, A, sys, 4)
    assert bat2.autotune(B, X, probe_iters=4) == (wg, r)
    its, rel, ms, rc = bat2.pcg(B, X, 1e-10, maxIters=2000)
    assert rc == 0


def test_coupled_systems_are_rejected(ctx):
    probs, A, sys, Bh = _batch(ctx, nsys=3, seed=1)
    bad = sys.copy()
    bad[1] += 1                        # moves a row of system 0 into system 1
    with pytest.raises(bsp.BspError):
        bsp.Batch(ctx, A, bad, 4)


@pytest.mark.parametrize("wg,r", [(256, 4), (1024, 2), (128, 1)])
def test_batched_equals_one_launch_per_system_bitwise(ctx, wg, r):
    """The batched kernel's per-item work depends only on that system and the tiles: each
    system launched on its own gives bitwise the same solution."""
    probs, A, sys, Bh = _batch(ctx, nsys=6, seed=13)
    bat = bsp.Batch(ctx, A, sys, 4)
    bat.set_tiles(wg, r)
    B = ctx.to_device(Bh)
    X = ctx.zeros(Bh.shape)
    its, _, _, rc = bat.pcg(B, X, 1e-10, maxIters=2000)
    assert rc == 0
    Xh = X.numpy()
    for s, p in enumerate(probs):
        As = bsp.BSR.from_problem(ctx, p)
        one = bsp.Batch(ctx, As, np.array([0, p.n], np.int32), 4)
        one.set_tiles(wg, r)
        Bs = ctx.to_device(np.stack([p.meta["rhs"][k] for k in range(4)]))
        Xs = ctx.zeros(Bs.shape)
        its1, _, _, rc1 = one.pcg(Bs, Xs, 1e-10, maxIters=2000)
        assert rc1 == 0
        np.testing.assert_array_equal(its1[0], its[s])
        np.testing.assert_array_equal(Xs.numpy(), Xh[:, sys[s] * 4:sys[s + 1] * 4])


def _permuted(p, perm):
    """The problem with rows and columns renumbered: new row t = old row perm[t], each row's
    blocks in its stored order with the columns renumbered (P A P^T)."""
    inv = np.empty_like(perm)
    inv[perm] = np.arange(len(perm))
    rp, ci, vals = [0], [], []
    for old in perm:
        ci.append(inv[p.colidx[p.rowptr[old]:p.rowptr[old + 1]]])
        vals.append(p.vals[p.rowptr[old]:p.rowptr[old + 1]])
        rp.append(rp[-1] + len(ci[-1]))
    return pr.Problem(p.n, p.bs, p.B, p.D, np.array(rp, np.int32), np.concatenate(ci).astype(np.int32),
                      np.concatenate(vals), p.b[np.repeat(perm * p.bs, p.bs) + np.tile(np.arange(p.bs), p.n)])


def test_coloured_ichol0_is_the_ichol0_of_the_permuted_matrix(ctx):
    # the colour-ordered IChol0 applied in A's numbering equals the oracle's natural IChol0 of
    # P A P^T applied to the permuted vector; its sweeps have 3 levels (2D parity colours)
    p = pr.batch_blocklap(nsys=1, lo=600, hi=700, seed=3)[0]
    col = pr.parity_colours(p.meta["dims"])
    perm = np.argsort(col, kind="stable")
    q = _permuted(p, perm)
    A = bsp.BSR.from_problem(ctx, p)
    M = bsp.IChol0(ctx, A, colour=col)
    r = np.random.default_rng(1).standard_normal(p.n * p.bs)
    lrp, lci, lv = po.ichol0_factor(q)
    rq = r.reshape(p.n, p.bs)[perm].reshape(-1)
    zq = po.ichol0_apply(q.n, q.bs, lrp, lci, lv, rq)
    rd = ctx.to_device(r)
    M(rd)
    z = rd.numpy().reshape(p.n, p.bs)
    np.testing.assert_allclose(z[perm].reshape(-1), zq, rtol=0, atol=1e-12 * np.abs(zq).max())


def test_coloured_batch_matches_oracle_pcg_on_permuted_systems(ctx):
    probs, A, sys, Bh = _batch(ctx, nsys=8, lo=60, hi=900, seed=4)
    col = np.concatenate([pr.parity_colours(p.meta["dims"]) for p in probs]).astype(np.int32)
    bat = bsp.Batch(ctx, A, sys, 4, colour=col)
    B = ctx.to_device(Bh)
    X = ctx.zeros(Bh.shape)
    its, rel, ms, rc = bat.pcg(B, X, 1e-10, maxIters=2000, tol_mode=bsp.TOL_REFERENCE)
    assert rc == 0
    Xh = X.numpy()
    for s, p in enumerate(probs):
        perm = np.argsort(pr.parity_colours(p.meta["dims"]), kind="stable")
        q = _permuted(p, perm)
        for k in range(4):
            b = p.meta["rhs"][k]
            bq = b.reshape(p.n, p.bs)[perm].reshape(-1)
            xo, st = po.pcg(q, bq, 1e-10, po.PRE_ICHOL0, maxIters=2000)
            x = _solution_of(Xh, probs, sys, s, k)
            assert abs(int(its[s, k]) - st.iters) <= 1, (s, k, its[s, k], st.iters)
            xq = x.reshape(p.n, p.bs)[perm].reshape(-1)
            assert np.linalg.norm(xq - xo) <= 1e-8 * np.linalg.norm(xo)
            assert np.linalg.norm(b - p.bsmv(x)) <= 1e-8 * np.linalg.norm(b)


def test_device_pcg_with_coloured_ichol0_matches_oracle_on_permuted_system(ctx):
    p = pr.batch_blocklap(nsys=1, lo=900, hi=1000, seed=6)[0]
    perm = np.argsort(pr.parity_colours(p.meta["dims"]), kind="stable")
    q = _permuted(p, perm)
    A = bsp.BSR.from_problem(ctx, p)
    M = bsp.IChol0(ctx, A, colour=pr.parity_colours(p.meta["dims"]))
    b = p.meta["rhs"][1]
    x, st, rc = bsp.pCG(ctx, A, ctx.to_device(b), 1e-10, M, maxIters=2000)
    bq = b.reshape(p.n, p.bs)[perm].reshape(-1)
    xo, sto = po.pcg(q, bq, 1e-10, po.PRE_ICHOL0, maxIters=2000)
    assert rc == 0 and abs(st.iters - sto.iters) <= 1, (st.iters, sto.iters)
    xq = x.numpy().reshape(p.n, p.bs)[perm].reshape(-1)
    assert np.linalg.norm(xq - xo) <= 1e-8 * np.linalg.norm(xo)


def _grid_problem(dims, offsets, bs, rng, nrhs=4):
    """SPD system on a grid: block (i, j) = -c I for every neighbour j = i + offset inside the
    grid (offsets closed under negation), diagonal block c ((deg + 1) I + S S^T / bs), which
    is diagonally dominant.  Rows keep their blocks in increasing column order."""
    dims = tuple(dims)
    n = int(np.prod(dims))
    coords = np.stack(np.unravel_index(np.arange(n), dims), axis=1)
    c = float(np.exp(0.5 * rng.standard_normal()))
    S = rng.standard_normal((bs, bs))
    rowptr, colidx, vals = [0], [], []
    for i in range(n):
        nb = []
        for off in offsets:
            q = coords[i] + np.asarray(off)
            if (q >= 0).all() and (q < np.asarray(dims)).all():
                nb.append(int(np.ravel_multi_index(tuple(q), dims)))
        for j in sorted(nb + [i]):
            colidx.append(j)
            vals.append(c * ((len(nb) + 1) * np.eye(bs) + S @ S.T / bs) if j == i else -c * np.eye(bs))
        rowptr.append(len(colidx))
    rhs = rng.standard_normal((nrhs, n * bs))
    return pr.Problem(n, bs, bs, 1, np.array(rowptr, np.int32), np.array(colidx, np.int32), np.array(vals),
                      rhs[0].copy(), name=f"grid{dims}", meta={"rhs": rhs})


_OFFS = {
    "chain": [(-1,), (1,)],                                                     # 1 strict block per row
    "7pt": [(-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1)],   # 3
    "9pt": [(a, b) for a in (-1, 0, 1) for b in (-1, 0, 1) if (a, b) != (0, 0)],     # 4
    "27pt": [(a, b, d) for a in (-1, 0, 1) for b in (-1, 0, 1) for d in (-1, 0, 1) if (a, b, d) != (0, 0, 0)],  # 13
}
_DIMS = {
    "chain": [(57,), (300,), (131,)],
    "7pt": [(5, 6, 7), (8, 8, 8), (3, 9, 4)],
    "9pt": [(10, 12), (20, 15), (7, 3)],
    "27pt": [(5, 5, 5), (6, 7, 4), (3, 3, 9)],
}


@pytest.mark.parametrize("pattern,bs,width", [("chain", 1, 2), ("7pt", 2, 3), ("9pt", 2, 4), ("27pt", 4, 0),
                                              ("9pt", 4, 4), ("7pt", 4, 3), ("7pt", 8, 0)])
def test_ell_sweeps_equal_csr_sweeps_bitwise(ctx, pattern, bs, width):
    """The ELL copy of the sweeps (rows with at most 4 strict blocks of at most 4 x 4) runs the
    CSR sweeps' operations in the same order: identical iterates for every tile; wider rows
    and 8 x 8 blocks keep CSR."""
    rng = np.random.Generator(np.random.PCG64(31))
    probs = [_grid_problem(d, _OFFS[pattern], bs, rng) for d in _DIMS[pattern]]
    rp, ci, vals, sys = bsp.block_diagonal(probs)
    n = int(sys[-1])
    A = bsp.BSR(ctx, n, n, bs, rp, ci, vals)
    Bh = np.stack([np.concatenate([p.meta["rhs"][k] for p in probs]) for k in range(4)])
    B = ctx.to_device(Bh)
    try:
        ctx.set_option("batch_ell", 0)
        csr = bsp.Batch(ctx, A, sys, 4)
    finally:
        ctx.set_option("batch_ell", 1)
    ell = bsp.Batch(ctx, A, sys, 4)
    assert csr.ell_width() == 0 and ell.ell_width() == width
    for wg, r in ((256, 1), (1024, 2), (128, 4)):
        out = []
        for bat in (csr, ell):
            bat.set_tiles(wg, r)
            X = ctx.zeros(Bh.shape)
            its, rel, _, rc = bat.pcg(B, X, 1e-10, maxIters=3000, tol_mode=bsp.TOL_TRUE_RESIDUAL)
            assert rc == 0 and (rel <= 1e-10).all()
            out.append((its, X.numpy()))
        np.testing.assert_array_equal(out[0][0], out[1][0])
        np.testing.assert_array_equal(out[0][1], out[1][1])
    Xh = out[1][1]
    for s, p in enumerate(probs):
        for k in range(4):
            b = p.meta["rhs"][k]
            x = Xh[k, sys[s] * bs:sys[s + 1] * bs]
            assert np.linalg.norm(b - p.bsmv(x)) <= 2e-10 * np.linalg.norm(b)


@pytest.mark.parametrize("pattern,bs", [("9pt", 2), ("7pt", 1)])
def test_ell_batch_matches_oracle_pcg(ctx, pattern, bs):
    """ELL sweeps with 4 and 3 strict blocks per row against the oracle's pCG with IChol0
    (reference stopping rule)."""
    rng = np.random.Generator(np.random.PCG64(37))
    probs = [_grid_problem(d, _OFFS[pattern], bs, rng) for d in _DIMS[pattern]]
    rp, ci, vals, sys = bsp.block_diagonal(probs)
    n = int(sys[-1])
    A = bsp.BSR(ctx, n, n, bs, rp, ci, vals)
    bat = bsp.Batch(ctx, A, sys, 4)
    assert bat.ell_width() > 0
    Bh = np.stack([np.concatenate([p.meta["rhs"][k] for p in probs]) for k in range(4)])
    X = ctx.zeros(Bh.shape)
    its, _, _, rc = bat.pcg(ctx.to_device(Bh), X, 1e-10, maxIters=3000, tol_mode=bsp.TOL_REFERENCE)
    assert rc == 0
    Xh = X.numpy()
    for s, p in enumerate(probs):
        for k in range(4):
            b = p.meta["rhs"][k]
            xo, st = po.pcg(p, b, 1e-10, po.PRE_ICHOL0, maxIters=3000)
            x = Xh[k, sys[s] * bs:sys[s + 1] * bs]
            assert abs(int(its[s, k]) - st.iters) <= 1, (s, k, its[s, k], st.iters)
            assert np.linalg.norm(x - xo) <= 1e-8 * np.linalg.norm(xo)


def test_autotune_solve_keeps_a_finalist_and_solves(ctx):
    """Complete-solve tuning keeps one of the probed tiles; the batch then solves with it."""
    probs, A, sys, Bh = _batch(ctx, nsys=10, seed=17)
    bat = bsp.Batch(ctx, A, sys, 4)
    B = ctx.to_device(Bh)
    X = ctx.zeros(Bh.shape)
    wg, r = bat.autotune_solve(B, X, 1e-10, maxIters=2000, tol_mode=bsp.TOL_TRUE_RESIDUAL, probe_iters=4)
    assert wg in (128, 256, 512, 1024) and r in (1, 2, 4)
    assert bat.info()[:2] == (wg, r)
    X.zero()
    its, rel, _, rc = bat.pcg(B, X, 1e-10, maxIters=2000, tol_mode=bsp.TOL_TRUE_RESIDUAL)
    assert rc == 0 and (rel <= 1e-10).all()
    with pytest.raises(bsp.BspError):
        bat.autotune_solve(B, X, 0.0)


def test_graph_cache_survives_tile_changes(ctx):
    """A graph captured for one RHS-per-item setting stays valid after solves with others.

    pcg(R=4) captures a graph over the R=4 work-item arrays; solves at R=1 and R=2 build their
    own arrays; the R=4 graph is then replayed and must give the first solve's bits.
    """
    probs, A, sys, Bh = _batch(ctx, nsys=6, seed=11)
    bat = bsp.Batch(ctx, A, sys, 4)
    B = ctx.to_device(Bh)
    bat.set_tiles(256, 4)
    X0 = ctx.zeros(Bh.shape)
    its0, _, _, rc = bat.pcg(B, X0, 1e-10, maxIters=2000)
    assert rc == 0
    ref = X0.numpy().copy()
    for wg, r in [(256, 1), (512, 2), (256, 4), (1024, 1), (256, 4)]:
        bat.set_tiles(wg, r)
        X = ctx.zeros(Bh.shape)
        its, _, _, rc = bat.pcg(B, X, 1e-10, maxIters=2000)
        assert rc == 0
        if (wg, r) == (256, 4):
            assert np.array_equal(its, its0)
            assert np.array_equal(X.numpy(), ref)
        # the same operands again replay the cached graph of this setting
        its2, _, _, rc = bat.pcg(B, X, 1e-10, maxIters=2000)
        assert rc == 0 and np.array_equal(its2, its)
    # autotune probes rebuild items for every R; the cached R=4 graph must still be valid
    Xt = ctx.zeros(Bh.shape)
    bat.autotune(B, Xt, probe_iters=4)
    bat.set_tiles(256, 4)
    X = ctx.zeros(Bh.shape)
    its, _, _, rc = bat.pcg(B, X, 1e-10, maxIters=2000)
    assert rc == 0 and np.array_equal(X.numpy(), ref)
