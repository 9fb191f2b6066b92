"""Complex systems (BlockSp<std::complex<double>, B>) on the device: the real-equivalent
matrix, its products, and the complex pGMRES against the complex oracle (oracle/zoracle.py,
pinned in tests/test_zoracle.py): iteration counts within max(2, 10%) and solutions within
1e-6 of the oracle's, for the preconditioners the reference allows for complex T (NoPre,
Diag and SOR with LU block inverses, ILU0), with and without (deflated) restarts."""
import numpy as np
import pytest

import blocksp_amd as bsp
import zoracle as zo
from blocksp_amd import problems as pr

pytestmark = pytest.mark.gpu


def rhs(n, seed=1):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(n) + 1j * rng.standard_normal(n)


def make(ctx, p, seed=0, scale=0.3):
    rp, ci, v = zo.complexify(p, seed, scale)
    return bsp.BSR.from_complex(ctx, p.n, p.n, p.bs, rp, ci, v), zo.ZBsr(p.n, p.bs, rp, ci, v)


def random_complex_bsr(n, bs, seed, dens=4):
    rng = np.random.default_rng(seed)
    rp, ci = [0], []
    for i in range(n):
        cols = list(rng.choice(n, size=min(n, int(rng.integers(1, dens + 1))), replace=False))
        if i not in cols:
            cols.append(i)
        rng.shuffle(cols)                      # insertion order, not sorted
        ci += cols
        rp.append(len(ci))
    v = rng.standard_normal((len(ci), bs, bs)) + 1j * rng.standard_normal((len(ci), bs, bs))
    for i in range(n):
        for q in range(rp[i], rp[i + 1]):
            if ci[q] == i:
                v[q] += (3 * bs) * np.eye(bs)
    return np.array(rp, np.int32), np.array(ci, np.int32), v


@pytest.mark.parametrize("bs", [1, 2, 3, 4, 8])
def test_products_and_storage(ctx, tmp_path, bs):
    print(f"complex products bs {bs}", flush=True)
    n = 60
    rp, ci, v = random_complex_bsr(n, bs, seed=bs)
    A = bsp.BSR.from_complex(ctx, n, n, bs, rp, ci, v)
    assert A.is_complex and A.cbs == bs and A.bs == 2 * bs
    Z = zo.ZBsr(n, bs, rp, ci, v)
    D = Z.dense()
    x = rhs(n * bs, 3)
    y = ctx.zeros(n * bs, np.complex128)
    bsp.bsmv(ctx, A, ctx.to_device(x), y)
    ref = D @ x
    bound = np.abs(D) @ np.abs(x)
    assert np.all(np.abs(y.numpy() - ref) <= 1e-13 * bound + 1e-300)
    # transposed product: the conjugate transpose, as the reference's for complex T
    yt = ctx.zeros(n * bs, np.complex128)
    bsp.bsmv_transpose(ctx, A, ctx.to_device(x), yt)
    np.testing.assert_allclose(yt.numpy(), D.conj().T @ x, rtol=0, atol=1e-12 * np.abs(D).sum(0).max() * np.abs(x).max())
    # download, cache file round trip
    rp2, ci2, v2 = A.download()
    np.testing.assert_array_equal(rp2, rp)
    np.testing.assert_array_equal(ci2, ci)
    np.testing.assert_array_equal(v2, v)
    A.save(tmp_path / "c.bsr")
    B = bsp.BSR.load(ctx, tmp_path / "c.bsr")
    assert B.cbs == bs
    np.testing.assert_array_equal(B.download()[2], v)
    # C = A A and C = A^H A, complex
    C = bsp.bsmm(ctx, A, A)
    assert C.cbs == bs
    crp, cci, cv = C.download()
    Cd = np.zeros_like(D)
    for i in range(n):
        for q in range(crp[i], crp[i + 1]):
            Cd[i * bs:(i + 1) * bs, cci[q] * bs:(cci[q] + 1) * bs] = cv[q]
    np.testing.assert_allclose(Cd, D @ D, rtol=0, atol=1e-12 * np.abs(D).max() ** 2 * n)
    Ch = bsp.bsmm(ctx, A, A, Atran=True)
    crp, cci, cv = Ch.download()
    Cd = np.zeros_like(D)
    for i in range(n):
        for q in range(crp[i], crp[i + 1]):
            Cd[i * bs:(i + 1) * bs, cci[q] * bs:(cci[q] + 1) * bs] = cv[q]
    np.testing.assert_allclose(Cd, D.conj().T @ D, rtol=0, atol=1e-12 * np.abs(D).max() ** 2 * n)


CASES = [("ilu0", {}, 30, 10), ("ilu0", {}, -1, -1), ("ilu0", {}, 12, 0), ("sor", dict(w=1.0, nu=2), 20, 4),
         ("diag", {}, 25, 6), ("none", {}, 40, 8)]


def precond(ctx, kind, A, kw):
    if kind == "ilu0":
        return bsp.ILU0(ctx, A), zo.PRE_ILU0
    if kind == "sor":
        return bsp.SOR(ctx, A, kw["w"], kw["nu"], 0), zo.PRE_SOR
    if kind == "diag":
        return bsp.Diag(ctx, A), zo.PRE_DIAG
    return bsp.NoPre(ctx, A), zo.PRE_NONE


@pytest.mark.parametrize("kind,kw,restart,k", CASES)
def test_pgmres_matches_complex_oracle(ctx, kind, kw, restart, k):
    print(f"complex pgmres {kind} restart {restart} k {k}", flush=True)
    p = pr.dg2d_veclap(10, 9, seed=2)
    A, Z = make(ctx, p, seed=4)
    b = rhs(p.n * p.bs, 5)
    M, zkind = precond(ctx, kind, A, kw)
    xd, st, rc = bsp.pGMRES(ctx, A, ctx.to_device(b), 1e-10, M, maxIters=3000, restart=restart, k=k)
    xo, sto = zo.pgmres(Z, b, 1e-10, zo.Precond(Z, zkind, **kw), maxIters=3000, restart=restart, k=k)
    assert rc == 0 and st.converged
    assert abs(st.iters - sto.iters) <= max(2, 0.1 * sto.iters), (st.iters, sto.iters)
    if restart > 0 and sto.iters > restart:
        assert st.restarts >= 1
    x = xd.numpy()
    assert x.dtype == np.complex128
    assert np.linalg.norm(x - xo) <= 1e-6 * np.linalg.norm(xo)
    assert st.true_relres == pytest.approx(np.linalg.norm(b - Z.dense() @ x) / np.linalg.norm(b), rel=1e-6)
    # (harmonic) Ritz values of the last cycle's complex Hessenberg
    ev, harm = bsp.pgmres_ritz(ctx)
    assert len(ev) > 0 and np.all(np.isfinite(ev)) and bool(harm) == (st.restarts > 0 and k > 0)


def test_unit_modulus_scaling(ctx):
    """b -> c b (|c| = 1) gives x -> c x in the same iteration count -- false when a
    Gram-Schmidt coefficient misses its conjugate."""
    p = pr.dg2d_veclap(10, 9, seed=2)
    A, Z = make(ctx, p, seed=6)
    M = bsp.ILU0(ctx, A)
    b = rhs(p.n * p.bs, 7)
    c = np.exp(1.1j)
    x1, s1, _ = bsp.pGMRES(ctx, A, ctx.to_device(b), 1e-10, M, maxIters=3000, restart=15, k=5)
    x2, s2, _ = bsp.pGMRES(ctx, A, ctx.to_device(c * b), 1e-10, M, maxIters=3000, restart=15, k=5)
    assert abs(s1.iters - s2.iters) <= 1
    np.testing.assert_allclose(x2.numpy(), c * x1.numpy(), rtol=0, atol=1e-8 * np.abs(x1.numpy()).max())


def test_true_residual_mode(ctx):
    p = pr.dg2d_veclap(10, 9, seed=2)
    A, Z = make(ctx, p, seed=8)
    M = bsp.ILU0(ctx, A)
    b = rhs(p.n * p.bs, 9)
    xd, st, rc = bsp.pGMRES(ctx, A, ctx.to_device(b), 1e-11, M, maxIters=3000, restart=20, k=5,
                            tol_mode=bsp.TOL_TRUE_RESIDUAL)
    x = xd.numpy()
    assert rc == 0 and np.linalg.norm(b - Z.dense() @ x) <= 1e-11 * np.linalg.norm(b)


def test_unsupported_routes(ctx):
    p = pr.dg2d_veclap(4, 4, seed=1)
    A, _ = make(ctx, p)
    b = ctx.to_device(rhs(p.n * p.bs))
    M = bsp.ILU0(ctx, A)
    with pytest.raises(bsp.BspError):
        bsp.pCG(ctx, A, b, 1e-8, M)
    with pytest.raises(bsp.BspError):
        bsp.Diag(ctx, A, cholesky=True)
    with pytest.raises(bsp.BspError):
        bsp.IChol0(ctx, A)
    with pytest.raises(bsp.BspError):
        bsp.ILU0Ex(ctx, A, precision=bsp.PREC_F32)
    with pytest.raises(bsp.BspError):
        bsp.FGMRES_mixed(ctx, A, b, 1e-8, M)
    R = bsp.BSR.from_problem(ctx, pr.dg2d_veclap(4, 4, seed=1))
    with pytest.raises(bsp.BspError):
        bsp.bsmm(ctx, A, R)
