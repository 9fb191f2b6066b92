"""Learned ILU0 correction on the GPU: kernels against the CPU oracle, training, ablation
and the FGMRES fallback."""
import numpy as np
import pytest

import blocksp_amd as bsp
from blocksp_amd import learned
from blocksp_amd import problems as pr
import mpnn_ref as mr

pytestmark = pytest.mark.gpu


def _grf(nx, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    kappa = np.exp(pr._gaussian_field((nx, nx, nx), nx / 4.0, rng))
    rp, ci, vals = pr.dg_laplacian((nx, nx, nx), kappa, D=1)
    b = rng.standard_normal(nx ** 3 * 8)
    return pr.Problem(nx ** 3, 8, 8, 1, rp, ci, vals, b, name=f"grf{nx}_{seed}")


def _colour(p):
    nx = round(p.n ** (1 / 3))
    return pr.parity_colours((nx, nx, nx))


def _theta(seed=1):
    rng = np.random.default_rng(seed)
    P = mr.unpack(mr.init_params(8, seed), 8)
    P["Wd"][:] = rng.uniform(-0.1, 0.1, P["Wd"].shape)
    P["bd"][:] = rng.uniform(-0.1, 0.1, P["bd"].shape)
    P["wg"][:] = rng.uniform(-0.1, 0.1, P["wg"].shape)
    P["bg"][:] = 0.05
    return mr.pack(P, 8).astype(np.float32)


def _setup(ctx, nx=4, seed=5, coloured=False):
    p = _grf(nx, seed)
    A = bsp.BSR.from_problem(ctx, p)
    col = _colour(p) if coloured else None
    M = bsp.ILU0Ex(ctx, A, colour=col)
    g = bsp.MPNNGraph(ctx, A, M)
    rp, ci, lu = M.factor()
    assert np.array_equal(rp, p.rowptr) and np.array_equal(ci, p.colidx)
    pos = None
    if col is not None:
        order = np.argsort(col, kind="stable")
        pos = np.empty_like(order)
        pos[order] = np.arange(len(order))
    og = mr.Graph(p.n, 8, p.rowptr, p.colidx, p.vals, lu=lu, elim_pos=pos)
    return p, A, M, g, og


@pytest.mark.parametrize("coloured", [False, True])
def test_forward_matches_oracle(ctx, coloured):
    p, A, M, g, og = _setup(ctx, coloured=coloured)
    th = _theta()
    net = bsp.MPNN(ctx, 8, seed=0)
    assert net.nparams == mr.num_params(8)
    net.set_params(th)
    dD, gl = net.infer(g)
    rdD, rgl, _ = mr.forward(th.astype(np.float64), og)
    np.testing.assert_allclose(dD, rdD, rtol=1e-4, atol=1e-5 * np.abs(rdD).max())
    np.testing.assert_allclose(gl, rgl, rtol=1e-4, atol=1e-5 * np.abs(rgl).max())


@pytest.mark.parametrize("coloured", [False, True])
def test_loss_and_gradient_match_oracle(ctx, coloured):
    p, A, M, g, og = _setup(ctx, coloured=coloured)
    th = _theta(seed=3)
    net = bsp.MPNN(ctx, 8, seed=0)
    net.set_params(th)
    Z = np.random.default_rng(11).standard_normal((4, p.n * 8))
    Zd = ctx.to_device(Z)
    L = net.loss_grad(g, Zd, 4)
    G = net.grads().astype(np.float64)
    Lr, Gr = mr.loss_and_grad(th.astype(np.float64), og, Z)
    assert abs(L - Lr) <= 1e-5 * Lr, (L, Lr)
    # whole gradient and each parameter group, relative to the group's norm (FP32 kernels)
    assert np.linalg.norm(G - Gr) <= 1e-4 * np.linalg.norm(Gr)
    off = 0
    for name, shp in mr.param_shapes(8):
        cnt = int(np.prod(shp))
        a, b = G[off:off + cnt], Gr[off:off + cnt]
        assert np.linalg.norm(a - b) <= 1e-3 * np.linalg.norm(b) + 1e-7, name
        off += cnt


def test_untrained_network_reproduces_ilu0(ctx):
    p, A, M, g, og = _setup(ctx)
    net = bsp.MPNN(ctx, 8, seed=0)          # zero heads
    dD, gl = net.infer(g)
    assert not dD.any() and not gl.any()
    Ml = bsp.LearnedILU0(ctx, A, net)
    np.testing.assert_array_equal(Ml.factor()[2], M.factor()[2])


def test_adam_step_matches_oracle(ctx):
    p, A, M, g, og = _setup(ctx)
    th = _theta(seed=2)
    net = bsp.MPNN(ctx, 8, seed=0)
    net.set_params(th)
    Zd = ctx.to_device(np.random.default_rng(1).standard_normal((2, p.n * 8)))
    net.loss_grad(g, Zd, 2)
    G = net.grads().astype(np.float64)
    net.adam_step(1e-3)
    ref, _, _ = mr.adam_step(th.astype(np.float64), G, np.zeros_like(G), np.zeros_like(G), 1, 1e-3)
    np.testing.assert_allclose(net.params(), ref, rtol=0, atol=2e-6)


def test_record_basis_gives_orthonormal_krylov_vectors(ctx):
    p = _grf(6, 7)
    A = bsp.BSR.from_problem(ctx, p)
    M = bsp.ILU0(ctx, A)
    Z, K = learned.record_krylov(ctx, A, M, p.b, 20)
    assert K == 20
    Q = Z.numpy()[:K]
    # modified Gram-Schmidt: orthogonality degrades like eps * cond(Krylov basis)
    np.testing.assert_allclose(Q @ Q.T, np.eye(K), atol=1e-6)
    np.testing.assert_allclose(np.diag(Q @ Q.T), np.ones(K), atol=1e-12)
    r = ctx.to_device(p.b)
    M(r)
    q0 = r.numpy() / np.linalg.norm(r.numpy())
    np.testing.assert_allclose(Q[0], q0, rtol=0, atol=1e-12)


def test_learned_correction_cuts_iterations(ctx):
    """Ablation: the trained network against the same FGMRES with the uncorrected ILU0."""
    train_set = [_grf(8, s) for s in range(8)]
    data = learned.training_set(ctx, train_set, count=64)
    net, losses = learned.train(ctx, data, epochs=40, batch=16, lr=1e-3)
    assert losses[-1] < 0.8 * losses[0], losses
    plain_its, learned_its = [], []
    for s in (100, 101, 102):
        p = _grf(12, s)
        A = bsp.BSR.from_problem(ctx, p)
        b = ctx.to_device(p.b)
        M0 = bsp.ILU0(ctx, A)
        Ml = bsp.LearnedILU0(ctx, A, net)
        _, st0, rc0 = bsp.FGMRES_mixed(ctx, A, b, 1e-10, M0, spmv_precision=bsp.PREC_F64)
        _, st1, rc1 = bsp.FGMRES_mixed(ctx, A, b, 1e-10, Ml, spmv_precision=bsp.PREC_F64, fallback=M0)
        assert rc0 == 0 and rc1 == 0
        assert st0.true_relres <= 1e-10 and st1.true_relres <= 1e-10
        assert st1.fallback == 0
        plain_its.append(st0.iters)
        learned_its.append(st1.iters)
    print("ablation iterations plain", plain_its, "learned", learned_its, "loss", losses[0], "->", losses[-1])
    assert sum(learned_its) <= 0.9 * sum(plain_its), (plain_its, learned_its)


def test_learned_correction_on_the_coloured_factor(ctx):
    """Ablation on the parity-coloured factor (the mixed solver's elimination order): an untrained
    network reproduces the coloured ILU0 bit for bit, and a network trained on Krylov vectors of
    coloured ILU0-pGMRES solves converges without the fallback in at most the plain coloured
    FGMRES's iterations + 5% -- its gain on this factor is a few percent: the fill that colouring
    drops couples rows of one colour, which a block-diagonal correction cannot represent."""
    p = _grf(12, 100)
    A = bsp.BSR.from_problem(ctx, p)
    col = _colour(p)
    Mc = bsp.ILU0Ex(ctx, A, colour=col)
    untrained = bsp.LearnedILU0(ctx, A, bsp.MPNN(ctx, 8, seed=0), colour=col)
    np.testing.assert_array_equal(untrained.factor()[2], Mc.factor()[2])
    data = learned.training_set(ctx, [_grf(8, s) for s in range(8)], count=64, colour_fn=_colour, krylov_colour=True)
    net, losses = learned.train(ctx, data, epochs=20, batch=16, lr=1e-3)
    assert losses[-1] <= losses[0], losses
    plain_its, learned_its = [], []
    for s in (100, 101, 102):
        p = _grf(12, s)
        A = bsp.BSR.from_problem(ctx, p)
        b = ctx.to_device(p.b)
        Mc = bsp.ILU0Ex(ctx, A, colour=col)
        Ml = bsp.LearnedILU0(ctx, A, net, colour=col)
        _, st0, rc0 = bsp.FGMRES_mixed(ctx, A, b, 1e-10, Mc, spmv_precision=bsp.PREC_F64)
        _, st1, rc1 = bsp.FGMRES_mixed(ctx, A, b, 1e-10, Ml, spmv_precision=bsp.PREC_F64, fallback=Mc)
        assert rc0 == 0 and rc1 == 0 and st0.true_relres <= 1e-10 and st1.true_relres <= 1e-10
        assert st1.fallback == 0
        plain_its.append(st0.iters)
        learned_its.append(st1.iters)
    print("coloured ablation iterations plain", plain_its, "learned", learned_its, "loss", losses[0], "->", losses[-1])
    assert sum(learned_its) <= 1.05 * sum(plain_its), (plain_its, learned_its)


def test_nonfinite_preconditioner_falls_back(ctx):
    p = _grf(6, 9)
    A = bsp.BSR.from_problem(ctx, p)
    b = ctx.to_device(p.b)
    M0 = bsp.ILU0(ctx, A)
    Mbad = bsp.HostPrecond(ctx, A, lambda r: r.fill(np.nan))
    x, st, rc = bsp.FGMRES_mixed(ctx, A, b, 1e-10, Mbad, spmv_precision=bsp.PREC_F64, fallback=M0)
    assert rc == 0 and st.fallback == 1 and st.true_relres <= 1e-10
    assert np.isfinite(x.numpy()).all()


def test_non_finite_network_is_refused(ctx):
    p = _grf(4, 10)
    A = bsp.BSR.from_problem(ctx, p)
    net = bsp.MPNN(ctx, 8, seed=0)
    th = net.params()
    th[-1] = np.nan                          # the edge head's bias
    net.set_params(th)
    with pytest.raises(bsp.BspError):
        bsp.LearnedILU0(ctx, A, net)


def test_saturated_network_still_converges(ctx):
    """Heads driven to their bounds everywhere (the worst a trained network can output):
    the corrected factor stays invertible and FGMRES, with the ILU0 fallback armed, reaches
    the FP64 target."""
    p = _grf(6, 10)
    A = bsp.BSR.from_problem(ctx, p)
    b = ctx.to_device(p.b)
    net = bsp.MPNN(ctx, 8, seed=0)
    rng = np.random.default_rng(0)
    P = mr.unpack(net.params().astype(np.float64), 8)
    P["bd"][:] = rng.choice([-300.0, 300.0], P["bd"].shape)
    P["wg"][:] = rng.choice([-100.0, 100.0], P["wg"].shape)
    net.set_params(mr.pack(P, 8).astype(np.float32))
    M0 = bsp.ILU0(ctx, A)
    Ml = bsp.LearnedILU0(ctx, A, net)
    x, st, rc = bsp.FGMRES_mixed(ctx, A, b, 1e-10, Ml, spmv_precision=bsp.PREC_F64, fallback=M0)
    assert rc == 0 and st.true_relres <= 1e-10, st.as_dict()
