"""CPU checks of the learned-correction network's oracle (oracle/mpnn_ref.py).

The analytic gradient is pinned two independent ways: central finite
differences of the loss, and torch autograd on a separate torch statement of
the same model.  A planted mistake in the backward pass (a dropped term, a
transposed block, a missing ReLU mask) changes the gradient by far more than
the tolerances used here.
"""
import numpy as np
import pytest

import mpnn_ref as mr
import pyoracle as po
from blocksp_amd import problems as pr


def _graph(nx=3, seed=5, coloured=False):
    rng = np.random.Generator(np.random.PCG64(seed))
    kappa = np.exp(rng.standard_normal((nx, nx, nx)))
    rp, ci, vals = pr.dg_laplacian((nx, nx, nx), kappa, D=1)
    p = pr.Problem(nx ** 3, 8, 8, 1, rp, ci, vals, np.zeros(nx ** 3 * 8))
    lu = po.ilu0_factor(p)[0]
    pos = None
    if coloured:
        col = pr.parity_colours((nx, nx, nx))
        order = np.argsort(col, kind="stable")
        pos = np.empty_like(order)
        pos[order] = np.arange(len(order))
    return mr.Graph(p.n, 8, rp, ci, vals, lu=lu, elim_pos=pos), p


def _theta(bs, seed=1):
    rng = np.random.default_rng(seed)
    th = mr.init_params(bs, seed)
    P = mr.unpack(th, bs)
    P["Wd"][:] = rng.uniform(-0.1, 0.1, P["Wd"].shape)
    P["bd"][:] = rng.uniform(-0.1, 0.1, P["bd"].shape)
    P["wg"][:] = rng.uniform(-0.1, 0.1, P["wg"].shape)
    P["bg"][:] = 0.05
    return mr.pack(P, bs)


def test_param_count_bs8():
    assert mr.num_params(8) == 62209


def test_diagonal_correction_stays_invertible():
    # worst case: every entry of tanh(D) at +-1 -> ||ALPHA T||_2 <= 8 ALPHA < 1
    g, p = _graph()
    rng = np.random.default_rng(0)
    dD = 1e3 * rng.choice([-1.0, 1.0], size=(g.n, 64))
    F = mr.corrected_factor(g, dD, np.zeros(len(g.e_slot)))
    for i in range(g.n):
        U = g.lu[g.diag_slot[i]]
        Fi = F[g.diag_slot[i]]
        assert np.linalg.norm(np.linalg.inv(Fi), 2) <= np.linalg.norm(np.linalg.inv(U), 2) / (1 - 8 * mr.ALPHA) + 1e-9


def test_zero_heads_reproduce_ilu0_factor():
    g, p = _graph()
    th = mr.init_params(8, 0)
    dD, gl, _ = mr.forward(th, g)
    assert np.all(dD == 0) and np.all(gl == 0)
    F = mr.corrected_factor(g, dD, gl)
    np.testing.assert_array_equal(F, g.lu)


def test_factor_product_matches_dense_lu():
    g, p = _graph(coloured=True)
    rng = np.random.default_rng(0)
    F = g.lu + 0.01 * rng.standard_normal(g.lu.shape)
    n, bs = g.n, g.bs
    Ld = np.eye(n * bs)
    Ud = np.zeros((n * bs, n * bs))
    for s in range(len(g.colidx)):
        i, j = g.rows[s], g.colidx[s]
        tgt = Ld if g.lower[s] else Ud
        tgt[i * bs:(i + 1) * bs, j * bs:(j + 1) * bs] = F[s]
    Z = rng.standard_normal((3, n * bs))
    y, u = mr.factor_product(g, F, Z)
    np.testing.assert_allclose(u, (Ud @ Z.T).T, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(y, (Ld @ Ud @ Z.T).T, rtol=1e-12, atol=1e-11)


def test_ilu0_factor_exact_on_pattern():
    # ILU0 reproduces A on its own pattern: (L U)_ij = A_ij for every stored block
    g, p = _graph()
    F = g.lu
    n, bs = g.n, g.bs
    Ld = np.eye(n * bs)
    Ud = np.zeros((n * bs, n * bs))
    for s in range(len(g.colidx)):
        i, j = g.rows[s], g.colidx[s]
        tgt = Ld if g.lower[s] else Ud
        tgt[i * bs:(i + 1) * bs, j * bs:(j + 1) * bs] = F[s]
    LU = Ld @ Ud
    A = pr.to_dense(p)
    for s in range(len(g.colidx)):
        i, j = g.rows[s], g.colidx[s]
        np.testing.assert_allclose(LU[i * bs:(i + 1) * bs, j * bs:(j + 1) * bs],
                                   A[i * bs:(i + 1) * bs, j * bs:(j + 1) * bs], atol=1e-10)


@pytest.mark.parametrize("coloured", [False, True])
def test_gradient_finite_differences(coloured):
    g, p = _graph(coloured=coloured)
    th = _theta(8)
    Z = np.random.default_rng(3).standard_normal((4, g.n * 8))
    L, G = mr.loss_and_grad(th, g, Z)
    assert abs(L - mr.loss(th, g, Z)) < 1e-12 * max(1.0, L)
    rng = np.random.default_rng(7)
    # random directions and individual coordinates from every parameter group
    dirs = [rng.standard_normal(th.size) for _ in range(3)]
    off = 0
    for name, shp in mr.param_shapes(8):
        cnt = int(np.prod(shp))
        e = np.zeros(th.size)
        e[off + rng.integers(cnt)] = 1.0
        dirs.append(e)
        off += cnt
    for d in dirs:
        d = d / np.linalg.norm(d)
        h = 1e-5
        fd = (mr.loss(th + h * d, g, Z) - mr.loss(th - h * d, g, Z)) / (2 * h)
        an = float(G @ d)
        assert abs(fd - an) <= 1e-6 * max(1.0, abs(an)) + 1e-8, (fd, an)


def test_gradient_torch_autograd():
    torch = pytest.importorskip("torch")
    g, p = _graph(coloured=True)
    th = _theta(8, seed=4)
    Z = np.random.default_rng(9).standard_normal((3, g.n * 8))
    L, G = mr.loss_and_grad(th, g, Z)
    T = torch.tensor(th, requires_grad=True)
    P, off = {}, 0
    for name, shp in mr.param_shapes(8):
        cnt = int(np.prod(shp))
        P[name] = T[off:off + cnt].reshape(shp)
        off += cnt
    t = lambda a: torch.tensor(a)
    x, e = t(g.x), t(g.e)
    src, dst = t(g.src), t(g.dst)
    h = torch.relu(x @ P["Wn"].T + P["bn"])
    phi = torch.relu(e @ P["We"].T + P["be"])
    for l in range(mr.NLAYERS):
        m = torch.relu(torch.cat([phi, h[dst]], 1) @ P[f"Wm{l}"].T + P[f"bm{l}"])
        a = torch.zeros(g.n, mr.H, dtype=torch.float64).index_add(0, src, m)
        h = h + torch.relu(torch.cat([h, a], 1) @ P[f"Wu{l}"].T + P[f"bu{l}"])
    dD = h @ P["Wd"].T + P["bd"]
    gl = (torch.cat([h[src], h[dst]], 1) @ P["wg"].T)[:, 0] + P["bg"][0]
    F = t(g.lu).clone()
    di, ei = t(g.diag_slot), t(g.e_slot)
    F = F.index_add(0, di, mr.ALPHA * torch.bmm(t(g.lu)[di], torch.tanh(dD.reshape(-1, 8, 8))))
    scale = torch.ones(len(g.colidx), dtype=torch.float64).index_copy(0, ei, 1 + mr.BETA * torch.tanh(gl))
    F = F * scale[:, None, None]
    # dense L', U' from F, then the loss
    n8 = g.n * 8
    Zt = t(Z)
    R = t(g.bsmv(Z))
    rows, cols = t(g.rows), t(g.colidx)
    low = t(g.lower)
    zb = Zt.reshape(-1, g.n, 8)
    K = Z.shape[0]
    u = torch.zeros(K, g.n, 8, dtype=torch.float64).index_add(
        1, rows[~low], torch.einsum("pij,kpj->kpi", F[~low], zb[:, cols[~low], :]))
    y = u.index_add(1, rows[low], torch.einsum("pij,kpj->kpi", F[low], u[:, cols[low], :]))
    res = y.reshape(K, -1) - R
    loss = ((res ** 2).sum(1) / (R ** 2).sum(1)).mean()
    loss.backward()
    assert abs(loss.item() - L) <= 1e-12 * L
    np.testing.assert_allclose(G, T.grad.numpy(), rtol=1e-9, atol=1e-12)


def test_adam_matches_closed_form_first_step():
    # first bias-corrected step moves every coordinate by lr * sign(g) (up to eps)
    rng = np.random.default_rng(0)
    th = rng.standard_normal(50)
    gr = rng.standard_normal(50)
    new, m, v = mr.adam_step(th, gr, np.zeros(50), np.zeros(50), 1, 1e-2)
    np.testing.assert_allclose(new, th - 1e-2 * np.sign(gr), rtol=0, atol=1e-6)
