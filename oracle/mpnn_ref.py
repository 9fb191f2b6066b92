"""CPU oracle for the learned ILU0 correction network (test infrastructure only).

Only tests, the package's smoke check and the benchmark's CPU-baseline leg use
this module, and only as the checker of the HIP kernels.  The learned
preconditioner has no counterpart in the reference BlockSp code; this module
states the model in plain numpy (float64) so the device forward pass, loss and
gradients can be compared against it.  It is pinned by central finite
differences and by torch autograd (tests/test_mpnn_oracle.py), which share no
code with the analytic backward pass below.

Model (hidden width H = 64, three message-passing layers, block size B <= 8):
  s_i  = ||A_ii||_F
  x_i  = vec(A_ii) / s_i                                   (B*B node features)
  e_ij = [vec(A_ij) / sqrt(s_i s_j), log s_j - log s_i]    (B*B + 1 edge features)
  h0   = relu(Wn x + bn),  phi = relu(We e + be)
  layer l: m_ij = relu(Wm_l [phi_ij ; h_j] + bm_l),  a_i = sum_j m_ij,
           h   <- h + relu(Wu_l [h_i ; a_i] + bu_l)
  heads:   dD_i = Wd h_i + bd (B*B values),  g_ij = wg . [h_i ; h_j] + bg
Correction of an ILU0 factor LU (stored in A's block pattern):
  F_ii = LU_ii (I + ALPHA tanh(D_i)),   F_ij = LU_ij (1 + BETA tanh g_ij)
with D_i = dD_i as a B x B matrix (row-major).  Every entry of ALPHA tanh(D_i) is at
most ALPHA in magnitude, so ||ALPHA tanh(D_i)||_2 <= B ALPHA = 0.4 for B = 8: the
corrected diagonal block stays invertible with ||F_ii^-1|| <= ||LU_ii^-1|| / 0.6.
Loss over sample vectors z_k (r_k = A z_k):
  U' = blocks of F at or right of the diagonal in elimination order, L' = I + the rest,
  loss = mean_k ||L' U' z_k - r_k||^2 / ||r_k||^2
Edges are the off-diagonal blocks in CSR order (row by row, stored order).
"""
from __future__ import annotations

import numpy as np

H = 64
NLAYERS = 3
ALPHA = 0.05
BETA = 0.25


def param_shapes(bs: int):
    b2 = bs * bs
    shapes = [("Wn", (H, b2)), ("bn", (H,)), ("We", (H, b2 + 1)), ("be", (H,))]
    for l in range(NLAYERS):
        shapes += [(f"Wm{l}", (H, 2 * H)), (f"bm{l}", (H,)), (f"Wu{l}", (H, 2 * H)), (f"bu{l}", (H,))]
    shapes += [("Wd", (b2, H)), ("bd", (b2,)), ("wg", (1, 2 * H)), ("bg", (1,))]
    return shapes


def num_params(bs: int) -> int:
    return int(sum(np.prod(s) for _, s in param_shapes(bs)))


def unpack(theta: np.ndarray, bs: int) -> dict:
    out, off = {}, 0
    for name, shp in param_shapes(bs):
        cnt = int(np.prod(shp))
        out[name] = theta[off:off + cnt].reshape(shp)
        off += cnt
    assert off == theta.size
    return out


def pack(d: dict, bs: int) -> np.ndarray:
    return np.concatenate([np.asarray(d[name], dtype=np.float64).ravel() for name, _ in param_shapes(bs)])


def init_params(bs: int, seed: int = 0) -> np.ndarray:
    """Uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) for the trunk, zero heads."""
    rng = np.random.default_rng(seed)
    d = {}
    for name, shp in param_shapes(bs):
        if name in ("Wd", "bd", "wg", "bg"):
            d[name] = np.zeros(shp)
            continue
        fan_in = shp[1] if len(shp) == 2 else dict(param_shapes(bs))["W" + name[1:]][1]
        lim = 1.0 / np.sqrt(fan_in)
        d[name] = rng.uniform(-lim, lim, size=shp)
    return pack(d, bs)


class Graph:
    """Block graph of a square BSR matrix plus the factor split for the loss."""

    def __init__(self, n, bs, rowptr, colidx, vals, lu=None, elim_pos=None):
        self.n, self.bs = int(n), int(bs)
        self.rowptr = np.asarray(rowptr, dtype=np.int64)
        self.colidx = np.asarray(colidx, dtype=np.int64)
        self.vals = np.asarray(vals, dtype=np.float64).reshape(-1, bs, bs)
        rows = np.repeat(np.arange(n), np.diff(self.rowptr))
        self.rows = rows
        diag = rows == self.colidx
        if diag.sum() != n:
            raise ValueError("every block row needs exactly one diagonal block")
        self.diag_slot = np.zeros(n, dtype=np.int64)
        self.diag_slot[rows[diag]] = np.nonzero(diag)[0]
        off = ~diag
        self.e_slot = np.nonzero(off)[0]
        self.src = rows[off]
        self.dst = self.colidx[off]
        D = self.vals[self.diag_slot]
        self.s = np.sqrt((D.reshape(n, -1) ** 2).sum(1))
        ls = np.log(self.s)
        self.x = D.reshape(n, -1) / self.s[:, None]
        Ae = self.vals[self.e_slot].reshape(len(self.e_slot), -1)
        self.e = np.concatenate([Ae / np.sqrt(self.s[self.src] * self.s[self.dst])[:, None],
                                 (ls[self.dst] - ls[self.src])[:, None]], axis=1)
        pos = np.arange(n) if elim_pos is None else np.asarray(elim_pos)
        self.lower = pos[self.colidx] < pos[rows]
        self.lu = None if lu is None else np.asarray(lu, dtype=np.float64).reshape(-1, bs, bs)

    def bsmv(self, z):
        bs = self.bs
        zb = z.reshape(-1, self.n, bs)
        out = np.zeros_like(zb)
        contrib = np.einsum("pij,kpj->kpi", self.vals, zb[:, self.colidx, :])
        np.add.at(out, (slice(None), self.rows), contrib)
        return out.reshape(z.shape)


def forward(theta, g: Graph):
    P = unpack(theta, g.bs)
    relu = lambda t: np.maximum(t, 0.0)
    act = {}
    h = relu(g.x @ P["Wn"].T + P["bn"])
    act["h"] = [h]
    phi = relu(g.e @ P["We"].T + P["be"])
    act["phi"] = phi
    act["m"], act["a"], act["t"] = [], [], []
    for l in range(NLAYERS):
        m = relu(np.concatenate([phi, h[g.dst]], 1) @ P[f"Wm{l}"].T + P[f"bm{l}"])
        a = np.zeros((g.n, H))
        np.add.at(a, g.src, m)
        t = relu(np.concatenate([h, a], 1) @ P[f"Wu{l}"].T + P[f"bu{l}"])
        h = h + t
        act["m"].append(m)
        act["a"].append(a)
        act["t"].append(t)
        act["h"].append(h)
    dD = h @ P["Wd"].T + P["bd"]
    gl = (np.concatenate([h[g.src], h[g.dst]], 1) @ P["wg"].T)[:, 0] + P["bg"][0]
    return dD, gl, act


def corrected_factor(g: Graph, dD, gl):
    F = g.lu.copy()
    Ld = g.lu[g.diag_slot]
    F[g.diag_slot] += ALPHA * np.einsum("nij,njk->nik", Ld, np.tanh(dD.reshape(g.n, g.bs, g.bs)))
    F[g.e_slot] *= (1.0 + BETA * np.tanh(gl))[:, None, None]
    return F


def factor_product(g: Graph, F, Z):
    """y = L' U' z for each row of Z (K x n*bs); returns (y, u)."""
    bs, n = g.bs, g.n
    K = Z.shape[0]
    zb = Z.reshape(K, n, bs)
    up = ~g.lower
    u = np.zeros((K, n, bs))
    np.add.at(u, (slice(None), g.rows[up]), np.einsum("pij,kpj->kpi", F[up], zb[:, g.colidx[up], :]))
    y = u.copy()
    lo = g.lower
    np.add.at(y, (slice(None), g.rows[lo]), np.einsum("pij,kpj->kpi", F[lo], u[:, g.colidx[lo], :]))
    return y.reshape(K, -1), u.reshape(K, -1)


def loss(theta, g: Graph, Z):
    dD, gl, _ = forward(theta, g)
    F = corrected_factor(g, dD, gl)
    R = g.bsmv(Z)
    y, _ = factor_product(g, F, Z)
    return float(np.mean(((y - R) ** 2).sum(1) / (R ** 2).sum(1)))


def loss_and_grad(theta, g: Graph, Z):
    """Loss and its analytic gradient with respect to theta."""
    bs, n = g.bs, g.n
    P = unpack(theta, bs)
    dD, gl, act = forward(theta, g)
    F = corrected_factor(g, dD, gl)
    R = g.bsmv(Z)
    y, u = factor_product(g, F, Z)
    K = Z.shape[0]
    rn = (R ** 2).sum(1)
    res = y - R
    L = float(np.mean((res ** 2).sum(1) / rn))
    gy = (2.0 / (K * rn))[:, None] * res
    gyb = gy.reshape(K, n, bs)
    ub = u.reshape(K, n, bs)
    zb = Z.reshape(K, n, bs)
    lo, up = g.lower, ~g.lower
    # dL/du = gy + L'^T gy
    gu = gyb.copy()
    np.add.at(gu, (slice(None), g.colidx[lo]), np.einsum("pij,kpi->kpj", F[lo], gyb[:, g.rows[lo], :]))
    gF = np.zeros_like(F)
    gF[lo] = np.einsum("kpi,kpj->pij", gyb[:, g.rows[lo], :], ub[:, g.colidx[lo], :])
    gF[up] = np.einsum("kpi,kpj->pij", gu[:, g.rows[up], :], zb[:, g.colidx[up], :])
    # to the network outputs
    T = np.tanh(dD.reshape(n, bs, bs))
    gT = ALPHA * np.einsum("nki,nkj->nij", g.lu[g.diag_slot], gF[g.diag_slot])
    g_dD = (gT * (1.0 - T ** 2)).reshape(n, -1)
    th = np.tanh(gl)
    g_g = BETA * (1.0 - th ** 2) * (g.lu[g.e_slot] * gF[g.e_slot]).sum((1, 2))
    G = {k: np.zeros_like(v) for k, v in P.items()}
    h3 = act["h"][NLAYERS]
    G["Wd"] = g_dD.T @ h3
    G["bd"] = g_dD.sum(0)
    cat = np.concatenate([h3[g.src], h3[g.dst]], 1)
    G["wg"] = (g_g @ cat)[None, :]
    G["bg"] = np.array([g_g.sum()])
    gh = g_dD @ P["Wd"]
    np.add.at(gh, g.src, np.outer(g_g, P["wg"][0, :H]))
    np.add.at(gh, g.dst, np.outer(g_g, P["wg"][0, H:]))
    gphi = np.zeros_like(act["phi"])
    for l in reversed(range(NLAYERS)):
        h, a, t, m = act["h"][l], act["a"][l], act["t"][l], act["m"][l]
        gpre_u = gh * (t > 0)
        G[f"Wu{l}"] = gpre_u.T @ np.concatenate([h, a], 1)
        G[f"bu{l}"] = gpre_u.sum(0)
        gin = gpre_u @ P[f"Wu{l}"]
        gh_new = gh + gin[:, :H]
        ga = gin[:, H:]
        gpre_m = ga[g.src] * (m > 0)
        G[f"Wm{l}"] = gpre_m.T @ np.concatenate([act["phi"], h[g.dst]], 1)
        G[f"bm{l}"] = gpre_m.sum(0)
        gin_m = gpre_m @ P[f"Wm{l}"]
        gphi += gin_m[:, :H]
        np.add.at(gh_new, g.dst, gin_m[:, H:])
        gh = gh_new
    gpre_e = gphi * (act["phi"] > 0)
    G["We"] = gpre_e.T @ g.e
    G["be"] = gpre_e.sum(0)
    gpre_n = gh * (act["h"][0] > 0)
    G["Wn"] = gpre_n.T @ g.x
    G["bn"] = gpre_n.sum(0)
    return L, pack(G, bs)


def adam_step(theta, grad, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One bias-corrected Adam step (Kingma & Ba 2015, Algorithm 1); t counts from 1."""
    m = b1 * m + (1 - b1) * grad
    v = b2 * v + (1 - b2) * grad * grad
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return theta - lr * mh / (np.sqrt(vh) + eps), m, v
