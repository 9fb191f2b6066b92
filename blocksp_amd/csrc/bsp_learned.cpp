// bsp_learned.cpp -- learned ILU0 correction: network, block graph, training and
// inference on the device, and the corrected ILU0 preconditioner.
//
// Graph: nodes are block rows; edges are A's off-diagonal blocks in CSR order (edge e
// of row i is its e - e_ptr[i]-th off-diagonal block).  The pattern must be
// structurally symmetric: e_rev[e] is the edge of the transposed block and tslot[p]
// the block slot of the transpose of slot p, so every transposed accumulation of the
// backward pass is a gather over a row (no atomics, deterministic).
#include <algorithm>
#include <cmath>
#include <cstring>

#include "bsp_internal.hpp"
#include "bsp_mpnn.hpp"

using namespace bsp;
using namespace bsp::mpnn;

struct bsp_mpnn {
    bsp_ctx* ctx = nullptr;
    int bs = 0, b2 = 0;
    int64_t nparams = 0;
    DevMem theta, grad, m, v;
    int64_t step = 0;
    // parameter offsets (floats)
    int64_t oWn, obn, oWe, obe, oWm[kLayers], obm[kLayers], oWu[kLayers], obu[kLayers], oWd, obd, owg, obg;
    const float* P(int64_t off) const { return theta.as<float>() + off; }
    float* G(int64_t off) const { return grad.as<float>() + off; }
};

struct bsp_mpnn_graph {
    bsp_ctx* ctx = nullptr;
    const bsp_bsr* A = nullptr;
    int n = 0, bs = 0, b2 = 0, E = 0;
    int64_t nnzb = 0;
    DevMem e_ptr, e_src, e_dst, e_slot, e_rev, diag_slot, slot_row, slot_edge, tslot, lower;
    DevMem sd, s32, x, ef, ef2, lu32;
    DevMem h[kLayers + 1], phi, msg[kLayers], agg[kLayers], upd[kLayers], dD, gl, F;
    DevMem gh, gh2, ga, ge1, ge2, gphi, g_dD, g_gl, gF, partial;
    size_t partial_floats = 0;
    int kcap = 0;
    DevMem z32, r32, r64, u, y, gy, gu, coef, sums;
    int sum_blocks = 0;
    int alloc_k(int K);
};

namespace {

uint64_t splitmix64(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

int graph_forward(bsp_ctx* ctx, const bsp_mpnn* net, bsp_mpnn_graph* g) {
    const int n = g->n, E = g->E, b2 = g->b2;
    const int* esrc = g->e_src.as<int>();
    const int* edst = g->e_dst.as<int>();
    int rc;
    Part px{g->x.as<float>(), nullptr, b2, b2};
    if ((rc = linear_fwd(ctx, n, kHidden, px, nullptr, b2, net->P(net->oWn), net->P(net->obn), g->h[0].as<float>(),
                         kHidden, true, nullptr, nullptr)))
        return rc;
    Part pe0{g->ef.as<float>(), nullptr, b2, b2}, pe1{g->ef2.as<float>(), nullptr, 1, 1};
    if ((rc = linear_fwd(ctx, E, kHidden, pe0, &pe1, b2 + 1, net->P(net->oWe), net->P(net->obe), g->phi.as<float>(),
                         kHidden, true, nullptr, nullptr)))
        return rc;
    for (int l = 0; l < kLayers; ++l) {
        Part a0{g->phi.as<float>(), nullptr, kHidden, kHidden}, a1{g->h[l].as<float>(), edst, kHidden, kHidden};
        if ((rc = linear_fwd(ctx, E, kHidden, a0, &a1, 2 * kHidden, net->P(net->oWm[l]), net->P(net->obm[l]),
                             g->msg[l].as<float>(), kHidden, true, nullptr, nullptr)))
            return rc;
        if ((rc = segsum(ctx, n, g->e_ptr.as<int>(), nullptr, g->msg[l].as<float>(), g->agg[l].as<float>(), false)))
            return rc;
        Part u0{g->h[l].as<float>(), nullptr, kHidden, kHidden}, u1{g->agg[l].as<float>(), nullptr, kHidden, kHidden};
        if ((rc = linear_fwd(ctx, n, kHidden, u0, &u1, 2 * kHidden, net->P(net->oWu[l]), net->P(net->obu[l]),
                             g->upd[l].as<float>(), kHidden, true, g->h[l].as<float>(), g->h[l + 1].as<float>())))
            return rc;
    }
    Part hd{g->h[kLayers].as<float>(), nullptr, kHidden, kHidden};
    if ((rc = linear_fwd(ctx, n, b2, hd, nullptr, kHidden, net->P(net->oWd), net->P(net->obd), g->dD.as<float>(), b2,
                         false, nullptr, nullptr)))
        return rc;
    Part hs{g->h[kLayers].as<float>(), esrc, kHidden, kHidden}, ht{g->h[kLayers].as<float>(), edst, kHidden, kHidden};
    return linear_fwd(ctx, E, 1, hs, &ht, 2 * kHidden, net->P(net->owg), net->P(net->obg), g->gl.as<float>(), 1, false,
                      nullptr, nullptr);
}

// Backpropagates g_dD / g_gl (already in g) through the network into net->grad.
int graph_backward(bsp_ctx* ctx, bsp_mpnn* net, bsp_mpnn_graph* g) {
    const int n = g->n, E = g->E, b2 = g->b2;
    const int* esrc = g->e_src.as<int>();
    const int* edst = g->e_dst.as<int>();
    const int* eptr = g->e_ptr.as<int>();
    const int* erev = g->e_rev.as<int>();
    float* part = g->partial.as<float>();
    const size_t pf = g->partial_floats;
    const float* h3 = g->h[kLayers].as<float>();
    int rc;
    // diagonal head
    GradIn gd{g->g_dD.as<float>(), nullptr, b2, nullptr, 0};
    Part ph{h3, nullptr, kHidden, kHidden};
    if ((rc = linear_bwd_w(ctx, n, b2, ph, nullptr, kHidden, gd, part, pf, net->G(net->oWd), net->G(net->obd))))
        return rc;
    float* gh = g->gh.as<float>();
    float* gh2 = g->gh2.as<float>();
    if ((rc = linear_bwd_in(ctx, n, b2, kHidden, 0, kHidden, net->P(net->oWd), gd, gh, kHidden, false, nullptr, 0,
                            false)))
        return rc;
    // edge-scaling head
    GradIn gg{g->g_gl.as<float>(), nullptr, 1, nullptr, 0};
    Part hs{h3, esrc, kHidden, kHidden}, ht{h3, edst, kHidden, kHidden};
    if ((rc = linear_bwd_w(ctx, E, 1, hs, &ht, 2 * kHidden, gg, part, pf, net->G(net->owg), net->G(net->obg))))
        return rc;
    if (E > 0) {
        if ((rc = linear_bwd_in(ctx, E, 1, kHidden, kHidden, 2 * kHidden, net->P(net->owg), gg, g->ge1.as<float>(),
                                kHidden, false, g->ge2.as<float>(), kHidden, false)))
            return rc;
        if ((rc = segsum(ctx, n, eptr, nullptr, g->ge1.as<float>(), gh, true))) return rc;
        if ((rc = segsum(ctx, n, eptr, erev, g->ge2.as<float>(), gh, true))) return rc;
    }
    BSP_HIP(hipMemsetAsync(g->gphi.p, 0, sizeof(float) * std::max<size_t>((size_t)E * kHidden, 1), ctx->stream));
    for (int l = kLayers - 1; l >= 0; --l) {
        const float* hl = g->h[l].as<float>();
        // update layer h_{l+1} = h_l + relu(Wu [h_l ; a_l] + bu)
        GradIn gu{gh, nullptr, kHidden, g->upd[l].as<float>(), kHidden};
        Part u0{hl, nullptr, kHidden, kHidden}, u1{g->agg[l].as<float>(), nullptr, kHidden, kHidden};
        if ((rc = linear_bwd_w(ctx, n, kHidden, u0, &u1, 2 * kHidden, gu, part, pf, net->G(net->oWu[l]),
                               net->G(net->obu[l]))))
            return rc;
        BSP_HIP(hipMemcpyAsync(gh2, gh, sizeof(float) * (size_t)n * kHidden, hipMemcpyDeviceToDevice, ctx->stream));
        if ((rc = linear_bwd_in(ctx, n, kHidden, kHidden, kHidden, 2 * kHidden, net->P(net->oWu[l]), gu, gh2, kHidden,
                                true, g->ga.as<float>(), kHidden, false)))
            return rc;
        // message layer m_e = relu(Wm [phi_e ; h_l[dst e]] + bm), a_i = sum over row i's edges
        if (E > 0) {
            GradIn gm{g->ga.as<float>(), esrc, kHidden, g->msg[l].as<float>(), kHidden};
            Part m0{g->phi.as<float>(), nullptr, kHidden, kHidden}, m1{hl, edst, kHidden, kHidden};
            if ((rc = linear_bwd_w(ctx, E, kHidden, m0, &m1, 2 * kHidden, gm, part, pf, net->G(net->oWm[l]),
                                   net->G(net->obm[l]))))
                return rc;
            if ((rc = linear_bwd_in(ctx, E, kHidden, kHidden, kHidden, 2 * kHidden, net->P(net->oWm[l]), gm,
                                    g->gphi.as<float>(), kHidden, true, g->ge2.as<float>(), kHidden, false)))
                return rc;
            if ((rc = segsum(ctx, n, eptr, erev, g->ge2.as<float>(), gh2, true))) return rc;
        } else {
            BSP_HIP(hipMemsetAsync(net->G(net->oWm[l]), 0, sizeof(float) * kHidden * 2 * kHidden, ctx->stream));
            BSP_HIP(hipMemsetAsync(net->G(net->obm[l]), 0, sizeof(float) * kHidden, ctx->stream));
        }
        std::swap(gh, gh2);
    }
    // encoders
    GradIn ge{g->gphi.as<float>(), nullptr, kHidden, g->phi.as<float>(), kHidden};
    Part pe0{g->ef.as<float>(), nullptr, b2, b2}, pe1{g->ef2.as<float>(), nullptr, 1, 1};
    if (E > 0) {
        if ((rc = linear_bwd_w(ctx, E, kHidden, pe0, &pe1, b2 + 1, ge, part, pf, net->G(net->oWe), net->G(net->obe))))
            return rc;
    } else {
        BSP_HIP(hipMemsetAsync(net->G(net->oWe), 0, sizeof(float) * kHidden * (b2 + 1), ctx->stream));
        BSP_HIP(hipMemsetAsync(net->G(net->obe), 0, sizeof(float) * kHidden, ctx->stream));
    }
    GradIn gn{gh, nullptr, kHidden, g->h[0].as<float>(), kHidden};
    Part px{g->x.as<float>(), nullptr, b2, b2};
    return linear_bwd_w(ctx, n, kHidden, px, nullptr, b2, gn, part, pf, net->G(net->oWn), net->G(net->obn));
}

int check_net_graph(bsp_ctx* ctx, const bsp_mpnn* net, const bsp_mpnn_graph* g) {
    BSP_CHECK(net->bs == g->bs, BSP_E_INVALID, "network and graph block sizes differ");
    return BSP_OK;
}

}  // namespace

int bsp_mpnn_graph::alloc_k(int K) {
    if (K <= kcap) return BSP_OK;
    const size_t len = (size_t)n * bs;
    const size_t cnt = len * (size_t)K;
    BSP_HIP(z32.alloc(sizeof(float) * cnt));
    BSP_HIP(r32.alloc(sizeof(float) * cnt));
    BSP_HIP(r64.alloc(sizeof(double) * len));
    BSP_HIP(u.alloc(sizeof(float) * cnt));
    BSP_HIP(y.alloc(sizeof(float) * cnt));
    BSP_HIP(gy.alloc(sizeof(float) * cnt));
    BSP_HIP(gu.alloc(sizeof(float) * cnt));
    BSP_HIP(coef.alloc(sizeof(float) * K));
    sum_blocks = (int)std::max<int64_t>(1, std::min<int64_t>(256, ((int64_t)len + 1023) / 1024));
    BSP_HIP(sums.alloc(sizeof(double) * 2 * (size_t)K * sum_blocks));
    kcap = K;
    return BSP_OK;
}

extern "C" {

int bsp_mpnn_create(bsp_ctx* ctx, int bs, uint64_t seed, bsp_mpnn** out) {
    if (!ctx || !out) return BSP_E_INVALID;
    *out = nullptr;
    BSP_CHECK(bs >= 1 && bs <= 8, BSP_E_UNSUPPORTED, "the learned correction supports block sizes 1..8");
    BSP_HIP(hipSetDevice(ctx->device));
    auto net = std::make_unique<bsp_mpnn>();
    net->ctx = ctx;
    net->bs = bs;
    net->b2 = bs * bs;
    const int b2 = net->b2, Hd = kHidden;
    int64_t off = 0;
    auto take = [&](int64_t cnt) {
        const int64_t o = off;
        off += cnt;
        return o;
    };
    net->oWn = take((int64_t)Hd * b2);
    net->obn = take(Hd);
    net->oWe = take((int64_t)Hd * (b2 + 1));
    net->obe = take(Hd);
    for (int l = 0; l < kLayers; ++l) {
        net->oWm[l] = take((int64_t)Hd * 2 * Hd);
        net->obm[l] = take(Hd);
        net->oWu[l] = take((int64_t)Hd * 2 * Hd);
        net->obu[l] = take(Hd);
    }
    net->oWd = take((int64_t)b2 * Hd);
    net->obd = take(b2);
    net->owg = take(2 * Hd);
    net->obg = take(1);
    net->nparams = off;
    // trunk: uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)); heads zero
    std::vector<float> th(off, 0.0f);
    uint64_t s = seed;
    auto fill = [&](int64_t o, int64_t cnt, int fan_in) {
        const double lim = 1.0 / std::sqrt((double)fan_in);
        for (int64_t t = 0; t < cnt; ++t) {
            const double u01 = (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0);
            th[o + t] = (float)((2.0 * u01 - 1.0) * lim);
        }
    };
    fill(net->oWn, (int64_t)Hd * b2, b2);
    fill(net->obn, Hd, b2);
    fill(net->oWe, (int64_t)Hd * (b2 + 1), b2 + 1);
    fill(net->obe, Hd, b2 + 1);
    for (int l = 0; l < kLayers; ++l) {
        fill(net->oWm[l], (int64_t)Hd * 2 * Hd, 2 * Hd);
        fill(net->obm[l], Hd, 2 * Hd);
        fill(net->oWu[l], (int64_t)Hd * 2 * Hd, 2 * Hd);
        fill(net->obu[l], Hd, 2 * Hd);
    }
    const size_t bytes = sizeof(float) * (size_t)off;
    BSP_HIP(net->theta.alloc(bytes));
    BSP_HIP(net->grad.alloc(bytes));
    BSP_HIP(net->m.alloc(bytes));
    BSP_HIP(net->v.alloc(bytes));
    BSP_HIP(hipMemcpy(net->theta.p, th.data(), bytes, hipMemcpyHostToDevice));
    BSP_HIP(hipMemset(net->grad.p, 0, bytes));
    BSP_HIP(hipMemset(net->m.p, 0, bytes));
    BSP_HIP(hipMemset(net->v.p, 0, bytes));
    *out = net.release();
    return BSP_OK;
}

int bsp_mpnn_destroy(bsp_mpnn* net) {
    delete net;
    return BSP_OK;
}

int bsp_mpnn_num_params(const bsp_mpnn* net, int64_t* count) {
    if (!net || !count) return BSP_E_INVALID;
    *count = net->nparams;
    return BSP_OK;
}

int bsp_mpnn_get_params(bsp_ctx* ctx, const bsp_mpnn* net, float* host) {
    if (!ctx || !net || !host) return BSP_E_INVALID;
    BSP_HIP(hipStreamSynchronize(ctx->stream));
    BSP_HIP(hipMemcpy(host, net->theta.p, sizeof(float) * net->nparams, hipMemcpyDeviceToHost));
    return BSP_OK;
}

int bsp_mpnn_set_params(bsp_ctx* ctx, bsp_mpnn* net, const float* host) {
    if (!ctx || !net || !host) return BSP_E_INVALID;
    const size_t bytes = sizeof(float) * net->nparams;
    BSP_HIP(hipStreamSynchronize(ctx->stream));
    BSP_HIP(hipMemcpy(net->theta.p, host, bytes, hipMemcpyHostToDevice));
    BSP_HIP(hipMemset(net->m.p, 0, bytes));
    BSP_HIP(hipMemset(net->v.p, 0, bytes));
    net->step = 0;
    return BSP_OK;
}

int bsp_mpnn_get_grads(bsp_ctx* ctx, const bsp_mpnn* net, float* host) {
    if (!ctx || !net || !host) return BSP_E_INVALID;
    BSP_HIP(hipStreamSynchronize(ctx->stream));
    BSP_HIP(hipMemcpy(host, net->grad.p, sizeof(float) * net->nparams, hipMemcpyDeviceToHost));
    return BSP_OK;
}

int bsp_mpnn_graph_create(bsp_ctx* ctx, const bsp_bsr* A, const bsp_precond* M, bsp_mpnn_graph** out) {
    if (!ctx || !A || !M || !out) return BSP_E_INVALID;
    *out = nullptr;
    BSP_CHECK(A->n_rows == A->n_cols, BSP_E_INVALID, "the learned correction needs a square matrix");
    BSP_CHECK(A->cbs == 0, BSP_E_UNSUPPORTED, "the learned correction does not take complex matrices");
    BSP_CHECK(A->bs >= 1 && A->bs <= 8, BSP_E_UNSUPPORTED, "the learned correction supports block sizes 1..8");
    BSP_CHECK(M->kind == BSP_PRE_ILU0 && M->n == A->n_rows && M->bs == A->bs && (!M->f_val.empty() || M->fdev.p),
              BSP_E_INVALID, "the factor to correct must be an ILU0 preconditioner of A");
    BSP_HIP(hipSetDevice(ctx->device));
    const int n = A->n_rows, bs = A->bs, b2 = bs * bs;
    const auto& rp = A->h_rp;
    const auto& ci = A->h_ci;
    const int64_t nnzb = A->nnzb;
    std::vector<int32_t> pos(n);
    for (int t = 0; t < n; ++t) pos[M->elim_order[t]] = t;
    std::vector<int32_t> e_ptr(n + 1, 0), e_src, e_dst, e_slot, diag(n, -1), slot_row(nnzb), slot_edge(nnzb, -1),
        tslot(nnzb, -1);
    std::vector<uint8_t> lower(nnzb, 0);
    for (int i = 0; i < n; ++i) {
        for (int p = rp[i]; p < rp[i + 1]; ++p) {
            const int j = ci[p];
            slot_row[p] = i;
            lower[p] = pos[j] < pos[i] ? 1 : 0;
            if (j == i) {
                BSP_CHECK(diag[i] < 0, BSP_E_INVALID, "duplicate diagonal block in row " + std::to_string(i));
                diag[i] = p;
            } else {
                slot_edge[p] = (int32_t)e_src.size();
                e_src.push_back(i);
                e_dst.push_back(j);
                e_slot.push_back(p);
            }
        }
        BSP_CHECK(diag[i] >= 0, BSP_E_INVALID, "missing diagonal block in row " + std::to_string(i));
        e_ptr[i + 1] = (int32_t)e_src.size();
    }
    const int E = (int)e_src.size();
    // transposed slots: rows' blocks sorted by column for lookup
    {
        std::vector<std::pair<int32_t, int32_t>> cols;
        std::vector<int32_t> cptr(n + 1, 0);
        cols.reserve(nnzb);
        for (int i = 0; i < n; ++i) {
            const size_t start = cols.size();
            for (int p = rp[i]; p < rp[i + 1]; ++p) cols.emplace_back(ci[p], p);
            std::sort(cols.begin() + start, cols.end());
            cptr[i + 1] = (int32_t)cols.size();
        }
        for (int i = 0; i < n; ++i)
            for (int p = rp[i]; p < rp[i + 1]; ++p) {
                const int j = ci[p];
                auto b = cols.begin() + cptr[j], e = cols.begin() + cptr[j + 1];
                auto it = std::lower_bound(b, e, std::make_pair((int32_t)i, (int32_t)-1));
                BSP_CHECK(it != e && it->first == i, BSP_E_UNSUPPORTED,
                          "the learned correction needs a structurally symmetric block pattern");
                tslot[p] = it->second;
            }
    }
    std::vector<int32_t> e_rev(E);
    for (int e = 0; e < E; ++e) e_rev[e] = slot_edge[tslot[e_slot[e]]];
    auto g = std::make_unique<bsp_mpnn_graph>();
    g->ctx = ctx;
    g->A = A;
    g->n = n;
    g->bs = bs;
    g->b2 = b2;
    g->E = E;
    g->nnzb = nnzb;
    auto up_i = [&](DevMem& m, const std::vector<int32_t>& v) -> hipError_t {
        hipError_t e = m.alloc(sizeof(int32_t) * std::max<size_t>(v.size(), 1));
        if (e != hipSuccess || v.empty()) return e;
        return hipMemcpy(m.p, v.data(), sizeof(int32_t) * v.size(), hipMemcpyHostToDevice);
    };
    BSP_HIP(up_i(g->e_ptr, e_ptr));
    BSP_HIP(up_i(g->e_src, e_src));
    BSP_HIP(up_i(g->e_dst, e_dst));
    BSP_HIP(up_i(g->e_slot, e_slot));
    BSP_HIP(up_i(g->e_rev, e_rev));
    BSP_HIP(up_i(g->diag_slot, diag));
    BSP_HIP(up_i(g->slot_row, slot_row));
    BSP_HIP(up_i(g->slot_edge, slot_edge));
    BSP_HIP(up_i(g->tslot, tslot));
    BSP_HIP(g->lower.alloc(std::max<size_t>(nnzb, 1)));
    BSP_HIP(hipMemcpy(g->lower.p, lower.data(), nnzb, hipMemcpyHostToDevice));
    if (M->fdev.p) {
        // the factor is on the device (A's block layout): FP32 copy there
        const int64_t cnt = nnzb * b2;
        BSP_HIP(g->lu32.alloc(sizeof(float) * std::max<int64_t>(cnt, 1)));
        if (int rc = launch_d2f(ctx, cnt, M->fdev.as<double>(), g->lu32.as<float>())) return rc;
        BSP_HIP(hipStreamSynchronize(ctx->stream));
    } else {
        std::vector<float> lu(M->f_val.begin(), M->f_val.end());
        BSP_HIP(g->lu32.alloc(sizeof(float) * std::max<size_t>(lu.size(), 1)));
        BSP_HIP(hipMemcpy(g->lu32.p, lu.data(), sizeof(float) * lu.size(), hipMemcpyHostToDevice));
    }
    const size_t nH = (size_t)n * kHidden, eH = std::max<size_t>((size_t)E * kHidden, 1);
    BSP_HIP(g->sd.alloc(sizeof(double) * n));
    BSP_HIP(g->s32.alloc(sizeof(float) * n));
    BSP_HIP(g->x.alloc(sizeof(float) * (size_t)n * b2));
    BSP_HIP(g->ef.alloc(sizeof(float) * std::max<size_t>((size_t)E * b2, 1)));
    BSP_HIP(g->ef2.alloc(sizeof(float) * std::max(E, 1)));
    for (int l = 0; l <= kLayers; ++l) BSP_HIP(g->h[l].alloc(sizeof(float) * nH));
    BSP_HIP(g->phi.alloc(sizeof(float) * eH));
    for (int l = 0; l < kLayers; ++l) {
        BSP_HIP(g->msg[l].alloc(sizeof(float) * eH));
        BSP_HIP(g->agg[l].alloc(sizeof(float) * nH));
        BSP_HIP(g->upd[l].alloc(sizeof(float) * nH));
    }
    BSP_HIP(g->dD.alloc(sizeof(float) * (size_t)n * b2));
    BSP_HIP(g->gl.alloc(sizeof(float) * std::max(E, 1)));
    BSP_HIP(g->F.alloc(sizeof(float) * (size_t)nnzb * b2));
    BSP_HIP(g->gh.alloc(sizeof(float) * nH));
    BSP_HIP(g->gh2.alloc(sizeof(float) * nH));
    BSP_HIP(g->ga.alloc(sizeof(float) * nH));
    BSP_HIP(g->ge1.alloc(sizeof(float) * eH));
    BSP_HIP(g->ge2.alloc(sizeof(float) * eH));
    BSP_HIP(g->gphi.alloc(sizeof(float) * eH));
    BSP_HIP(g->g_dD.alloc(sizeof(float) * (size_t)n * b2));
    BSP_HIP(g->g_gl.alloc(sizeof(float) * std::max(E, 1)));
    BSP_HIP(g->gF.alloc(sizeof(float) * (size_t)nnzb * b2));
    g->partial_floats = (size_t)kMaxPartialBlocks * kHidden * (2 * kHidden + 1);
    BSP_HIP(g->partial.alloc(sizeof(float) * g->partial_floats));
    int rc = features(ctx, n, E, b2, g->diag_slot.as<int>(), g->e_src.as<int>(), g->e_dst.as<int>(),
                      g->e_slot.as<int>(), A->val.as<double>(), g->sd.as<double>(), g->s32.as<float>(),
                      g->x.as<float>(), g->ef.as<float>(), g->ef2.as<float>());
    if (rc) return rc;
    BSP_HIP(hipStreamSynchronize(ctx->stream));
    *out = g.release();
    return BSP_OK;
}

int bsp_mpnn_graph_destroy(bsp_mpnn_graph* g) {
    delete g;
    return BSP_OK;
}

int bsp_mpnn_infer(bsp_ctx* ctx, const bsp_mpnn* net, bsp_mpnn_graph* g, float* dD_host, float* g_host) {
    if (!ctx || !net || !g) return BSP_E_INVALID;
    int rc = check_net_graph(ctx, net, g);
    if (rc) return rc;
    BSP_HIP(hipSetDevice(ctx->device));
    if ((rc = graph_forward(ctx, net, g))) return rc;
    BSP_HIP(hipStreamSynchronize(ctx->stream));
    if (dD_host) BSP_HIP(hipMemcpy(dD_host, g->dD.p, sizeof(float) * (size_t)g->n * g->b2, hipMemcpyDeviceToHost));
    if (g_host && g->E) BSP_HIP(hipMemcpy(g_host, g->gl.p, sizeof(float) * (size_t)g->E, hipMemcpyDeviceToHost));
    return BSP_OK;
}

int bsp_mpnn_loss_grad(bsp_ctx* ctx, bsp_mpnn* net, bsp_mpnn_graph* g, const double* Z_dev, int K, double* loss) {
    if (!ctx || !net || !g || !Z_dev || K <= 0) return BSP_E_INVALID;
    int rc = check_net_graph(ctx, net, g);
    if (rc) return rc;
    BSP_HIP(hipSetDevice(ctx->device));
    if ((rc = g->alloc_k(K))) return rc;
    const int n = g->n, bs = g->bs, b2 = g->b2;
    const int64_t len = (int64_t)n * bs;
    const bsp_bsr* A = g->A;
    float* z32 = g->z32.as<float>();
    float* r32 = g->r32.as<float>();
    if ((rc = launch_d2f(ctx, len * K, Z_dev, z32))) return rc;
    for (int k = 0; k < K; ++k) {
        if ((rc = bsmv_matrix(ctx, A, Z_dev + k * len,
                              g->r64.as<double>(), 0)))
            return rc;
        if ((rc = launch_d2f(ctx, len, g->r64.as<double>(), r32 + k * len))) return rc;
    }
    if ((rc = graph_forward(ctx, net, g))) return rc;
    if ((rc = assemble(ctx, g->nnzb, bs, g->slot_row.as<int>(), g->slot_edge.as<int>(), g->lu32.as<float>(),
                       g->dD.as<float>(), g->gl.as<float>(), kAlpha, kBeta, g->F.as<float>())))
        return rc;
    const int* rp = A->rp.as<int>();
    const int* ci = A->ci.as<int>();
    const unsigned char* lo = g->lower.as<unsigned char>();
    const float* F = g->F.as<float>();
    if ((rc = factor_prod(ctx, true, n, bs, K, rp, ci, lo, F, z32, nullptr, g->u.as<float>()))) return rc;
    if ((rc = factor_prod(ctx, false, n, bs, K, rp, ci, lo, F, g->u.as<float>(), g->u.as<float>(), g->y.as<float>())))
        return rc;
    if ((rc = loss_sums(ctx, len, K, g->sum_blocks, g->y.as<float>(), r32, g->sums.as<double>()))) return rc;
    std::vector<double> sums(2 * (size_t)K * g->sum_blocks);
    BSP_HIP(hipMemcpyAsync(sums.data(), g->sums.p, sizeof(double) * sums.size(), hipMemcpyDeviceToHost, ctx->stream));
    BSP_HIP(hipStreamSynchronize(ctx->stream));
    std::vector<float> coef(K);
    double L = 0.0;
    for (int k = 0; k < K; ++k) {
        double a = 0.0, b = 0.0;
        for (int t = 0; t < g->sum_blocks; ++t) {
            a += sums[((size_t)k * g->sum_blocks + t) * 2];
            b += sums[((size_t)k * g->sum_blocks + t) * 2 + 1];
        }
        BSP_CHECK(b > 0.0 && std::isfinite(b), BSP_E_NUMERIC, "a sample vector has A z = 0 or non-finite entries");
        L += a / b;
        coef[k] = (float)(2.0 / (K * b));
    }
    L /= K;
    if (loss) *loss = L;
    BSP_HIP(hipMemcpy(g->coef.p, coef.data(), sizeof(float) * K, hipMemcpyHostToDevice));
    if ((rc = grad_y(ctx, len, K, g->coef.as<float>(), g->y.as<float>(), r32, g->gy.as<float>()))) return rc;
    if ((rc = grad_u(ctx, n, bs, K, rp, ci, g->tslot.as<int>(), lo, F, g->gy.as<float>(), g->gu.as<float>())))
        return rc;
    if ((rc = grad_F(ctx, g->nnzb, bs, K, len, g->slot_row.as<int>(), ci, lo, g->gy.as<float>(), g->u.as<float>(),
                     g->gu.as<float>(), z32, g->gF.as<float>())))
        return rc;
    if ((rc = grad_heads(ctx, n, g->E, bs, g->diag_slot.as<int>(), g->e_slot.as<int>(), g->lu32.as<float>(),
                         g->dD.as<float>(), g->gF.as<float>(), g->gl.as<float>(), kAlpha, kBeta,
                         g->g_dD.as<float>(), g->g_gl.as<float>())))
        return rc;
    if ((rc = graph_backward(ctx, net, g))) return rc;
    BSP_HIP(hipStreamSynchronize(ctx->stream));
    return std::isfinite(L) ? BSP_OK : fail(ctx, BSP_E_NUMERIC, "training loss is not finite");
}

int bsp_mpnn_adam_step(bsp_ctx* ctx, bsp_mpnn* net, float lr) {
    if (!ctx || !net) return BSP_E_INVALID;
    BSP_HIP(hipSetDevice(ctx->device));
    const float b1 = 0.9f, b2 = 0.999f;
    ++net->step;
    const float c1 = (float)(1.0 / (1.0 - std::pow((double)b1, (double)net->step)));
    const float c2 = (float)(1.0 / (1.0 - std::pow((double)b2, (double)net->step)));
    int rc = adam(ctx, net->nparams, net->theta.as<float>(), net->grad.as<float>(), net->m.as<float>(),
                  net->v.as<float>(), lr, b1, b2, 1e-8f, c1, c2);
    if (rc) return rc;
    BSP_HIP(hipStreamSynchronize(ctx->stream));
    return BSP_OK;
}

int bsp_mpnn_train(bsp_ctx* ctx, bsp_mpnn* net, int ngraphs, bsp_mpnn_graph* const* graphs, const double* const* Z_dev,
                   const int* K, int epochs, int batch, float lr, double* epoch_loss) {
    if (!ctx || !net || ngraphs <= 0 || !graphs || !Z_dev || !K || epochs < 0 || batch <= 0) return BSP_E_INVALID;
    for (int e = 0; e < epochs; ++e) {
        double tot = 0.0;
        int steps = 0;
        int maxk = 0;
        for (int gi = 0; gi < ngraphs; ++gi) maxk = std::max(maxk, K[gi]);
        for (int b0 = 0; b0 < maxk; b0 += batch) {
            for (int gi = 0; gi < ngraphs; ++gi) {
                if (b0 >= K[gi]) continue;
                const int kk = std::min(batch, K[gi] - b0);
                const int64_t len = (int64_t)graphs[gi]->n * graphs[gi]->bs;
                double L = 0.0;
                int rc = bsp_mpnn_loss_grad(ctx, net, graphs[gi], Z_dev[gi] + (size_t)b0 * len, kk, &L);
                if (rc) return rc;
                if ((rc = bsp_mpnn_adam_step(ctx, net, lr))) return rc;
                tot += L;
                ++steps;
            }
        }
        if (epoch_loss) epoch_loss[e] = steps ? tot / steps : 0.0;
        if (ctx->verbose > 0) {
            std::fprintf(stderr, "[mpnn] epoch %d loss %.5f\n", e, steps ? tot / steps : 0.0);
            std::fflush(stderr);
        }
    }
    return BSP_OK;
}

int bsp_precond_create_learned(bsp_ctx* ctx, const bsp_bsr* A, const bsp_mpnn* net, const int32_t* colour,
                               int precision, bsp_precond** out) {
    if (!ctx || !A || !net || !out) return BSP_E_INVALID;
    *out = nullptr;
    BSP_CHECK(precision == BSP_PREC_F64 || precision == BSP_PREC_F32 || precision == BSP_PREC_BF16, BSP_E_INVALID,
              "bad precision");
    BSP_CHECK(A->cbs == 0, BSP_E_UNSUPPORTED, "the learned correction does not take complex matrices");
    BSP_CHECK(net->bs == A->bs, BSP_E_INVALID, "network and matrix block sizes differ");
    BSP_HIP(hipSetDevice(ctx->device));
    std::vector<int32_t> elim, pos;
    std::vector<double> LU;
    int rc = ilu0_factor_host(ctx, A, colour, elim, pos, LU);
    if (rc) return rc;
    // a lightweight ILU0 holder for the graph (factor and order only)
    bsp_precond base;
    base.ctx = ctx;
    base.kind = BSP_PRE_ILU0;
    base.n = A->n_rows;
    base.bs = A->bs;
    base.elim_order = elim;
    base.f_val.swap(LU);
    bsp_mpnn_graph* graw = nullptr;
    if ((rc = bsp_mpnn_graph_create(ctx, A, &base, &graw))) return rc;
    std::unique_ptr<bsp_mpnn_graph> g(graw);
    LU.swap(base.f_val);
    const int n = A->n_rows, bs = A->bs, b2 = bs * bs;
    std::vector<float> dD((size_t)n * b2), gl(std::max(g->E, 1));
    if ((rc = bsp_mpnn_infer(ctx, net, g.get(), dD.data(), gl.data()))) return rc;
    const auto& rp = A->h_rp;
    const auto& ci = A->h_ci;
    int e = 0;
    std::vector<double> T(b2), L0(b2);
    for (int i = 0; i < n; ++i)
        for (int p = rp[i]; p < rp[i + 1]; ++p) {
            double* f = LU.data() + (size_t)p * b2;
            if (ci[p] == i) {
                // F_ii = LU_ii (I + alpha tanh(D_i))
                for (int q = 0; q < b2; ++q) T[q] = std::tanh((double)dD[(size_t)i * b2 + q]);
                std::copy(f, f + b2, L0.begin());
                for (int r = 0; r < bs; ++r)
                    for (int c = 0; c < bs; ++c) {
                        double acc = 0.0;
                        for (int k = 0; k < bs; ++k) acc += L0[r * bs + k] * T[k * bs + c];
                        f[r * bs + c] = L0[r * bs + c] + (double)kAlpha * acc;
                    }
            } else {
                const double sc = 1.0 + (double)kBeta * std::tanh((double)gl[e++]);
                for (int q = 0; q < b2; ++q) f[q] *= sc;
            }
        }
    for (double v : LU) BSP_CHECK(std::isfinite(v), BSP_E_NUMERIC, "learned factor correction is not finite");
    return ilu0_build(ctx, A, elim, pos, LU, precision, out);
}

int bsp_ctx_record_basis(bsp_ctx* ctx, double* buf_dev, int64_t len, int cap, int every) {
    if (!ctx || cap < 0 || (cap > 0 && (!buf_dev || len <= 0))) return BSP_E_INVALID;
    ctx->rec_buf = cap > 0 ? buf_dev : nullptr;
    ctx->rec_cap = cap;
    ctx->rec_len = cap > 0 ? len : -1;
    ctx->rec_every = std::max(1, every);
    ctx->rec_count = 0;
    ctx->rec_seen = 0;
    return BSP_OK;
}

int bsp_ctx_recorded(const bsp_ctx* ctx, int* count) {
    if (!ctx || !count) return BSP_E_INVALID;
    *count = ctx->rec_count;
    return BSP_OK;
}

}  // extern "C"
