// Exercises the C-ABI entry points the drop-in program does not reach -- mixed-precision
// FGMRES over coloured FP32 / BF16 factors, the learned correction with the fallback solver,
// SpGEMM on 8 x 8 blocks, the BF16 block SpMM, the batched pCG with its tile choices and
// autotuners, BSR save / load, the orthogonalisation step (with and without allocation),
// coloured IChol0 / SOR, and the error returns -- so that the library's host code runs
// under the host sanitizers (make -C tests/cpp asan builds it as capi_test_asan).
// Prints "CAPI PASSED" when every check holds.
#include <blocksp_c.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <unistd.h>
#include <vector>

static int failures = 0;
static void report(const char* name, double value, double limit) {
    const bool ok = std::isfinite(value) && value <= limit;
    std::printf("%s %s %.3e (limit %.1e)\n", ok ? "OK" : "FAIL", name, value, limit);
    if (!ok) ++failures;
}
static void expect_rc(const char* name, int rc, int want) {
    std::printf("%s %s rc=%d (want %d)\n", rc == want ? "OK" : "FAIL", name, rc, want);
    if (rc != want) ++failures;
}
#define MUST(call)                                                                        \
    do {                                                                                  \
        const int rc_ = (call);                                                           \
        if (rc_ != BSP_OK) {                                                              \
            std::printf("FATAL %s -> %d (%s)\n", #call, rc_, bsp_last_error(ctx));        \
            std::exit(2);                                                                 \
        }                                                                                 \
    } while (0)

static unsigned long long g_state = 0x9e3779b97f4a7c15ull;
static double urand() {   // xorshift64 in [-1, 1)
    g_state ^= g_state << 13; g_state ^= g_state >> 7; g_state ^= g_state << 17;
    return (double)(g_state >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

struct Host {   // a BSR matrix on the host
    int n = 0, bs = 0;
    std::vector<int32_t> rp{0}, ci;
    std::vector<double> v;
};

// Block 5-point stencil on an nx x ny grid placed at block row `off` of a matrix with n rows:
// symmetric (A_ji = A_ij^T) and block-row diagonally dominant, hence SPD.
static void stencil(Host& H, int nx, int ny, int off) {
    const int B = H.bs, B2 = B * B;
    const double eps = 0.1;
    const double c = 4.0 * (1.0 + B * eps) + 1.0 + 2.0 * B * eps;
    auto cell = [&](int x, int y) { return off + y * nx + x; };
    for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) {
            const int i = cell(x, y);
            const int nb[4][2] = {{x, y - 1}, {x - 1, y}, {x + 1, y}, {x, y + 1}};
            std::vector<std::pair<int, int>> cols;   // (column, neighbour index or -1 for the diagonal)
            for (int k = 0; k < 4; ++k) {
                const int xx = nb[k][0], yy = nb[k][1];
                if (xx >= 0 && xx < nx && yy >= 0 && yy < ny) cols.push_back({cell(xx, yy), k});
            }
            cols.insert(cols.begin() + (int)cols.size() / 2, {i, -1});
            for (auto [j, k] : cols) {
                H.ci.push_back(j);
                const size_t base = H.v.size();
                H.v.resize(base + B2);
                if (k < 0) {
                    for (int r = 0; r < B; ++r)
                        for (int s = 0; s <= r; ++s) {
                            const double e = r == s ? c + eps * urand() : eps * urand();
                            H.v[base + r * B + s] = e;
                            H.v[base + s * B + r] = e;
                        }
                } else if (j > i) {
                    std::vector<double> blk(B2);
                    for (int q = 0; q < B2; ++q) blk[q] = (q % (B + 1) == 0 ? -1.0 : 0.0) + eps * urand();
                    for (int q = 0; q < B2; ++q) H.v[base + q] = blk[q];
                } else {
                    // transpose of block (j, i), already stored in row j
                    int p = H.rp[j];
                    while (H.ci[p] != i) ++p;
                    for (int r = 0; r < B; ++r)
                        for (int s = 0; s < B; ++s) H.v[base + r * B + s] = H.v[(size_t)p * B2 + s * B + r];
                }
            }
            H.rp.push_back((int32_t)H.ci.size());
        }
    H.n = (int)H.rp.size() - 1;
}

static bsp_ctx* ctx = nullptr;

static bsp_bsr* upload(const Host& H) {
    bsp_bsr* A = nullptr;
    MUST(bsp_bsr_create(ctx, H.n, H.n, H.bs, H.rp.data(), H.ci.data(), H.v.data(), &A));
    return A;
}
static double* dev(const std::vector<double>& h) {
    void* p = nullptr;
    MUST(bsp_malloc(ctx, h.size() * sizeof(double), &p));
    MUST(bsp_memcpy_h2d(ctx, p, h.data(), h.size() * sizeof(double)));
    return (double*)p;
}
static std::vector<double> host(const double* d, size_t len) {
    std::vector<double> h(len);
    MUST(bsp_memcpy_d2h(ctx, h.data(), d, len * sizeof(double)));
    return h;
}
static std::vector<double> random_vec(size_t len) {
    std::vector<double> h(len);
    for (auto& e : h) e = urand();
    return h;
}
// ||b - A x|| / ||b|| through the library's own bsmv (FP64)
static double relres(const bsp_bsr* A, int64_t len, const double* b, const double* x) {
    const std::vector<double> bh = host(b, len);
    double* t = dev(std::vector<double>(len, 0.0));
    MUST(bsp_bsmv(ctx, A, x, t, 0));
    const std::vector<double> th = host(t, len);
    MUST(bsp_free(ctx, t));
    double rr = 0, bb = 0;
    for (int64_t i = 0; i < len; ++i) {
        rr += (bh[i] - th[i]) * (bh[i] - th[i]);
        bb += bh[i] * bh[i];
    }
    return std::sqrt(rr / bb);
}

int main() {
    std::setvbuf(stdout, nullptr, _IOLBF, 0);   // keep the report if a sanitizer ends the run
    if (bsp_ctx_create(0, nullptr, &ctx) != BSP_OK) {
        std::printf("FATAL bsp_ctx_create failed (is a GPU visible?)\n");
        return 2;
    }
    // ---- 8 x 8 blocks on a 24 x 20 grid, parity colouring
    const int nx = 24, ny = 20;
    Host H;
    H.bs = 8;
    stencil(H, nx, ny, 0);
    const int64_t len = (int64_t)H.n * H.bs;
    std::vector<int32_t> parity(H.n);
    for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) parity[y * nx + x] = (x + y) & 1;
    bsp_bsr* A = upload(H);
    double* b = dev(random_vec(len));
    double* x = dev(std::vector<double>(len, 0.0));
    bsp_stats st{};

    for (int prec : {BSP_PREC_F32, BSP_PREC_BF16}) {
        bsp_precond* M = nullptr;
        MUST(bsp_precond_create_ilu0_ex(ctx, A, parity.data(), prec, &M));
        MUST(bsp_fgmres_mixed(ctx, A, M, b, x, 1e-10, 2000, 30, 10, BSP_PREC_F32, 0.0, &st));
        report(prec == BSP_PREC_F32 ? "fgmres_mixed_f32_factor" : "fgmres_mixed_bf16_factor",
               relres(A, len, b, x), 1.05e-10);
        MUST(bsp_precond_destroy(M));
    }
    {   // coloured IChol0-pCG and coloured SOR-pGMRES
        bsp_precond* M = nullptr;
        MUST(bsp_precond_create_ichol0_ex(ctx, A, parity.data(), &M));
        MUST(bsp_pcg(ctx, A, M, b, x, 1e-10, 2000, BSP_TOL_TRUE_RESIDUAL, &st));
        report("ichol0_coloured_pcg", relres(A, len, b, x), 1e-9);
        MUST(bsp_precond_destroy(M));
        MUST(bsp_precond_create_sor_ex(ctx, A, 1.0, 2, 2, 0, parity.data(), &M));
        MUST(bsp_pgmres(ctx, A, M, b, x, 1e-10, 2000, 30, 10, BSP_TOL_TRUE_RESIDUAL, &st));
        report("sor_coloured_pgmres", relres(A, len, b, x), 1e-9);
        int nlev = 0;
        MUST(bsp_precond_levels(M, 1, nullptr, &nlev));
        std::vector<int32_t> lev(H.n);
        MUST(bsp_precond_levels(M, 1, lev.data(), &nlev));
        report("sor_coloured_two_levels", std::abs(nlev - 2.0), 0.0);
        MUST(bsp_precond_destroy(M));
    }
    {   // learned correction: record the Krylov basis of a pGMRES solve, train, solve with fallback
        bsp_precond* M0 = nullptr;
        MUST(bsp_precond_create_ilu0_ex(ctx, A, parity.data(), BSP_PREC_F64, &M0));
        const int cap = 24;
        double* rec = dev(std::vector<double>((size_t)cap * len, 0.0));
        MUST(bsp_ctx_record_basis(ctx, rec, len, cap, 1));
        MUST(bsp_pgmres(ctx, A, M0, b, x, 1e-10, 2000, 30, 10, BSP_TOL_REFERENCE, &st));
        int count = 0;
        MUST(bsp_ctx_recorded(ctx, &count));
        MUST(bsp_ctx_record_basis(ctx, nullptr, 0, 0, 1));
        report("recorded_basis_vectors", count > 0 && count <= cap ? 0.0 : 1.0, 0.0);
        bsp_mpnn* net = nullptr;
        bsp_mpnn_graph* g = nullptr;
        MUST(bsp_mpnn_create(ctx, H.bs, 7, &net));
        MUST(bsp_mpnn_graph_create(ctx, A, M0, &g));
        int64_t np = 0;
        MUST(bsp_mpnn_num_params(net, &np));
        std::vector<float> params(np), grads(np);
        MUST(bsp_mpnn_get_params(ctx, net, params.data()));
        double loss0 = 0, loss1 = 0;
        MUST(bsp_mpnn_loss_grad(ctx, net, g, rec, count, &loss0));
        MUST(bsp_mpnn_get_grads(ctx, net, grads.data()));
        MUST(bsp_mpnn_adam_step(ctx, net, 1e-3f));
        const double* zs[1] = {rec};
        const int ks[1] = {count};
        std::vector<double> epoch_loss(3);
        bsp_mpnn_graph* gs[1] = {g};
        MUST(bsp_mpnn_train(ctx, net, 1, gs, zs, ks, 3, 8, 1e-3f, epoch_loss.data()));
        MUST(bsp_mpnn_loss_grad(ctx, net, g, rec, count, &loss1));
        report("mpnn_losses_finite", std::isfinite(loss0) && std::isfinite(loss1) ? 0.0 : 1.0, 0.0);
        std::vector<float> dD((size_t)H.n * H.bs * H.bs), gg(H.ci.size());
        MUST(bsp_mpnn_infer(ctx, net, g, dD.data(), gg.data()));
        bsp_precond* ML = nullptr;
        MUST(bsp_precond_create_learned(ctx, A, net, parity.data(), BSP_PREC_F32, &ML));
        bsp_precond* MF = nullptr;
        MUST(bsp_precond_create_ilu0_ex(ctx, A, parity.data(), BSP_PREC_F32, &MF));
        MUST(bsp_fgmres_fallback(ctx, A, ML, MF, b, x, 1e-10, 2000, 30, 10, BSP_PREC_F32, 0.0, &st));
        report("learned_fgmres_fallback", relres(A, len, b, x), 1.05e-10);
        MUST(bsp_mpnn_set_params(ctx, net, params.data()));
        MUST(bsp_precond_destroy(MF));
        MUST(bsp_precond_destroy(ML));
        MUST(bsp_mpnn_graph_destroy(g));
        MUST(bsp_mpnn_destroy(net));
        MUST(bsp_free(ctx, rec));
        MUST(bsp_precond_destroy(M0));
    }
    {   // SpGEMM on 8 x 8 blocks: (A A) v = A (A v), A^T A likewise, recomputed after new values
        bsp_bsr *C = nullptr, *Ct = nullptr;
        MUST(bsp_bsmm(ctx, A, A, &C));
        MUST(bsp_bsmm_trans_a(ctx, A, A, &Ct));
        double* v = dev(random_vec(len));
        double* t1 = dev(std::vector<double>(len, 0.0));
        double* t2 = dev(std::vector<double>(len, 0.0));
        auto check = [&](const char* name, bsp_bsr* P) {
            MUST(bsp_bsmv(ctx, A, v, t1, 0));
            MUST(bsp_bsmv(ctx, A, t1, t2, 0));   // A is symmetric, so A^T A = A A
            MUST(bsp_bsmv(ctx, P, v, t1, 0));
            const auto a = host(t1, len), e = host(t2, len);
            double num = 0, den = 0;
            for (int64_t i = 0; i < len; ++i) { num = std::max(num, std::abs(a[i] - e[i])); den = std::max(den, std::abs(e[i])); }
            report(name, num / den, 1e-13);
        };
        check("bsmm_AA", C);
        check("bsmm_AtA", Ct);
        std::vector<double> v2 = H.v;
        for (auto& e : v2) e *= 0.5;
        double* vals = dev(v2);
        MUST(bsp_bsr_set_values_dev(ctx, A, vals));
        MUST(bsp_bsmm_numeric(ctx, A, A, C));
        check("bsmm_numeric_new_values", C);
        MUST(bsp_bsr_set_values_dev(ctx, A, dev(H.v)));   // back to the original values (freed with the context)
        MUST(bsp_free(ctx, vals));
        MUST(bsp_bsr_destroy(Ct));
        MUST(bsp_bsr_destroy(C));
        MUST(bsp_free(ctx, v));
        MUST(bsp_free(ctx, t1));
        MUST(bsp_free(ctx, t2));
    }
    {   // BF16 block SpMM against the FP64 multi-RHS product
        const int nrhs = 32;
        double* X = dev(random_vec((size_t)len * nrhs));
        double* Y = dev(std::vector<double>((size_t)len * nrhs, 0.0));
        double* Y2 = dev(std::vector<double>((size_t)len * nrhs, 0.0));
        void *Xb = nullptr, *Yb = nullptr;
        MUST(bsp_malloc(ctx, (size_t)len * nrhs * 2, &Xb));
        MUST(bsp_malloc(ctx, (size_t)len * nrhs * 4, &Yb));
        MUST(bsp_bsmm_rhs(ctx, A, nrhs, X, Y));
        MUST(bsp_rhs_pack_bf16(ctx, H.n, H.bs, nrhs, X, (uint16_t*)Xb));
        MUST(bsp_bsmm_rhs_bf16(ctx, A, nrhs, (const uint16_t*)Xb, (float*)Yb));
        MUST(bsp_rhs_unpack_f32(ctx, H.n, H.bs, nrhs, (const float*)Yb, Y2));
        const auto a = host(Y2, (size_t)len * nrhs), e = host(Y, (size_t)len * nrhs);
        double num = 0, den = 0;
        for (size_t i = 0; i < a.size(); ++i) { num = std::max(num, std::abs(a[i] - e[i])); den = std::max(den, std::abs(e[i])); }
        report("bsmm_rhs_bf16_vs_fp64", num / den, 2e-2);
        expect_rc("bsmm_rhs_bf16_rejects_8_rhs", bsp_bsmm_rhs_bf16(ctx, A, 8, (const uint16_t*)Xb, (float*)Yb), BSP_E_INVALID);
        MUST(bsp_free(ctx, Xb));
        MUST(bsp_free(ctx, Yb));
        MUST(bsp_free(ctx, X));
        MUST(bsp_free(ctx, Y));
        MUST(bsp_free(ctx, Y2));
    }
    {   // BSR save / load round trip
        char path[] = "/tmp/bsp_capi_XXXXXX";
        const int fd = mkstemp(path);
        if (fd >= 0) close(fd);
        MUST(bsp_bsr_save(ctx, A, path));
        bsp_bsr* L = nullptr;
        MUST(bsp_bsr_load(ctx, path, &L));
        std::remove(path);
        int n2 = 0, m2 = 0, bs2 = 0;
        int64_t nnzb = 0;
        MUST(bsp_bsr_info(L, &n2, &m2, &bs2, &nnzb));
        std::vector<int32_t> rp(n2 + 1), ci(nnzb);
        std::vector<double> vv(nnzb * bs2 * bs2);
        MUST(bsp_bsr_download(ctx, L, rp.data(), ci.data(), vv.data()));
        report("bsr_save_load_bitwise", rp == H.rp && ci == H.ci && vv == H.v ? 0.0 : 1.0, 0.0);
        MUST(bsp_bsr_destroy(L));
        expect_rc("bsr_load_missing_file", bsp_bsr_load(ctx, "/nonexistent/dir/m.bsr", &L), BSP_E_INVALID);
    }
    {   // orthogonalisation step, allocating and allocation-free forms
        const int m = 5;
        std::vector<double*> Q(m + 1);
        for (int j = 0; j <= m; ++j) Q[j] = dev(std::vector<double>(len, 0.0));
        double* w = dev(random_vec(len));
        double nrm = 0;
        MUST(bsp_nrm2(ctx, len, w, &nrm));
        MUST(bsp_scal(ctx, len, 1.0 / nrm, w));
        MUST(bsp_memcpy_d2d(ctx, Q[0], w, len * sizeof(double)));
        std::vector<double> h(m + 2);
        for (int j = 1; j < m; ++j) {
            double* wj = dev(random_vec(len));
            MUST(bsp_orthogonalize(ctx, len, Q.data(), j, wj, h.data(), Q[j]));
            MUST(bsp_free(ctx, wj));
        }
        double worst = 0;
        for (int i = 0; i < m; ++i)
            for (int j = 0; j < m; ++j) {
                double d = 0;
                MUST(bsp_dot(ctx, len, Q[i], Q[j], &d));
                worst = std::max(worst, std::abs(d - (i == j ? 1.0 : 0.0)));
            }
        report("orthogonalize_orthonormal", worst, 1e-13);
        const std::vector<double> wh = random_vec(len);
        double* wa = dev(wh);
        double* wb = dev(wh);
        double* hd = dev(std::vector<double>(m + 1, 0.0));
        MUST(bsp_orthogonalize(ctx, len, Q.data(), m, wa, h.data(), Q[m]));
        const auto qa = host(Q[m], len);
        MUST(bsp_orthogonalize_noalloc(ctx, len, Q.data(), m, wb, hd, Q[m]));
        const auto qb = host(Q[m], len), hb = host(hd, m + 1);
        double diff = 0;
        for (int64_t i = 0; i < len; ++i) diff = std::max(diff, std::abs(qa[i] - qb[i]));
        for (int j = 0; j <= m; ++j) diff = std::max(diff, std::abs(h[j] - hb[j]));
        report("orthogonalize_noalloc_equals_alloc", diff, 1e-15);
        for (auto q : Q) MUST(bsp_free(ctx, q));
        MUST(bsp_free(ctx, w));
        MUST(bsp_free(ctx, wa));
        MUST(bsp_free(ctx, wb));
        MUST(bsp_free(ctx, hd));
    }
    {   // batched IChol0-pCG over uneven 4 x 4-block systems, tile choices and autotuners
        Host Bt;
        Bt.bs = 4;
        const int dims[5][2] = {{8, 6}, {13, 7}, {10, 5}, {20, 9}, {7, 7}};
        std::vector<int32_t> srp{0};
        for (auto& d : dims) {
            stencil(Bt, d[0], d[1], Bt.n);
            srp.push_back(Bt.n);
        }
        const int nsys = 5, nrhs = 2;
        const int64_t blen = (int64_t)Bt.n * Bt.bs;
        bsp_bsr* AB = upload(Bt);
        bsp_batch* bt = nullptr;
        MUST(bsp_batch_create(ctx, AB, nsys, srp.data(), nrhs, &bt));
        double* Bd = dev(random_vec((size_t)blen * nrhs));
        double* Xd = dev(std::vector<double>((size_t)blen * nrhs, 0.0));
        std::vector<int32_t> its(nsys * nrhs);
        std::vector<double> rel(nsys * nrhs);
        double ms = 0;
        auto check = [&](const char* name) {
            double worst = 0;
            for (int k = 0; k < nrhs; ++k) worst = std::max(worst, relres(AB, blen, Bd + k * blen, Xd + k * blen));
            report(name, worst, 1e-9);
        };
        MUST(bsp_batch_pcg(ctx, bt, Bd, Xd, 1e-10, 1000, BSP_TOL_TRUE_RESIDUAL, its.data(), rel.data(), &ms));
        check("batch_pcg_default_tiles");
        MUST(bsp_batch_set_tiles(ctx, bt, 256, 1));
        MUST(bsp_batch_pcg(ctx, bt, Bd, Xd, 1e-10, 1000, BSP_TOL_TRUE_RESIDUAL, its.data(), rel.data(), &ms));
        check("batch_pcg_256x1");
        MUST(bsp_batch_set_tiles(ctx, bt, 128, 2));
        MUST(bsp_batch_pcg(ctx, bt, Bd, Xd, 1e-10, 1000, BSP_TOL_TRUE_RESIDUAL, its.data(), rel.data(), &ms));
        check("batch_pcg_128x2");
        int wg = 0, rpi = 0, nitems = 0;
        MUST(bsp_batch_autotune(ctx, bt, Bd, Xd, 4, &wg, &rpi));
        MUST(bsp_batch_autotune_solve(ctx, bt, Bd, Xd, 1e-10, 1000, BSP_TOL_TRUE_RESIDUAL, 4, &wg, &rpi));
        check("batch_autotune_solve");
        MUST(bsp_batch_info(bt, &wg, &rpi, &nitems));
        MUST(bsp_batch_pcg(ctx, bt, Bd, Xd, 1e-10, 1000, BSP_TOL_TRUE_RESIDUAL, its.data(), rel.data(), &ms));
        check("batch_pcg_after_autotune");
        expect_rc("batch_pcg_not_converged",
                  bsp_batch_pcg(ctx, bt, Bd, Xd, 1e-14, 2, BSP_TOL_TRUE_RESIDUAL, its.data(), rel.data(), &ms),
                  BSP_NOT_CONVERGED);
        expect_rc("batch_set_tiles_rejects_3_rhs", bsp_batch_set_tiles(ctx, bt, 256, 3), BSP_E_INVALID);
        MUST(bsp_batch_destroy(bt));
        std::vector<int32_t> colour(Bt.n);
        for (int i = 0; i < Bt.n; ++i) colour[i] = i & 1;   // any colouring is a valid elimination order
        MUST(bsp_batch_create_ex(ctx, AB, nsys, srp.data(), nrhs, colour.data(), &bt));
        MUST(bsp_batch_pcg(ctx, bt, Bd, Xd, 1e-10, 1000, BSP_TOL_TRUE_RESIDUAL, its.data(), rel.data(), &ms));
        check("batch_pcg_coloured");
        MUST(bsp_batch_destroy(bt));
        MUST(bsp_free(ctx, Bd));
        MUST(bsp_free(ctx, Xd));
        MUST(bsp_bsr_destroy(AB));
    }
    {   // error returns
        bsp_bsr* E = nullptr;
        bsp_precond* M = nullptr;
        expect_rc("bsr_create_bs0", bsp_bsr_create(ctx, H.n, H.n, 0, H.rp.data(), H.ci.data(), H.v.data(), &E),
                  BSP_E_INVALID);
        expect_rc("precond_null_matrix", bsp_precond_create_ilu0(ctx, nullptr, &M), BSP_E_INVALID);
        Host S = H;
        // row 0 eliminates first, so its pivot is A's own diagonal block: a zero column makes
        // the ILU0 pivot singular (a later row's pivot would be changed by the elimination)
        for (int p = S.rp[0]; p < S.rp[1]; ++p)
            if (S.ci[p] == 0)
                for (int r = 0; r < S.bs; ++r) S.v[(size_t)p * 64 + r * 8 + 2] = 0.0;
        bsp_bsr* AS = upload(S);
        expect_rc("ilu0_singular_pivot", bsp_precond_create_ilu0(ctx, AS, &M), BSP_E_NUMERIC);
        if (M) MUST(bsp_precond_destroy(M));
        report("last_error_message", std::strlen(bsp_last_error(ctx)) > 0 ? 0.0 : 1.0, 0.0);
        M = nullptr;
        expect_rc("diag_singular_diagonal", bsp_precond_create_diag(ctx, AS, 0, &M), BSP_E_NUMERIC);
        if (M) MUST(bsp_precond_destroy(M));
        MUST(bsp_bsr_destroy(AS));
        MUST(bsp_precond_create_ilu0(ctx, A, &M));
        expect_rc("pgmres_not_converged", bsp_pgmres(ctx, A, M, b, x, 1e-14, 3, 30, 10, BSP_TOL_REFERENCE, &st),
                  BSP_NOT_CONVERGED);
        std::vector<double> bn = host(b, len);
        bn[7] = NAN;
        double* bnan = dev(bn);
        expect_rc("pgmres_nan_rhs", bsp_pgmres(ctx, A, M, bnan, x, 1e-10, 200, 30, 10, BSP_TOL_REFERENCE, &st),
                  BSP_E_NUMERIC);
        expect_rc("fgmres_restart_too_large",
                  bsp_fgmres_mixed(ctx, A, M, b, x, 1e-10, 200, 80, 10, BSP_PREC_F32, 0.0, &st), BSP_E_UNSUPPORTED);
        MUST(bsp_free(ctx, bnan));
        MUST(bsp_precond_destroy(M));
    }
    MUST(bsp_free(ctx, b));
    MUST(bsp_free(ctx, x));
    MUST(bsp_bsr_destroy(A));
    MUST(bsp_ctx_destroy(ctx));
    if (failures) {
        std::printf("CAPI FAILED (%d)\n", failures);
        return 1;
    }
    std::printf("CAPI PASSED\n");
    return 0;
}
