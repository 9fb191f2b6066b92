// bsp_eig.cpp -- Ritz-value diagnostics: the printEigInfo of pGMRES
// (blocksp_solvers.hpp:272-288) and arnoldi_eigenvalues (blocksp_solvers.hpp:298-333).
// The Arnoldi process runs on the device (bsmv, preconditioner, FP64 modified
// Gram-Schmidt); the small Hessenberg eigenproblem is solved on the host by the same
// complex Schur routine the deflated restart uses.
#include <cmath>
#include <limits>

#include "bsp_dense.hpp"
#include "bsp_zdense.hpp"
#include "bsp_internal.hpp"

using namespace bsp;

namespace {

int write_values(const std::vector<dense::cplx>& ev, double* re, double* im, int cap, int* count) {
    const int n = (int)ev.size();
    if (count) *count = n;
    for (int i = 0; i < n && i < cap; ++i) {
        if (re) re[i] = ev[i].real();
        if (im) im[i] = ev[i].imag();
    }
    return BSP_OK;
}

}  // namespace

extern "C" {

int bsp_pgmres_ritz(bsp_ctx* ctx, double* re, double* im, int cap, int* count, int* harmonic) {
    if (!ctx || cap < 0) return BSP_E_INVALID;
    if (count) *count = 0;
    if (harmonic) *harmonic = ctx->last_restarted ? 1 : 0;
    const int m = ctx->last_Hm, n = ctx->last_Hn;
    std::vector<dense::cplx> ev;
    if (!ctx->last_Hc.empty()) {   // the last run solved a complex system
        BSP_CHECK(n > 0 && m >= n && (int64_t)ctx->last_Hc.size() >= (int64_t)m * n, BSP_E_INVALID,
                  "no pGMRES Hessenberg recorded on this context");
        BSP_CHECK(dense::zritz_values(m, n, ctx->last_Hc, ctx->last_restarted, ev), BSP_E_NUMERIC,
                  "Hessenberg eigenproblem failed");
        return write_values(ev, re, im, cap, count);
    }
    BSP_CHECK(n > 0 && m >= n && (int64_t)ctx->last_H.size() >= (int64_t)m * n, BSP_E_INVALID,
              "no pGMRES Hessenberg recorded on this context");
    // the square part of H; harmonic Ritz values when the run restarted
    BSP_CHECK(dense::ritz_values(m, n, ctx->last_H.data(), ctx->last_restarted, ev), BSP_E_NUMERIC,
              "Hessenberg eigenproblem failed");
    return write_values(ev, re, im, cap, count);
}

int bsp_arnoldi_eigenvalues(bsp_ctx* ctx, const bsp_bsr* A, const bsp_precond* M, int iters, int harmonic,
                            double* re, double* im, int cap, int* count) {
    if (!ctx || !A || iters < 1 || cap < 0) return BSP_E_INVALID;
    BSP_CHECK(A->n_rows == A->n_cols, BSP_E_INVALID, "arnoldi_eigenvalues needs a square matrix");
    BSP_CHECK(A->cbs == 0, BSP_E_UNSUPPORTED, "arnoldi_eigenvalues: complex matrices are not supported");
    BSP_CHECK(!M || (M->n == A->n_rows && M->bs == A->bs), BSP_E_INVALID, "preconditioner does not match");
    BSP_HIP(hipSetDevice(ctx->device));
    const int n = A->n_rows, bs = A->bs;
    const int64_t len = (int64_t)n * bs;
    iters = (int)std::min<int64_t>(iters, len);
    // zero threshold 10 M^2 eps (blocksp_solvers.hpp:319)
    const double zero_thresh = 10.0 * bs * bs * std::numeric_limits<double>::epsilon();
    std::vector<DevMem> Qm(iters + 1);
    std::vector<double*> Q;
    for (auto& q : Qm) BSP_HIP(q.alloc(sizeof(double) * len));
    DevMem v, hcol;
    BSP_HIP(v.alloc(sizeof(double) * len));
    BSP_HIP(hcol.alloc(sizeof(double) * (iters + 2)));
    // q_0 = e_0 (solvers.hpp:315-316)
    BSP_HIP(hipMemsetAsync(Qm[0].p, 0, sizeof(double) * len, ctx->stream));
    const double one = 1.0;
    BSP_HIP(hipMemcpyAsync(Qm[0].p, &one, sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    Q.push_back(Qm[0].as<double>());
    std::vector<double> H((size_t)(iters + 1) * iters, 0.0), col(iters + 2);
    int done = 0;
    for (int it = 0; it < iters; ++it) {
        int rc = bsmv_matrix(ctx, A, Q[it], v.as<double>(), 0);
        if (!rc && M) rc = bsp_precond_apply(ctx, M, v.as<double>());
        if (!rc)
            rc = launch_arnoldi_mgs(ctx, len, v.as<double>(), Q.data(), it, hcol.as<double>(), Qm[it + 1].as<double>(),
                                    ctx->red_partials.as<double>());
        if (rc) return rc;
        BSP_HIP(hipMemcpyAsync(col.data(), hcol.p, sizeof(double) * (it + 2), hipMemcpyDeviceToHost, ctx->stream));
        BSP_HIP(hipStreamSynchronize(ctx->stream));
        for (int j = 0; j < it + 2; ++j) H[(size_t)j * iters + it] = col[j];
        done = it + 1;
        if (std::fabs(col[it + 1]) < zero_thresh || !std::isfinite(col[it + 1])) break;   // invariant subspace
        Q.push_back(Qm[it + 1].as<double>());
    }
    // compact to (done + 1) x done
    std::vector<double> Hc((size_t)(done + 1) * done);
    for (int i = 0; i <= done; ++i)
        for (int j = 0; j < done; ++j) Hc[(size_t)i * done + j] = H[(size_t)i * iters + j];
    std::vector<dense::cplx> ev;
    BSP_CHECK(dense::ritz_values(done + 1, done, Hc.data(), harmonic != 0, ev), BSP_E_NUMERIC,
              "Hessenberg eigenproblem failed");
    return write_values(ev, re, im, cap, count);
}

}  // extern "C"
