// bsp_zgmres.cpp -- pGMRES for complex systems (BlockSp<std::complex<double>, B>): the
// reference's left-preconditioned GMRES with deflated restart (blocksp_solvers.hpp:109-295,
// blocksp_deflated_restart.hpp:38-269) with T complex.
//
// The matrix is the real-equivalent form made by bsp_bsr_create_complex (each complex entry
// a + ib as the 2 x 2 block [[a, -b], [b, a]]), so the product and every preconditioner
// apply are the real kernels on the interleaved complex vectors.  The Arnoldi step needs the
// complex inner product, which the real one is not: it runs classical Gram-Schmidt twice
// with the complex kernels of bsp_complex.hip (the coefficients of the two passes are added,
// the "twice is enough" form of the reference's modified Gram-Schmidt, one_arnoldi,
// solvers.hpp:127-131).  The Hessenberg matrix, its least squares (solved afresh every step,
// as the reference's zgels is) and the harmonic-Ritz restart live on the host in complex
// arithmetic (bsp_zdense.hpp); the basis rotation and the solution updates are device kernels.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <vector>

#include "bsp_internal.hpp"
#include "bsp_zdense.hpp"

namespace bsp {

using cplx = std::complex<double>;

int zdots_blocks(int64_t nc);
int launch_zdots(const bsp_ctx* ctx, int64_t nc, int m, const double* Q, int64_t stride, const double* v,
                 double* part);
int launch_zupdate(const bsp_ctx* ctx, int64_t nc, int m, const double* Q, int64_t stride, const double* h,
                   double sign, double* v);
int launch_zrotate(const bsp_ctx* ctx, int64_t nc, int m, int kp, const double* Q, int64_t stride, const double* P,
                   double* Qout, int64_t ostride);

namespace {

// out[j] = (Q_j, v) for j < m: device partials added in workgroup order on the host
int zdots(bsp_ctx* ctx, int64_t nc, int m, const double* Q, int64_t stride, const double* v, DevMem& part,
          std::vector<double>& hpart, std::vector<cplx>& out) {
    const int nblk = zdots_blocks(nc);
    const size_t cnt = (size_t)m * nblk * 2;
    if (part.bytes < sizeof(double) * cnt) BSP_HIP(part.alloc(sizeof(double) * cnt));
    int rc = launch_zdots(ctx, nc, m, Q, stride, v, part.as<double>());
    if (rc) return rc;
    hpart.resize(cnt);
    BSP_HIP(hipMemcpyAsync(hpart.data(), part.p, sizeof(double) * cnt, hipMemcpyDeviceToHost, ctx->stream));
    BSP_HIP(hipStreamSynchronize(ctx->stream));
    out.assign(m, cplx(0.0));
    for (int j = 0; j < m; ++j) {
        double re = 0.0, im = 0.0;
        for (int b = 0; b < nblk; ++b) {
            re += hpart[((size_t)j * nblk + b) * 2];
            im += hpart[((size_t)j * nblk + b) * 2 + 1];
        }
        out[j] = cplx(re, im);
    }
    return BSP_OK;
}

int upload(bsp_ctx* ctx, DevMem& dst, const std::vector<cplx>& src) {
    const size_t bytes = sizeof(cplx) * std::max<size_t>(src.size(), 1);
    if (dst.bytes < bytes) BSP_HIP(dst.alloc(bytes));
    if (!src.empty())
        BSP_HIP(hipMemcpyAsync(dst.p, src.data(), sizeof(cplx) * src.size(), hipMemcpyHostToDevice, ctx->stream));
    // the host vector may change before the copy runs: wait for it
    BSP_HIP(hipStreamSynchronize(ctx->stream));
    return BSP_OK;
}

}  // namespace

// Same contract as gmres_core (bsp_krylov.cpp) for a complex matrix (A->cbs > 0): solves from
// the initial guess in x (r_0 = M^{-1}(b - A x)), updates x in place.
int zgmres_core(bsp_ctx* ctx, const bsp_bsr* A, const bsp_precond* M, const double* b, double* x, double tol,
                bool tol_is_relative, int maxIters, int restart, int k, int& iters_out, int& restarts_out,
                double& res0_out, double& res_out, bool& converged, int& applies, int stag_window,
                bool& stagnated) {
    stagnated = false;
    const int n = A->n_rows, bs = A->bs;
    const int64_t len = (int64_t)n * bs, nc = len / 2;
    const double eps = std::numeric_limits<double>::epsilon();
    const int maxm = restart > 0 ? restart : maxIters + 1;
    const bool deflate = restart > 0 && k > 0;
    DevMem Qa, Qb, vbuf, part, hbuf;
    BSP_HIP(Qa.alloc(sizeof(double) * (size_t)(maxm + 1) * std::max<int64_t>(len, 1)));
    if (deflate) BSP_HIP(Qb.alloc(sizeof(double) * (size_t)(std::min(k, maxm) + 1) * std::max<int64_t>(len, 1)));
    BSP_HIP(vbuf.alloc(sizeof(double) * std::max<int64_t>(len, 1)));
    double* Q = Qa.as<double>();
    double* v = vbuf.as<double>();
    auto q = [&](int j) { return Q + (size_t)j * len; };
    std::vector<double> hpart;
    // r0 = M^{-1}(b - A x0)
    int rc = launch_residual(ctx, n, bs, A->rp.as<int>(), A->ci.as<int>(), A->val.as<double>(), b, x, q(0));
    if (rc) return rc;
    if ((rc = bsp_precond_apply(ctx, M, q(0)))) return rc;
    ++applies;
    double beta = 0.0;
    if ((rc = reduce(ctx, 1, len, q(0), q(0), SLOT_TMP0, &beta))) return rc;
    res0_out = beta;
    iters_out = 0;
    restarts_out = 0;
    converged = false;
    if (!(beta > 0.0) || !std::isfinite(beta)) {
        res_out = beta;
        converged = beta == 0.0;
        return beta == 0.0 ? BSP_OK : fail(ctx, BSP_E_NUMERIC, "initial residual is not finite");
    }
    if ((rc = launch_scal(ctx, len, 1.0 / beta, q(0)))) return rc;
    const double zero_thresh = std::min(beta * eps, eps);
    const double tolerance = tol_is_relative ? beta * tol : tol;
    double res = beta, cycle_beta = beta;
    std::vector<cplx> H;   // (it + 2) x (it + 1), row-major, rebuilt as it grows
    int Hm = 1, Hn = 0;
    std::vector<cplx> c_r, y, res_vec, hcol, hpass;
    bool restarted = false;
    int it = 0, tot = 0;
    std::vector<double> res_at_restart;
    while (res > tolerance) {
        // one_arnoldi (solvers.hpp:109-136) on column it
        const int m = it + 2, nn = it + 1;
        {
            std::vector<cplx> Hg((size_t)m * nn, cplx(0.0));
            for (int i = 0; i < std::min(Hm, m); ++i)
                for (int j = 0; j < std::min(Hn, nn); ++j) Hg[(size_t)i * nn + j] = H[(size_t)i * Hn + j];
            H.swap(Hg);
            Hm = m;
            Hn = nn;
        }
        if ((rc = bsmv_matrix(ctx, A, q(it), v, 0))) return rc;
        if ((rc = bsp_precond_apply(ctx, M, v))) return rc;
        ++applies;
        hcol.assign(it + 1, cplx(0.0));
        for (int pass = 0; pass < 2; ++pass) {
            if ((rc = zdots(ctx, nc, it + 1, Q, len, v, part, hpart, hpass))) return rc;
            if ((rc = upload(ctx, hbuf, hpass))) return rc;
            if ((rc = launch_zupdate(ctx, nc, it + 1, Q, len, hbuf.as<double>(), -1.0, v))) return rc;
            for (int j = 0; j <= it; ++j) hcol[j] += hpass[j];
        }
        double hnorm = 0.0;
        if ((rc = reduce(ctx, 1, len, v, v, SLOT_TMP1, &hnorm))) return rc;
        for (int j = 0; j <= it; ++j) H[(size_t)j * nn + it] = hcol[j];
        H[(size_t)(it + 1) * nn + it] = hnorm;
        const bool breakdown = std::fabs(hnorm) < zero_thresh || !std::isfinite(hnorm);
        if (!breakdown) {
            BSP_HIP(hipMemcpyAsync(q(it + 1), v, sizeof(double) * len, hipMemcpyDeviceToDevice, ctx->stream));
            if ((rc = launch_scal(ctx, len, 1.0 / hnorm, q(it + 1)))) return rc;
        }
        // least squares min || c - H y || (solvers.hpp:192-232)
        std::vector<cplx> c(m, cplx(0.0));
        if (!restarted)
            c[0] = cycle_beta;
        else
            for (int i = 0; i < (int)c_r.size() && i < m; ++i) c[i] = c_r[i];
        if (!dense::zlstsq(m, nn, H.data(), c.data(), y))
            return fail(ctx, BSP_E_NUMERIC, "complex pGMRES: singular least-squares problem");
        res_vec.assign(m, cplx(0.0));
        double r2 = 0.0;
        for (int i = 0; i < m; ++i) {
            cplx s = c[i];
            for (int j = 0; j < nn; ++j) s -= H[(size_t)i * nn + j] * y[j];
            res_vec[i] = s;
            r2 += std::norm(s);
        }
        res = std::sqrt(r2);
        ++it;
        ++tot;
        if (ctx->verbose > 1) {
            std::fprintf(stderr, "[zgmres] it %d res %.6e\n", tot, res);
            std::fflush(stderr);
        }
        if (tot > maxIters || breakdown || !std::isfinite(res)) break;
        if (it == restart && res > tolerance) {
            // x_0 += Q y, then the deflated restart (solvers.hpp:238-263)
            if ((rc = upload(ctx, hbuf, y))) return rc;
            if ((rc = launch_zupdate(ctx, nc, (int)y.size(), Q, len, hbuf.as<double>(), 1.0, x))) return rc;
            ++restarts_out;
            res_at_restart.push_back(res);
            if (stag_window > 0 && (int)res_at_restart.size() > stag_window &&
                res > 0.5 * res_at_restart[res_at_restart.size() - 1 - stag_window]) {
                stagnated = true;
                y.clear();
                break;
            }
            if (!deflate) {
                // pure restart: r_0 = M^{-1}(b - A x_0)
                if ((rc = launch_residual(ctx, n, bs, A->rp.as<int>(), A->ci.as<int>(), A->val.as<double>(), b, x,
                                          q(0))))
                    return rc;
                if ((rc = bsp_precond_apply(ctx, M, q(0)))) return rc;
                ++applies;
                double nb = 0.0;
                if ((rc = reduce(ctx, 1, len, q(0), q(0), SLOT_TMP0, &nb))) return rc;
                res = nb;
                if (!(nb > 0.0) || !std::isfinite(nb)) {
                    y.clear();
                    break;
                }
                if ((rc = launch_scal(ctx, len, 1.0 / nb, q(0)))) return rc;
                cycle_beta = nb;
                H.clear();
                Hm = 1;
                Hn = 0;
                it = 0;
                restarted = false;
            } else {
                std::vector<cplx> P, Hnew;
                const int kk = dense::zdeflate(it, H, res_vec, k, P, Hnew, c_r);
                if (kk < 0) return fail(ctx, BSP_E_NUMERIC, "complex pGMRES: deflated restart failed");
                if ((rc = upload(ctx, hbuf, P))) return rc;
                if ((rc = launch_zrotate(ctx, nc, it + 1, kk + 1, Q, len, hbuf.as<double>(), Qb.as<double>(), len)))
                    return rc;
                BSP_HIP(hipMemcpyAsync(Q, Qb.p, sizeof(double) * (size_t)(kk + 1) * len, hipMemcpyDeviceToDevice,
                                       ctx->stream));
                H.swap(Hnew);
                Hm = kk + 1;
                Hn = kk;
                it = kk;
                restarted = true;
            }
            y.clear();
        }
    }
    if (!y.empty()) {
        if ((rc = upload(ctx, hbuf, y))) return rc;
        if ((rc = launch_zupdate(ctx, nc, (int)y.size(), Q, len, hbuf.as<double>(), 1.0, x))) return rc;
    }
    // the Hessenberg matrix of the last cycle, for bsp_pgmres_ritz
    ctx->last_Hc = H;
    ctx->last_Hm = Hm;
    ctx->last_Hn = Hn;
    ctx->last_restarted = restarted;
    ctx->last_H.clear();
    iters_out = tot;
    res_out = res;
    converged = res <= tolerance;
    BSP_HIP(hipStreamSynchronize(ctx->stream));
    return BSP_OK;
}

}  // namespace bsp
