// bsp_krylov.cpp -- pGMRES with deflated restart and pCG on the device.
//
// Vectors of length n*bs live on the device for the whole solve; every
// vector operation (bsmv, preconditioner, modified Gram-Schmidt, updates,
// norms) is a kernel on the context stream.  Per GMRES iteration the host
// receives only the new Hessenberg column (iter+2 doubles) to run the
// least-squares update and the stopping test; per pCG iteration one scalar.
//
// pGMRES follows blocksp_solvers.hpp:156-295: left preconditioning, Arnoldi
// with modified Gram-Schmidt (one_arnoldi, 109-136), least squares on the full
// Hessenberg each iteration (the reference calls gels, 192-232), restart at
// iter == restart through deflated_restart (deflated_restart.hpp:38-269),
// final x = x0 + Q y.  pCG follows blocksp_solvers.hpp:51-105.
#include <chrono>
#include <cmath>
#include <deque>
#include <limits>

#include "bsp_dense.hpp"
#include "bsp_internal.hpp"

namespace bsp {

namespace {

struct VecPool {
    std::deque<DevMem> bufs;
    size_t len = 0;
    hipError_t grow(size_t count) {
        while (bufs.size() < count) {
            bufs.emplace_back();
            hipError_t e = bufs.back().alloc(sizeof(double) * std::max<size_t>(len, 1));
            if (e != hipSuccess) return e;
        }
        return hipSuccess;
    }
};

struct Timer {
    bsp_ctx* ctx;
    explicit Timer(bsp_ctx* c) : ctx(c) { (void)hipEventRecord(ctx->events[14], ctx->stream); }
    double stop() {
        float ms = 0.f;
        (void)hipEventRecord(ctx->events[15], ctx->stream);
        (void)hipEventSynchronize(ctx->events[15]);
        (void)hipEventElapsedTime(&ms, ctx->events[14], ctx->events[15]);
        return ms;
    }
};

int check_sweep_errors(bsp_ctx* ctx) {
    uint32_t err = 0;
    BSP_HIP(hipMemcpyAsync(&err, ctx->err_flag.p, sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
    BSP_HIP(hipStreamSynchronize(ctx->stream));
    if (err) {
        BSP_HIP(hipMemsetAsync(ctx->err_flag.p, 0, sizeof(uint32_t), ctx->stream));
        return fail(ctx, BSP_E_TIMEOUT, "a sync-free sweep wait exceeded the spin limit");
    }
    return BSP_OK;
}

// One pGMRES run in the reference's semantics, starting from x (x0 = x on entry).
// tolerance is absolute on the preconditioned least-squares residual estimate.
// The least squares stays on the device (bsp_restart.hip): each Arnoldi step's column is
// rotated into the device triangle by the step's last kernel, which also writes the residual
// estimate to host-visible memory.  The host keeps up to kAhead steps in flight and stops at
// the first step whose estimate meets the tolerance (or breaks down, or reaches maxIters):
// steps enqueued after it only add columns beyond the ones the solution uses.  At a restart
// the solution update, the residual vector and the deflation's eigenproblem run on the device;
// the host reads back kout once per restart.
int gmres_core(bsp_ctx* ctx, const bsp_bsr* A, const bsp_precond* M, const double* b, double* x, double tol,
               bool tol_is_relative, int maxIters, int restart, int k, int& iters_out, int& restarts_out,
               double& res0_out, double& res_out, bool& converged, int& applies, int stag_window,
               bool& stagnated) {
    constexpr int kAhead = 3;
    stagnated = false;
    std::vector<double> res_hist;
    const int n = A->n_rows, bs = A->bs;
    const int64_t len = (int64_t)n * bs;
    const double eps = std::numeric_limits<double>::epsilon();
    double* sc = ctx->scalars.as<double>();
    const int maxm = restart > 0 ? restart : maxIters + 1;
    DevMem hcol_mem;
    BSP_HIP(hcol_mem.alloc(sizeof(double) * (maxm + 2)));
    DeviceLS ls;
    if (int rc = ls_alloc(ctx, ls, maxm, restart > 0 && k > 0 ? restart : 0)) return rc;
    // per-column {residual estimate, |h_{j+1,j}|}, written by the device (coherent, mapped)
    struct Pinned {
        double* p = nullptr;
        ~Pinned() { if (p) (void)hipHostFree(p); }
    } outp;
    BSP_HIP(hipHostMalloc((void**)&outp.p, sizeof(double) * 2 * (maxm + 1), hipHostMallocCoherent | hipHostMallocMapped));
    double* out_dev = nullptr;
    BSP_HIP(hipHostGetDevicePointer((void**)&out_dev, outp.p, 0));
    hipEvent_t ev[kAhead + 1] = {};
    struct EvGuard {
        hipEvent_t* e;
        ~EvGuard() {
            for (int i = 0; i <= kAhead; ++i)
                if (e[i]) (void)hipEventDestroy(e[i]);
        }
    } evg{ev};
    for (hipEvent_t& e : ev) BSP_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    VecPool pool;
    pool.len = (size_t)len;
    BSP_HIP(pool.grow(2));
    std::vector<double*> Q;             // current basis (device pointers into pool)
    std::vector<double*> spare;
    auto take = [&](double*& p) -> int {
        if (!spare.empty()) {
            p = spare.back();
            spare.pop_back();
            return BSP_OK;
        }
        BSP_HIP(pool.grow(pool.bufs.size() + 1));
        p = pool.bufs.back().as<double>();
        return BSP_OK;
    };
    double* v = pool.bufs[0].as<double>();
    double* q0 = pool.bufs[1].as<double>();
    // r0 = M^{-1}(b - A x0)
    int rc = launch_residual(ctx, n, bs, A->rp.as<int>(), A->ci.as<int>(), A->val.as<double>(), b, x, q0);
    if (rc) return rc;
    rc = bsp_precond_apply(ctx, M, q0);
    ++applies;
    if (rc) return rc;
    double beta = 0.0;
    rc = reduce(ctx, 1, len, q0, q0, SLOT_TMP0, &beta);
    if (rc) return rc;
    res0_out = beta;
    iters_out = 0;
    restarts_out = 0;
    converged = false;
    if (!(beta > 0.0) || !std::isfinite(beta)) {
        res_out = beta;
        converged = beta == 0.0;
        return beta == 0.0 ? BSP_OK : fail(ctx, BSP_E_NUMERIC, "initial residual is not finite");
    }
    rc = launch_scale_div(ctx, len, q0, sc + SLOT_TMP0, q0);
    if (rc) return rc;
    Q.push_back(q0);
    // the cycle's beta stays in its own slot (SLOT_TMP0 is reused by the reductions)
    BSP_HIP(hipMemcpyAsync(sc + SLOT_BETA, sc + SLOT_TMP0, sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    if ((rc = ls_fresh(ctx, ls, sc + SLOT_BETA))) return rc;
    const double zero_thresh = std::min(beta * eps, eps);
    const double tolerance = tol_is_relative ? beta * tol : tol;
    double res = beta;
    int iter = 0, tot = 0;
    double* hcol_dev = hcol_mem.as<double>();
    // One Arnoldi step for column j (one_arnoldi, solvers.hpp:109-136), enqueued without waiting.
    auto enqueue = [&](int j) -> int {
        if (ctx->rec_buf && ctx->rec_len == len && ctx->rec_count < ctx->rec_cap &&
            ctx->rec_seen++ % ctx->rec_every == 0) {
            BSP_HIP(hipMemcpyAsync(ctx->rec_buf + (size_t)ctx->rec_count * len, Q[j], sizeof(double) * len,
                                   hipMemcpyDeviceToDevice, ctx->stream));
            ++ctx->rec_count;
        }
        int r2 = bsmv_matrix(ctx, A, Q[j], v, 0);
        if (r2) return r2;
        r2 = bsp_precond_apply(ctx, M, v);
        ++applies;
        if (r2) return r2;
        double* qn = nullptr;
        if ((r2 = take(qn))) return r2;
        r2 = launch_arnoldi_mgs(ctx, len, v, Q.data(), j, hcol_dev, qn, ctx->red_partials.as<double>());
        if (r2) return r2;
        Q.push_back(qn);   // speculative: dropped again if the run stops before this column
        if ((r2 = ls_column(ctx, ls, j, hcol_dev, out_dev + 2 * j, nullptr))) return r2;
        BSP_HIP(hipEventRecord(ev[j % (kAhead + 1)], ctx->stream));
        return BSP_OK;
    };
    bool done = false;
    int nc = 0;   // columns of the cycle the solution uses
    while (!done) {
        const int cycle_end = restart > 0 ? restart : maxm;
        int enq = iter;
        bool stop = false;
        while (iter < cycle_end) {
            // keep up to kAhead steps in flight, never past maxIters + 1 steps in total
            while (enq < cycle_end && enq - iter < kAhead && tot + (enq - iter) <= maxIters) {
                if ((rc = enqueue(enq))) return rc;
                ++enq;
            }
            if (enq == iter) break;   // nothing left to run (maxIters)
            BSP_HIP(hipEventSynchronize(ev[iter % (kAhead + 1)]));
            res = outp.p[2 * iter];
            const double hlast = outp.p[2 * iter + 1];
            const bool breakdown = hlast < zero_thresh || !std::isfinite(hlast);
            ++iter;
            ++tot;
            if (ctx->verbose > 0 && tot % ctx->verbose == 0) {
                std::fprintf(stderr, "[pgmres] iter %d res %.3e (target %.3e) restarts %d\n", tot, res, tolerance,
                             restarts_out);
                std::fflush(stderr);
            }
            if (res <= tolerance || tot > maxIters || breakdown || !std::isfinite(res)) {
                stop = true;
                if (breakdown) {
                    // one_arnoldi leaves the basis one vector short (solvers.hpp:133)
                }
                break;
            }
        }
        // drop the speculative basis vectors past the columns in use (steps still in flight
        // write them; the stream orders those writes before any later use of the buffers)
        while ((int)Q.size() > iter + 1) {
            spare.push_back(Q.back());
            Q.pop_back();
        }
        nc = iter;
        if (stop || iter < cycle_end) {
            done = true;
            break;
        }
        // ---- restart (iter == restart, not converged) ----
        if (stag_window > 0) {
            // stagnation of the restarted recurrence (true-residual mode only): less than a
            // factor 2 gained over the last stag_window restarts
            res_hist.push_back(res);
            const size_t h = res_hist.size();
            if (h > (size_t)stag_window && res > 0.5 * res_hist[h - 1 - stag_window]) {
                stagnated = true;
                done = true;
                break;
            }
        }
        // x0 += Q y (solvers.hpp:242-245); the reference's r0 recomputation (248) is only
        // consumed by a pure restart and is skipped for k > 0.
        if ((rc = ls_solve(ctx, ls, nc, sc + SLOT_BETA, k > 0))) return rc;
        if (k > 0 && (rc = ls_deflate_begin(ctx, ls, nc, k))) return rc;
        // the solution update runs on the device while the host solves the restart's eigenproblem
        rc = launch_combine_add(ctx, len, nc, Q.data(), ls.y.as<double>(), x);
        if (rc) return rc;
        ++restarts_out;
        if (k == 0) {
            // pure restart: r0 = M^{-1}(b - A x0), Q = [r0/||r0||]
            for (double* p : Q) spare.push_back(p);
            Q.clear();
            double* r0 = nullptr;
            if ((rc = take(r0))) return rc;
            rc = launch_residual(ctx, n, bs, A->rp.as<int>(), A->ci.as<int>(), A->val.as<double>(), b, x, r0);
            if (rc) return rc;
            rc = bsp_precond_apply(ctx, M, r0);
            ++applies;
            if (rc) return rc;
            rc = reduce(ctx, 1, len, r0, r0, SLOT_TMP0, &beta);
            if (rc) return rc;
            rc = launch_scale_div(ctx, len, r0, sc + SLOT_TMP0, r0);
            if (rc) return rc;
            Q.push_back(r0);
            BSP_HIP(hipMemcpyAsync(sc + SLOT_BETA, sc + SLOT_TMP0, sizeof(double), hipMemcpyDeviceToDevice,
                                   ctx->stream));
            if ((rc = ls_fresh(ctx, ls, sc + SLOT_BETA))) return rc;
            iter = 0;
            res = beta;
            continue;
        }
        int kout = 0;
        if ((rc = ls_deflate_end(ctx, ls, nc, k, false, kout))) return rc;
        const int kp = kout + 1;
        const int m = nc + 1;
        BSP_CHECK((int)Q.size() == m, BSP_E_NUMERIC, "basis size mismatch at restart");
        std::vector<double*> Qnew(kp);
        for (int j = 0; j < kp; ++j)
            if ((rc = take(Qnew[j]))) return rc;
        // Q_new = Q P (P: m x kp on the device)
        rc = launch_basis_rotate(ctx, len, m, kp, Q.data(), ls.P, Qnew.data());
        if (rc) return rc;
        if ((rc = ls_restart(ctx, ls, kout))) return rc;
        for (double* p : Q) spare.push_back(p);
        Q = Qnew;
        iter = kout;
    }
    // x = x0 + Q y (solvers.hpp:290-294)
    if (nc > 0 && std::isfinite(res)) {
        if ((rc = ls_solve(ctx, ls, nc, sc + SLOT_BETA, false))) return rc;
        rc = launch_combine_add(ctx, len, nc, Q.data(), ls.y.as<double>(), x);
        if (rc) return rc;
    }
    BSP_HIP(hipStreamSynchronize(ctx->stream));
    iters_out = tot;
    res_out = res;
    converged = res <= tolerance;
    std::vector<double> Hh;
    if ((rc = ls_download_h(ctx, ls, nc, Hh))) return rc;
    ctx->last_H = Hh;
    ctx->last_Hc.clear();
    ctx->last_Hm = nc + 1;
    ctx->last_Hn = nc;
    ctx->last_restarted = ls.restarted;
    return check_sweep_errors(ctx);
}

int true_relres(bsp_ctx* ctx, const bsp_bsr* A, const double* b, const double* x, double* work, double bnorm,
                double& out) {
    const int64_t len = (int64_t)A->n_rows * A->bs;
    int rc = launch_residual(ctx, A->n_rows, A->bs, A->rp.as<int>(), A->ci.as<int>(), A->val.as<double>(), b, x, work);
    if (rc) return rc;
    double rn = 0.0;
    rc = reduce(ctx, 1, len, work, work, SLOT_TMP1, &rn);
    if (rc) return rc;
    out = bnorm > 0.0 ? rn / bnorm : rn;
    return BSP_OK;
}

}  // namespace

}  // namespace bsp

using namespace bsp;

extern "C" {

int bsp_pgmres(bsp_ctx* ctx, const bsp_bsr* A, const bsp_precond* M, const double* b, double* x, double tol,
               int maxIters, int restart, int k, int tol_mode, bsp_stats* stats) {
    if (!ctx || !A || !M || !b || !x) return BSP_E_INVALID;
    BSP_CHECK(A->n_rows == A->n_cols, BSP_E_INVALID, "pGMRES needs a square matrix");
    BSP_CHECK(M->n == A->n_rows && M->bs == A->bs, BSP_E_INVALID, "preconditioner does not match the matrix");
    BSP_CHECK(b != x, BSP_E_INVALID, "b and x must not alias");
    BSP_HIP(hipSetDevice(ctx->device));
    const int64_t len = (int64_t)A->n_rows * A->bs;
    Timer timer(ctx);
    BSP_HIP(hipMemsetAsync(x, 0, sizeof(double) * len, ctx->stream));   // x0 = 0
    int iters = 0, restarts = 0, applies = 0;
    double res0 = 0.0, res = 0.0;
    bool conv = false, stagnated = false;
    // Reference semantics: one run with the reference's stopping rule.  True-residual mode
    // additionally leaves a run whose restarted recurrence stagnates (8 restarts without a
    // factor 2) and restarts from the current iterate with the exact residual.
    const bool true_mode = tol_mode == BSP_TOL_TRUE_RESIDUAL;
    const int stag_window = true_mode && restart > 0 ? 8 : 0;
    // complex systems: the same driver over the complex-arithmetic core
    auto core = A->cbs > 0 ? zgmres_core : gmres_core;
    int rc = core(ctx, A, M, b, x, tol, true, maxIters, restart, k, iters, restarts, res0, res, conv, applies,
                  stag_window, stagnated);
    if (rc < 0) return rc;
    double bnorm = 0.0, rr = 0.0;
    DevMem work;
    BSP_HIP(work.alloc(sizeof(double) * std::max<int64_t>(len, 1)));
    rc = reduce(ctx, 1, len, b, b, SLOT_TMP2, &bnorm);
    if (rc) return rc;
    rc = true_relres(ctx, A, b, x, work.as<double>(), bnorm, rr);
    if (rc) return rc;
    int total_iters = iters, total_restarts = restarts, passes = 1;
    if (true_mode) {
        // FP64 refinement: restart from x with the exact residual until ||b - Ax|| <= tol ||b||
        for (; passes < 50 && rr > tol && total_iters < maxIters && std::isfinite(rr); ++passes) {
            const double pass_tol = std::min(0.5, std::max(0.25 * tol / rr, 1e-12));
            int it2 = 0, rs2 = 0;
            double r02 = 0.0, res2 = 0.0;
            bool c2 = false;
            rc = core(ctx, A, M, b, x, pass_tol, true, maxIters - total_iters, restart, k, it2, rs2, r02, res2, c2,
                      applies, stag_window, stagnated);
            if (rc < 0) return rc;
            total_iters += it2;
            total_restarts += rs2;
            rc = true_relres(ctx, A, b, x, work.as<double>(), bnorm, rr);
            if (rc) return rc;
            res = res2;
            if (ctx->verbose > 0) {
                std::fprintf(stderr, "[pgmres] refinement pass %d: +%d iterations, true relres %.3e\n", passes, it2,
                             rr);
                std::fflush(stderr);
            }
        }
        conv = rr <= tol;
    }
    const double ms = timer.stop();
    if (stats) {
        stats->iters = total_iters;
        stats->restarts = total_restarts;
        stats->converged = conv ? 1 : 0;
        stats->precond_applies = applies;
        stats->res0 = res0;
        stats->res_est = res;
        stats->true_relres = rr;
        stats->solve_ms = ms;
        stats->fallback = 0;
        stats->reorth = 0;
    }
    return conv ? BSP_OK : BSP_NOT_CONVERGED;
}

int bsp_pcg(bsp_ctx* ctx, const bsp_bsr* A, const bsp_precond* M, const double* b, double* x, double tol,
            int maxIters, int tol_mode, bsp_stats* stats) {
    if (!ctx || !A || !M || !b || !x) return BSP_E_INVALID;
    BSP_CHECK(A->n_rows == A->n_cols, BSP_E_INVALID, "pCG needs a square matrix");
    // the reference's pCG solvers are real only (SOR_pCG, IChol0_pCG: blocksp_solvers.hpp:411, 485)
    BSP_CHECK(A->cbs == 0, BSP_E_UNSUPPORTED, "pCG: complex matrices are not supported");
    BSP_CHECK(M->n == A->n_rows && M->bs == A->bs, BSP_E_INVALID, "preconditioner does not match the matrix");
    BSP_CHECK(b != x, BSP_E_INVALID, "b and x must not alias");
    BSP_HIP(hipSetDevice(ctx->device));
    const int n = A->n_rows, bs = A->bs;
    const int64_t len = (int64_t)n * bs;
    Timer timer(ctx);
    DevMem mr, mp, mz, mAp;
    BSP_HIP(mr.alloc(sizeof(double) * std::max<int64_t>(len, 1)));
    BSP_HIP(mp.alloc(sizeof(double) * std::max<int64_t>(len, 1)));
    BSP_HIP(mz.alloc(sizeof(double) * std::max<int64_t>(len, 1)));
    BSP_HIP(mAp.alloc(sizeof(double) * std::max<int64_t>(len, 1)));
    double *r = mr.as<double>(), *p = mp.as<double>(), *z = mz.as<double>(), *Ap = mAp.as<double>();
    double* sc = ctx->scalars.as<double>();
    const int* rp = A->rp.as<int>();
    const int* ci = A->ci.as<int>();
    const double* av = A->val.as<double>();
    BSP_HIP(hipMemsetAsync(x, 0, sizeof(double) * len, ctx->stream));
    double bnorm = 0.0;
    int rc = reduce(ctx, 1, len, b, b, SLOT_TMP2, &bnorm);
    if (rc) return rc;
    int total = 0, applies = 0;
    double res0 = 0.0, res = 0.0, rr = 0.0;
    bool conv = false;
    const double target = tol;
    for (int outer = 0; outer < (tol_mode == BSP_TOL_TRUE_RESIDUAL ? 20 : 1); ++outer) {
        // r = b - A x (x = 0 on the first pass: r = b, solvers.hpp:60)
        rc = launch_residual(ctx, n, bs, rp, ci, av, b, x, r);
        if (rc) return rc;
        rc = launch_copy(ctx, len, r, p);
        if (rc) return rc;
        rc = bsp_precond_apply(ctx, M, p);
        ++applies;
        if (rc) return rc;
        rc = launch_copy(ctx, len, p, z);
        if (rc) return rc;
        double pn = 0.0;
        rc = reduce(ctx, 1, len, p, p, SLOT_TMP0, &pn);
        if (rc) return rc;
        if (outer == 0) {
            res0 = pn;
            // as pGMRES: a right-hand side (or preconditioner) that gives a non-finite initial
            // residual is an error, not a run that stops at once
            if (!std::isfinite(pn)) return fail(ctx, BSP_E_NUMERIC, "initial residual is not finite");
        }
        double inner_tol = tol;
        if (outer > 0) inner_tol = std::max(0.25 * target / rr, 1e-15);
        const double tolerance = pn * inner_tol;
        res = pn;
        int iter = 0;
        // rz in SLOT 0.. : s0 = rz, s1 = pAp, s3 = rz_new
        while (res > tolerance) {
            rc = bsmv_matrix(ctx, A, p, Ap, 0);
            if (rc) return rc;
            rc = reduce(ctx, 0, len, r, z, SLOT_TMP0, nullptr);
            if (rc) return rc;
            rc = reduce(ctx, 0, len, p, Ap, SLOT_TMP1, nullptr);
            if (rc) return rc;
            rc = launch_cg_update_xr(ctx, len, sc + SLOT_TMP0, sc + SLOT_TMP1, p, Ap, x, r);
            if (rc) return rc;
            rc = reduce(ctx, 1, len, r, r, SLOT_TMP2 + 1, &res);
            if (rc) return rc;
            if (!std::isfinite(res)) break;
            if (res > tolerance) {
                rc = launch_copy(ctx, len, r, z);
                if (rc) return rc;
                rc = bsp_precond_apply(ctx, M, z);
                ++applies;
                if (rc) return rc;
                rc = reduce(ctx, 0, len, r, z, SLOT_TMP3, nullptr);
                if (rc) return rc;
                rc = launch_cg_update_p(ctx, len, sc + SLOT_TMP3, sc + SLOT_TMP0, z, p);
                if (rc) return rc;
                ++iter;
                if (ctx->verbose > 0 && iter % ctx->verbose == 0) {
                    std::fprintf(stderr, "[pcg] iter %d res %.3e (target %.3e)\n", total + iter, res, tolerance);
                    std::fflush(stderr);
                }
                if (total + iter > maxIters) break;
            }
        }
        total += iter;
        rc = true_relres(ctx, A, b, x, Ap, bnorm, rr);
        if (rc) return rc;
        conv = tol_mode == BSP_TOL_TRUE_RESIDUAL ? rr <= target : res <= tolerance;
        if (tol_mode != BSP_TOL_TRUE_RESIDUAL || conv || total > maxIters || !std::isfinite(rr)) break;
    }
    rc = check_sweep_errors(ctx);
    if (rc) return rc;
    const double ms = timer.stop();
    if (stats) {
        stats->iters = total;
        stats->restarts = 0;
        stats->converged = conv ? 1 : 0;
        stats->precond_applies = applies;
        stats->res0 = res0;
        stats->res_est = res;
        stats->true_relres = rr;
        stats->solve_ms = ms;
        stats->fallback = 0;
        stats->reorth = 0;
    }
    return conv ? BSP_OK : BSP_NOT_CONVERGED;
}

}  // extern "C"

namespace bsp {

int true_relres_ext(bsp_ctx* ctx, const bsp_bsr* A, const double* b, const double* x, double* work, double bnorm,
                    double& out) {
    return true_relres(ctx, A, b, x, work, bnorm, out);
}

int check_sweep_errors_ext(bsp_ctx* ctx) { return check_sweep_errors(ctx); }

}  // namespace bsp
