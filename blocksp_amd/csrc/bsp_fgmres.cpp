// bsp_fgmres.cpp -- mixed-precision flexible GMRES with deflated restart and FP64
// iterative refinement.
//
// Outer loop (FP64): r = b - A x with the FP64 matrix; stop when ||r|| <= tol ||b||.
// Inner solve of A e = r: right-preconditioned flexible GMRES (Saad, "A flexible
// inner-outer preconditioned GMRES", 1993) with Morgan's deflated restart:
//   z_j = P(v_j)            preconditioner in FP32 (or FP64 for an FP64 M), stored FP32
//   w   = A z_j             SpMV with FP32 values and FP32 accumulation (or FP64)
//   Gram-Schmidt against V (FP64 arithmetic; V stored in FP64, or in FP32 with the
//   fgmres_basis_f32 option), H column, v_{j+1} = w / h_{j+1,j}
//   least squares on H (FP64, host), x += Z y with FP64 accumulation
//   restart: harmonic Ritz subspace of H (FP64) -> V' = V P_{k+1}, Z' = Z P_k,
//            H' = P_{k+1}^T H P_k (A Z' = V' H' is preserved).
#include <cmath>
#include <cstdio>
#include <deque>
#include <limits>
#include <type_traits>

#include "bsp_dense.hpp"
#include "bsp_internal.hpp"

namespace bsp {

// FP32 copy of A's values for the inner SpMV: in the ELL order of A's ELL structure (chunk-major
// blocks) when the block size allows chunks and the context's bsmv_layout is 2, else in the
// chunk-major order of the chunked kernel when the block size allows it.
int ensure_values_f32(bsp_ctx* ctx, const bsp_bsr* A) {
    if (A->val32.p) return BSP_OK;
    if (ctx->bsmv_layout == 2 && bsmv_f32_chunked_supported(A->bs) && ell_width(A) > 0) {
        int rc = ensure_ell_columns(ctx, A);
        if (rc) return rc;
        BSP_HIP(A->val32.alloc(sizeof(float) * A->bs * A->bs * std::max<int64_t>((int64_t)A->n_rows * A->ell_w, 1)));
        A->val32_ell = true;
        A->val32_chunked = false;
        return launch_ell_values_f32(ctx, A->n_rows, A->bs, A->ell_w, A->rp.as<int>(), A->val.as<double>(),
                                     A->val32.as<float>());
    }
    const int64_t cnt = A->nnzb * (int64_t)A->bs * A->bs;
    BSP_HIP(A->val32.alloc(sizeof(float) * std::max<int64_t>(cnt, 1)));
    A->val32_ell = false;
    A->val32_chunked = bsmv_f32_chunked_supported(A->bs);
    if (A->val32_chunked) return launch_d2f_chunked(ctx, A->nnzb, A->bs, A->val.as<double>(), A->val32.as<float>());
    return launch_d2f(ctx, cnt, A->val.as<double>(), A->val32.as<float>());
}

namespace {

template <class T>
struct Pool {
    std::deque<DevMem> bufs;
    std::vector<T*> free_;
    size_t len = 0;
    int take(bsp_ctx* ctx, T*& p) {
        if (!free_.empty()) {
            p = free_.back();
            free_.pop_back();
            return BSP_OK;
        }
        bufs.emplace_back();
        BSP_HIP(bufs.back().alloc(sizeof(T) * std::max<size_t>(len, 1)));
        p = bufs.back().template as<T>();
        return BSP_OK;
    }
    void give(T* p) { free_.push_back(p); }
};

struct Pinned {
    double* p = nullptr;
    ~Pinned() {
        if (p) (void)hipHostFree(p);
    }
};

struct InnerStats {
    int iters = 0, restarts = 0;
    double res = 0.0, beta = 0.0;
    bool stagnated = false;
    bool nonfinite = false;   // the pass met a non-finite value; its last update was dropped
    int reorth = 0;           // steps that needed a second Gram-Schmidt pass
};

// One inner FGMRES-DR pass on A e = r for the current x (x += e on return).  VT is the
// storage type of the Arnoldi basis V (FP64, or FP32 with FP64 Gram-Schmidt arithmetic); ZT
// that of the preconditioned basis Z (FP32, or FP64 when the preconditioner and the SpMV are
// both FP64: the all-FP64 flexible GMRES).
template <typename VT, typename ZT>
int fgmres_pass(bsp_ctx* ctx, const bsp_bsr* A, const bsp_precond* M, const double* b, double* x, double reduction,
                int maxIters, int restart, int k, bool spmv_f32, InnerStats& st, double* r_work, double* t64) {
    const int n = A->n_rows, bs = A->bs;
    const int64_t len = (int64_t)n * bs;
    const double eps = std::numeric_limits<double>::epsilon();
    const int m_max = restart > 0 ? restart : maxIters + 1;
    constexpr bool vf32 = std::is_same_v<VT, float>;
    constexpr bool z64 = std::is_same_v<ZT, double>;
    if constexpr (z64) {
        BSP_CHECK(!M->f32 && !spmv_f32, BSP_E_INVALID, "an FP64 Z basis needs an FP64 preconditioner and SpMV");
    }
    Pool<VT> V;
    Pool<ZT> Z;
    V.len = Z.len = (size_t)len;
    DevMem hcol;
    BSP_HIP(hcol.alloc(sizeof(double) * 2 * (m_max + 2)));
    // least squares and deflated restart on the device
    DeviceLS ls;
    if (int rc = ls_alloc(ctx, ls, m_max, restart > 0 ? restart : 0)) return rc;
    Pinned hpin, outp;
    // H column slots (for the reorthogonalisation flag); the classical Gram-Schmidt kernels write
    // them directly (coherent, mapped, through the device address of the same memory)
    BSP_HIP(hipHostMalloc((void**)&hpin.p, sizeof(double) * 2 * (m_max + 2),
                          hipHostMallocCoherent | hipHostMallocMapped));
    double* hpin_dev = nullptr;
    BSP_HIP(hipHostGetDevicePointer((void**)&hpin_dev, hpin.p, 0));
    // per-column {residual estimate, |h_{j+1,j}|} from the device least squares
    BSP_HIP(hipHostMalloc((void**)&outp.p, sizeof(double) * 2 * (m_max + 1), hipHostMallocCoherent | hipHostMallocMapped));
    double* out_dev = nullptr;
    BSP_HIP(hipHostGetDevicePointer((void**)&out_dev, outp.p, 0));
    // classical Gram-Schmidt (default): scratch and the reorthogonalisation flag
    const bool use_cgs = vf32 || ctx->fgmres_ortho == 0;
    DevMem cgs_scratch, cgs_flag;
    if (use_cgs) {
        BSP_HIP(cgs_scratch.alloc(sizeof(double) * cgs_scratch_doubles()));
        BSP_HIP(cgs_flag.alloc(sizeof(int)));
    }
    double* sc = ctx->scalars.as<double>();
    // r = b - A x (FP64), beta = ||r||, V[0] = r / beta
    int rc = launch_residual(ctx, n, bs, A->rp.as<int>(), A->ci.as<int>(), A->val.as<double>(), b, x, r_work);
    if (rc) return rc;
    double beta = 0.0;
    rc = reduce(ctx, 1, len, r_work, r_work, SLOT_TMP0, &beta);
    if (rc) return rc;
    st = InnerStats{};
    st.beta = beta;
    st.res = beta;
    if (!(beta > 0.0) || !std::isfinite(beta)) return std::isfinite(beta) ? BSP_OK : fail(ctx, BSP_E_NUMERIC, "residual not finite");
    std::vector<VT*> Vb;
    std::vector<ZT*> Zb;
    VT* v0 = nullptr;
    rc = V.take(ctx, v0);
    if (rc) return rc;
    rc = launch_scale_div(ctx, len, r_work, sc + SLOT_TMP0, v0);
    if (rc) return rc;
    Vb.push_back(v0);
    BSP_HIP(hipMemcpyAsync(sc + SLOT_BETA, sc + SLOT_TMP0, sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
    if ((rc = ls_fresh(ctx, ls, sc + SLOT_BETA))) return rc;
    const double zero_thresh = std::min(beta * eps, eps);
    const double target = beta * reduction;
    std::vector<double> res_hist;
    const bool f32_pre = M->f32;
    const size_t hslot = (size_t)m_max + 2;   // H column slots: step j uses slot j % 2
    // e[slot]: the step's Gram-Schmidt done (main stream); e[2 + slot]: its least-squares update
    // done (second stream)
    struct Events {
        hipEvent_t e[4] = {nullptr, nullptr, nullptr, nullptr};
        ~Events() {
            for (hipEvent_t x : e)
                if (x) (void)hipEventDestroy(x);
        }
    } ev;
    // e[0..1] only order the second stream after the main one on this device: no system-scope
    // fence (host-visible results are written to pinned memory by the kernels themselves, and
    // e[2..3], which the host waits on, keep the fence); measured 6.6 -> 5.1 us between steps
    for (int i = 0; i < 4; ++i)
        BSP_HIP(hipEventCreateWithFlags(&ev.e[i], hipEventDisableTiming | (i < 2 ? hipEventDisableSystemFence : 0)));
    // the least-squares updates run on a second stream: the next Arnoldi step does not need them,
    // only the host (the residual estimate) and the next update do
    if (!ctx->aux_stream) BSP_HIP(hipStreamCreateWithFlags(&ctx->aux_stream, hipStreamNonBlocking));
    hipStream_t ls_stream = ctx->aux_stream;
    struct SyncAux {
        hipStream_t s;
        ~SyncAux() { (void)hipStreamSynchronize(s); }
    } sync_aux{ls_stream};
    double res = beta;
    int iter = 0, tot = 0;
    // w of the step in slot 0 reuses r_work (r is not needed after V[0] is formed); slot 1 has
    // its own, so a step's w survives while the next step runs
    DevMem w1;
    if (use_cgs) BSP_HIP(w1.alloc(sizeof(double) * std::max<int64_t>(len, 1)));
    double* wbuf[2] = {r_work, use_cgs ? w1.as<double>() : r_work};
    // One Arnoldi step for column `col`, enqueued without waiting: z = P(V[col]), w = A z,
    // Gram-Schmidt of w against V[0..col], q = w / h -> V[col + 1] (appended to Vb at once), event
    // ev.e[slot]; then, on the second stream after that event, the column rotated into the device
    // least squares (which writes the residual estimate for the host), event ev.e[2 + slot].
    // Classical Gram-Schmidt runs its first pass only; the flag it writes with the H column to
    // host slot `slot` says whether the rare second pass is needed -- then the least-squares update
    // of this column skips itself and the host applies the corrected column.
    // With an FP32 preconditioner the Gram-Schmidt step also writes the FP32 copy of V[col + 1]
    // (znext), which the next step preconditions in place instead of converting V[col + 1].
    struct Step {
        ZT* z = nullptr;
        ZT* znext = nullptr;
        VT* vn = nullptr;
        int slot = 0;
    };
    // the Gram-Schmidt kernels write an FP32 copy of the new basis vector only for FP32 Z
    auto z32 = [](ZT* p) -> float* {
        if constexpr (z64) return nullptr;
        else return p;
    };
    auto enqueue_step = [&](int col, int slot, Step& s, ZT* zin) -> int {
        s.slot = slot;
        int r2 = BSP_OK;
        if (zin) s.z = zin;
        else if ((r2 = Z.take(ctx, s.z))) return r2;
        ZT* z = s.z;
        if constexpr (z64) {
            // all-FP64: z = M^{-1} V[col] in place on the FP64 Z vector, w = A z (FP64)
            if constexpr (vf32) r2 = launch_f2d(ctx, len, Vb[col], z);
            else r2 = launch_copy(ctx, len, Vb[col], z);
            if (!r2) r2 = bsp_precond_apply(ctx, M, z);
        } else if (f32_pre && zin) {
            r2 = precond_apply_f32(ctx, M, z);
        } else if (f32_pre) {
            if constexpr (vf32) {
                BSP_HIP(hipMemcpyAsync(z, Vb[col], sizeof(float) * len, hipMemcpyDeviceToDevice, ctx->stream));
            } else {
                r2 = launch_d2f(ctx, len, Vb[col], z);
            }
            if (!r2) r2 = precond_apply_f32(ctx, M, z);
        } else {
            if constexpr (vf32) r2 = launch_f2d(ctx, len, Vb[col], t64);
            else r2 = launch_copy(ctx, len, Vb[col], t64);
            if (!r2) r2 = bsp_precond_apply(ctx, M, t64);
            if (!r2) r2 = launch_d2f(ctx, len, t64, z);
        }
        if (r2) return r2;
        double* w = wbuf[slot];
        // w = A z
        if constexpr (z64) {
            r2 = bsmv_matrix(ctx, A, z, w, 0);
        } else if (spmv_f32) {
            if (A->val32_ell)
                r2 = launch_bsmv_f32_ell(ctx, n, bs, A->ell_w, A->ell_ci.as<int>(), A->val32.as<float>(), z, w);
            else if (A->val32_chunked)
                r2 = launch_bsmv_f32_chunked(ctx, n, bs, A->rp.as<int>(), A->ci.as<int>(), A->val32.as<float>(), z, w);
            else
                r2 = launch_bsmv_f32(ctx, n, bs, A->rp.as<int>(), A->ci.as<int>(), A->val32.as<float>(), z, w);
        } else {
            r2 = launch_f2d(ctx, len, z, t64);
            if (!r2) r2 = bsmv_matrix(ctx, A, t64, w, 0);
        }
        if (r2) return r2;
        r2 = V.take(ctx, s.vn);
        if (r2) return r2;
        s.znext = nullptr;
        double* hd = hcol.as<double>() + slot * hslot;
        if (use_cgs) {
            if constexpr (!z64)
                if (f32_pre && (r2 = Z.take(ctx, s.znext))) return r2;
            r2 = launch_arnoldi_cgs(ctx, len, w, Vb.data(), col + 1, hd, s.vn, z32(s.znext), cgs_scratch.as<double>(),
                                    cgs_flag.as<int>(), 0, hpin_dev + slot * hslot);
        } else {
            if constexpr (!vf32)
                r2 = launch_arnoldi_mgs(ctx, len, w, Vb.data(), col, hd, s.vn, ctx->red_partials.as<double>());
        }
        if (r2) return r2;
        BSP_HIP(hipEventRecord(ev.e[slot], ctx->stream));
        BSP_HIP(hipStreamWaitEvent(ls_stream, ev.e[slot], 0));
        r2 = ls_column(ctx, ls, col, hd, out_dev + 2 * col, use_cgs ? hd + col + 2 : nullptr, ls_stream);
        if (r2) return r2;
        BSP_HIP(hipEventRecord(ev.e[2 + slot], ls_stream));
        Vb.push_back(s.vn);
        return BSP_OK;
    };
    // The step after the current one is enqueued before the host waits for the current step, so
    // the device works on step j + 1 while the host reads step j's residual estimate.  A step
    // enqueued past the end of the pass (convergence, breakdown, the iteration limit) is dropped
    // unused; none is enqueued past a restart.
    Step cur, nxt;
    if (res > target) {
        rc = enqueue_step(iter, 0, cur, nullptr);
        if (rc) return rc;
    }
    int nc = 0;   // columns of the current cycle
    while (res > target) {
        const bool spec = (restart <= 0 || iter + 1 < restart) && tot + 1 < maxIters;
        if (spec) {
            rc = enqueue_step(iter + 1, cur.slot ^ 1, nxt, cur.znext);
            if (rc) return rc;
        }
        BSP_HIP(hipEventSynchronize(ev.e[2 + cur.slot]));
        const double* hp = hpin.p + cur.slot * hslot;
        if (use_cgs && hp[iter + 2] != 0.0) {
            // cancellation in the first Gram-Schmidt pass: drop the step enqueued on the
            // unfinished vector, orthogonalise once more, apply the corrected column to the
            // least squares, then enqueue that step again
            BSP_HIP(hipStreamSynchronize(ctx->stream));
            BSP_HIP(hipStreamSynchronize(ls_stream));
            if (spec) {
                if (nxt.z != cur.znext) Z.give(nxt.z);
                if (nxt.znext) Z.give(nxt.znext);
                V.give(nxt.vn);
                Vb.pop_back();
            }
            double* hd = hcol.as<double>() + cur.slot * hslot;
            rc = launch_arnoldi_cgs(ctx, len, wbuf[cur.slot], Vb.data(), iter + 1, hd, cur.vn, z32(cur.znext),
                                    cgs_scratch.as<double>(), cgs_flag.as<int>(), 2, nullptr);
            if (rc) return rc;
            if ((rc = ls_column(ctx, ls, iter, hd, out_dev + 2 * iter, nullptr))) return rc;
            BSP_HIP(hipStreamSynchronize(ctx->stream));
            ++st.reorth;
            if (spec) {
                rc = enqueue_step(iter + 1, cur.slot ^ 1, nxt, cur.znext);
                if (rc) return rc;
            }
        }
        res = outp.p[2 * iter];
        const double hlast = outp.p[2 * iter + 1];
        const bool breakdown = hlast < zero_thresh || !std::isfinite(hlast);
        Zb.push_back(cur.z);
        ++iter;
        ++tot;
        nc = iter;
        if (ctx->verbose > 0 && tot % ctx->verbose == 0) {
            std::fprintf(stderr, "[fgmres] iter %d res %.3e (target %.3e) restarts %d\n", tot, res, target, st.restarts);
            std::fflush(stderr);
        }
        if (!std::isfinite(res)) {
            st.nonfinite = true;
            break;
        }
        if (tot >= maxIters || breakdown) break;
        if (iter == restart && res > target) {
            res_hist.push_back(res);
            const size_t hsz = res_hist.size();
            if (hsz > 8 && res > 0.5 * res_hist[hsz - 9]) {
                st.stagnated = true;
                break;
            }
            // x += Z y, then the deflated restart on the device (after the cycle's last update)
            BSP_HIP(hipStreamWaitEvent(ctx->stream, ev.e[2 + cur.slot], 0));
            if ((rc = ls_solve(ctx, ls, nc, sc + SLOT_BETA, true))) return rc;
            if ((rc = ls_deflate_begin(ctx, ls, nc, k))) return rc;
            // the solution update runs on the device while the host solves the eigenproblem
            if constexpr (z64) rc = launch_combine_add(ctx, len, nc, Zb.data(), ls.y.as<double>(), x);
            else rc = launch_combine_add_f32(ctx, len, nc, Zb.data(), ls.y.as<double>(), x);
            if (rc) return rc;
            ++st.restarts;
            int kout = 0;
            if ((rc = ls_deflate_end(ctx, ls, nc, k, true, kout))) return rc;
            const int kp = kout + 1, m = nc + 1;
            BSP_CHECK((int)Vb.size() == m && (int)Zb.size() == nc, BSP_E_NUMERIC, "basis size mismatch");
            std::vector<VT*> Vn(kp);
            std::vector<ZT*> Zn(kout);
            for (int j = 0; j < kp; ++j)
                if ((rc = V.take(ctx, Vn[j]))) return rc;
            for (int j = 0; j < kout; ++j)
                if ((rc = Z.take(ctx, Zn[j]))) return rc;
            // V' = V P_{k+1} (m x kp), Z' = Z P_k (nc x kout), both on the device
            if constexpr (vf32) rc = launch_basis_rotate_f32(ctx, len, m, kp, Vb.data(), ls.P, Vn.data());
            else rc = launch_basis_rotate(ctx, len, m, kp, Vb.data(), ls.P, Vn.data());
            if (!rc && kout > 0) {
                if constexpr (z64) rc = launch_basis_rotate(ctx, len, nc, kout, Zb.data(), ls.Pk, Zn.data());
                else rc = launch_basis_rotate_f32(ctx, len, nc, kout, Zb.data(), ls.Pk, Zn.data());
            }
            if (rc) return rc;
            if ((rc = ls_restart(ctx, ls, kout))) return rc;
            // the old vectors go back to the pools at once: the kernels that will write them
            // again are enqueued on the same stream after the rotations that read them
            for (VT* p : Vb) V.give(p);
            for (ZT* p : Zb) Z.give(p);
            if (cur.znext) Z.give(cur.znext);   // an FP32 copy of a basis vector the rotation replaced
            Vb = Vn;
            Zb = Zn;
            iter = kout;
            nc = 0;
            rc = enqueue_step(iter, 0, cur, nullptr);
            if (rc) return rc;
            continue;
        }
        if (res <= target) break;
        BSP_CHECK(spec, BSP_E_NUMERIC, "fgmres: no step enqueued");
        cur = nxt;
    }
    if (!st.nonfinite && nc > 0) {
        BSP_HIP(hipStreamWaitEvent(ctx->stream, ev.e[2 + cur.slot], 0));
        if ((rc = ls_solve(ctx, ls, nc, sc + SLOT_BETA, false))) return rc;
        if constexpr (z64) rc = launch_combine_add(ctx, len, nc, Zb.data(), ls.y.as<double>(), x);
        else rc = launch_combine_add_f32(ctx, len, nc, Zb.data(), ls.y.as<double>(), x);
        if (rc) return rc;
    }
    BSP_HIP(hipStreamSynchronize(ctx->stream));
    BSP_HIP(hipStreamSynchronize(ls_stream));
    st.iters = tot;
    st.res = res;
    return BSP_OK;
}

}  // namespace

}  // namespace bsp

using namespace bsp;

extern "C" {

int bsp_fgmres_mixed(bsp_ctx* ctx, const bsp_bsr* A, const bsp_precond* M, const double* b, double* x, double tol,
                     int maxIters, int restart, int k, int spmv_precision, double inner_reduction, bsp_stats* stats) {
    return bsp_fgmres_fallback(ctx, A, M, nullptr, b, x, tol, maxIters, restart, k, spmv_precision, inner_reduction,
                               stats);
}

int bsp_fgmres_fallback(bsp_ctx* ctx, const bsp_bsr* A, const bsp_precond* M, const bsp_precond* M_fallback,
                        const double* b, double* x, double tol, int maxIters, int restart, int k, int spmv_precision,
                        double inner_reduction, bsp_stats* stats) {
    if (!ctx || !A || !M || !b || !x) return BSP_E_INVALID;
    BSP_CHECK(A->n_rows == A->n_cols, BSP_E_INVALID, "FGMRES needs a square matrix");
    BSP_CHECK(A->cbs == 0, BSP_E_UNSUPPORTED, "the mixed-precision FGMRES does not take complex matrices");
    BSP_CHECK(M->n == A->n_rows && M->bs == A->bs, BSP_E_INVALID, "preconditioner does not match the matrix");
    BSP_CHECK(!M_fallback || (M_fallback->n == A->n_rows && M_fallback->bs == A->bs), BSP_E_INVALID,
              "fallback preconditioner does not match the matrix");
    BSP_CHECK(b != x, BSP_E_INVALID, "b and x must not alias");
    BSP_CHECK(restart >= 1 && restart <= 62, BSP_E_UNSUPPORTED,
              "FGMRES needs a restart length in [1, 62] (the basis lives in the solver's pools)");
    BSP_CHECK(k >= 0 && k < restart, BSP_E_INVALID, "FGMRES deflation size must be in [0, restart)");
    BSP_CHECK(spmv_precision == BSP_PREC_F64 || spmv_precision == BSP_PREC_F32, BSP_E_INVALID, "bad SpMV precision");
    BSP_HIP(hipSetDevice(ctx->device));
    const int64_t len = (int64_t)A->n_rows * A->bs;
    const bool spmv_f32 = spmv_precision == BSP_PREC_F32;
    if (!(inner_reduction > 0.0 && inner_reduction < 1.0)) inner_reduction = 1e-6;
    int rc;
    if (spmv_f32 && (rc = ensure_values_f32(ctx, A))) return rc;
    float ms = 0.f;
    BSP_HIP(hipEventRecord(ctx->events[12], ctx->stream));
    BSP_HIP(hipMemsetAsync(x, 0, sizeof(double) * len, ctx->stream));
    DevMem rw, t64;
    BSP_HIP(rw.alloc(sizeof(double) * std::max<int64_t>(len, 1)));
    BSP_HIP(t64.alloc(sizeof(double) * std::max<int64_t>(len, 1)));
    double bnorm = 0.0;
    rc = reduce(ctx, 1, len, b, b, SLOT_TMP2, &bnorm);
    if (rc) return rc;
    int total = 0, restarts = 0, passes = 0, reorth = 0;
    double rr = 1.0, res0 = 0.0, res_est = 0.0;
    const bsp_precond* cur = M;
    bool switched = false;
    for (; passes < 100 && total < maxIters; ++passes) {
        const double rr_before = rr;
        // the pass reduces the current residual by `inner_reduction` (or to the target)
        const double pass_red = std::max(inner_reduction, 0.5 * tol / std::max(rr, 1e-300));
        InnerStats st;
        const double red = std::min(pass_red, 0.5);
        // all-FP64 flexible GMRES (FP64 preconditioner and SpMV): Z is stored in FP64 too
        const bool z64 = !cur->f32 && !spmv_f32;
        if (ctx->fgmres_basis_f32 && z64)
            rc = fgmres_pass<float, double>(ctx, A, cur, b, x, red, maxIters - total, restart, k, spmv_f32, st,
                                            rw.as<double>(), t64.as<double>());
        else if (ctx->fgmres_basis_f32)
            rc = fgmres_pass<float, float>(ctx, A, cur, b, x, red, maxIters - total, restart, k, spmv_f32, st,
                                           rw.as<double>(), t64.as<double>());
        else if (z64)
            rc = fgmres_pass<double, double>(ctx, A, cur, b, x, red, maxIters - total, restart, k, spmv_f32, st,
                                             rw.as<double>(), t64.as<double>());
        else
            rc = fgmres_pass<double, float>(ctx, A, cur, b, x, red, maxIters - total, restart, k, spmv_f32, st,
                                            rw.as<double>(), t64.as<double>());
        if (rc) return rc;
        if (passes == 0) res0 = st.beta;
        total += st.iters;
        restarts += st.restarts;
        reorth += st.reorth;
        res_est = st.res;
        rc = true_relres_ext(ctx, A, b, x, rw.as<double>(), bnorm, rr);
        if (rc) return rc;
        if (ctx->verbose > 0) {
            std::fprintf(stderr, "[fgmres] refinement pass %d: %d iterations, true relres %.3e%s\n", passes, st.iters,
                         rr, st.stagnated ? " (stagnated)" : "");
            std::fflush(stderr);
        }
        const bool bad = st.stagnated || st.nonfinite || !std::isfinite(rr) || rr > rr_before;
        if (bad && M_fallback && !switched) {
            switched = true;
            cur = M_fallback;
            if (!std::isfinite(rr)) {
                BSP_HIP(hipMemsetAsync(x, 0, sizeof(double) * len, ctx->stream));
                rr = 1.0;
            }
            if (ctx->verbose > 0) {
                std::fprintf(stderr, "[fgmres] switching to the fallback preconditioner\n");
                std::fflush(stderr);
            }
            continue;
        }
        if (rr <= tol || !std::isfinite(rr) || st.iters == 0) break;
    }
    rc = check_sweep_errors_ext(ctx);
    if (rc) return rc;
    BSP_HIP(hipEventRecord(ctx->events[13], ctx->stream));
    BSP_HIP(hipEventSynchronize(ctx->events[13]));
    BSP_HIP(hipEventElapsedTime(&ms, ctx->events[12], ctx->events[13]));
    const bool conv = rr <= tol;
    if (stats) {
        stats->iters = total;
        stats->restarts = restarts;
        stats->converged = conv ? 1 : 0;
        stats->precond_applies = total;
        stats->res0 = res0;
        stats->res_est = res_est;
        stats->true_relres = rr;
        stats->solve_ms = ms;
        stats->fallback = switched ? 1 : 0;
        stats->reorth = reorth;
    }
    return conv ? BSP_OK : BSP_NOT_CONVERGED;
}

}  // extern "C"
