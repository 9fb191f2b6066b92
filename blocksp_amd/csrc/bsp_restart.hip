// bsp_restart.hip -- launchers of the device least squares and deflated restart of GMRES
// (team routines in bsp_restart.hpp; one 64-thread workgroup per call, workspace in LDS).
#include <algorithm>
#include <cstring>

#include "bsp_dense.hpp"
#include "bsp_ls.hpp"

namespace bsp {

namespace {

constexpr int kTeam = 64;
// per-column scratch of ls_column held in LDS up to this many doubles, else in ls.tbuf
constexpr int kLdsColumn = 4096;

rs::LSState state_of(const DeviceLS& ls) { return ls_state_of(ls); }

__global__ __launch_bounds__(kTeam) void k_ls_fresh(rs::LSState st, const double* __restrict__ beta) {
    rs::ls_fresh_team(rs::DevTeam{}, st, *beta);
}

// skip (optional): the column is not applied when *skip != 0 (a first Gram-Schmidt pass that
// asked for a second one: the caller applies the corrected column later).
__global__ __launch_bounds__(kTeam) void k_ls_column(rs::LSState st, int j, const double* __restrict__ hcol,
                                                     double* out, const double* skip, double* tglobal) {
    extern __shared__ double lds[];
    if (skip && *skip != 0.0) return;
    double* t = tglobal ? tglobal : lds;
    rs::ls_column_team(rs::DevTeam{}, st, j, hcol, t, out);
    if (out && threadIdx.x == 0) __threadfence_system();
}

__global__ __launch_bounds__(kTeam) void k_ls_solve(rs::LSState st, int nc, const double* __restrict__ beta,
                                                    const double* __restrict__ cr, int use_cr,
                                                    const int* __restrict__ kout, double* resv, double* wglobal,
                                                    int staged) {
    extern __shared__ double lds[];
    const int kp = use_cr ? *kout + 1 : 0;
    double* w = staged ? (wglobal ? wglobal : lds) : nullptr;
    rs::ls_solve_team(rs::DevTeam{}, st, nc, *beta, use_cr ? cr : nullptr, kp, resv, w);
}

__global__ __launch_bounds__(kTeam) void k_deflate(rs::LSState st, int n, int k, const double* __restrict__ resv,
                                                   double* P, double* Pk, double* Hnew, double* cr, int* kout,
                                                   double* wsglobal) {
    extern __shared__ double lds[];
    double* ws = wsglobal ? wsglobal : lds;
    const int ko = rs::deflate_team(rs::DevTeam{}, n, st.H, st.ld, resv, k, ws, P, Pk, Hnew, cr);
    if (threadIdx.x == 0) *kout = ko;
}

__global__ __launch_bounds__(kTeam) void k_ls_restart(rs::LSState st, const int* __restrict__ kout,
                                                      const double* __restrict__ Hnew, const double* __restrict__ cr,
                                                      double* wglobal) {
    extern __shared__ double lds[];
    const int ko = *kout;
    if (ko < 0) return;
    rs::ls_restart_team(rs::DevTeam{}, st, ko, Hnew, cr, wglobal ? wglobal : lds);
}

int launched(const bsp_ctx* ctx, const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(ctx, BSP_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
    return BSP_OK;
}

}  // namespace

int ls_alloc(bsp_ctx* ctx, DeviceLS& ls, int maxm, int restart) {
    ls.maxm = maxm;
    ls.ld = maxm + 1;
    const size_t hsz = (size_t)ls.ld * maxm;
    BSP_HIP(ls.H.alloc(sizeof(double) * hsz));
    BSP_HIP(ls.R.alloc(sizeof(double) * hsz));
    BSP_HIP(ls.cs.alloc(sizeof(double) * maxm));
    BSP_HIP(ls.sn.alloc(sizeof(double) * maxm));
    BSP_HIP(ls.g.alloc(sizeof(double) * ls.ld));
    BSP_HIP(ls.Q0t.alloc(sizeof(double) * (rs::kMaxN + 1) * (rs::kMaxN + 1)));
    BSP_HIP(ls.y.alloc(sizeof(double) * maxm));
    BSP_HIP(ls.res.alloc(sizeof(double) * (maxm + 1)));
    BSP_HIP(ls.meta.alloc(sizeof(int) * 2));
    BSP_HIP(ls.resv.alloc(sizeof(double) * ls.ld));
    if (rs::ls_column_ws(maxm) > kLdsColumn) BSP_HIP(ls.tbuf.alloc(sizeof(double) * rs::ls_column_ws(maxm)));
    // workspace of the staged least-squares solve and of the restart QR: LDS when it fits, else
    // device memory (the solve of very long unrestarted cycles keeps the row-by-row route)
    const int64_t wsz = sizeof(double) * std::max<int64_t>(maxm <= 256 ? rs::ls_solve_ws(maxm) : 0,
                                                            restart > 0 ? rs::ls_restart_ws(maxm) : 0);
    if (wsz > ctx->lds_per_block) BSP_HIP(ls.wsg.alloc(wsz));
    BSP_HIP(hipEventCreateWithFlags(&ls.ev, hipEventDisableTiming));
    if (restart > 0) {
        const int n = restart, m = n + 1;
        ls.rbuf_doubles = (size_t)m * (n + 1) + (size_t)n * n + (size_t)(n + 1) * n + (n + 1) + 1;
        BSP_HIP(ls.rbuf.alloc(sizeof(double) * ls.rbuf_doubles));
        ls.P = ls.rbuf.as<double>();
        ls.Pk = ls.P + (size_t)m * (n + 1);
        ls.Hnew = ls.Pk + (size_t)n * n;
        ls.cr = ls.Hnew + (size_t)(n + 1) * n;
        ls.kout = reinterpret_cast<int*>(ls.cr + (n + 1));
        BSP_HIP(hipHostMalloc((void**)&ls.hpin, sizeof(double) * (ls.rbuf_doubles + (size_t)m * (m + 1)),
                              hipHostMallocDefault));
        const size_t ws = sizeof(double) * rs::deflate_ws_doubles(n);
        ls.ws_in_lds = n <= rs::kMaxN && (int64_t)ws <= ctx->lds_per_block;
        if (!ls.ws_in_lds) BSP_HIP(ls.ws.alloc(ws));
    }
    return BSP_OK;
}

int ls_fresh(bsp_ctx* ctx, DeviceLS& ls, const double* beta_dev) {
    hipLaunchKernelGGL(k_ls_fresh, dim3(1), dim3(kTeam), 0, ctx->stream, state_of(ls), beta_dev);
    ls.restarted = false;
    return launched(ctx, "ls_fresh");
}

int ls_column(bsp_ctx* ctx, DeviceLS& ls, int j, const double* hcol_dev, double* out_dev, const double* skip_dev,
              hipStream_t stream) {
    BSP_CHECK(j >= 0 && j < ls.maxm, BSP_E_INVALID, "least squares: column beyond the Hessenberg's capacity");
    const bool lds = rs::ls_column_ws(j) <= kLdsColumn;
    hipLaunchKernelGGL(k_ls_column, dim3(1), dim3(kTeam), lds ? sizeof(double) * rs::ls_column_ws(j) : 0,
                       stream ? stream : ctx->stream,
                       state_of(ls), j, hcol_dev, out_dev, skip_dev, lds ? nullptr : ls.tbuf.as<double>());
    return launched(ctx, "ls_column");
}

int ls_solve(bsp_ctx* ctx, DeviceLS& ls, int nc, const double* beta_dev, bool want_resv) {
    const int64_t wsz = sizeof(double) * (int64_t)rs::ls_solve_ws(nc);
    const bool staged = nc <= 256;
    const bool in_lds = staged && wsz <= ctx->lds_per_block;
    hipLaunchKernelGGL(k_ls_solve, dim3(1), dim3(kTeam), in_lds ? (size_t)wsz : 0, ctx->stream, state_of(ls), nc,
                       beta_dev, ls.cr, ls.restarted ? 1 : 0, ls.kout, want_resv ? ls.resv.as<double>() : nullptr,
                       staged && !in_lds ? ls.wsg.as<double>() : nullptr, staged ? 1 : 0);
    return launched(ctx, "ls_solve");
}

int ls_deflate_begin(bsp_ctx* ctx, DeviceLS& ls, int n, int k) {
    BSP_CHECK(ls.P && n <= ls.maxm, BSP_E_INVALID, "deflated restart: no restart storage");
    ls.defl_device = ctx->deflate_device && n <= rs::kMaxN;
    if (ls.defl_device) {
        const size_t ws = sizeof(double) * rs::deflate_ws_doubles(n);
        const bool in_lds = ls.ws_in_lds && (int64_t)ws <= ctx->lds_per_block;
        hipLaunchKernelGGL(k_deflate, dim3(1), dim3(kTeam), in_lds ? ws : 0, ctx->stream, state_of(ls), n, k,
                           ls.resv.as<double>(), ls.P, ls.Pk, ls.Hnew, ls.cr, ls.kout,
                           in_lds ? nullptr : ls.ws.as<double>());
        if (int rc = launched(ctx, "deflate")) return rc;
        BSP_HIP(hipMemcpyAsync(ls.hpin, ls.kout, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    }
    // the cycle's Hessenberg and residual vector down (for the host route, or for the Schur
    // route if the device route declines)
    const int m = n + 1;
    double* down = ls.hpin + ls.rbuf_doubles;
    BSP_CHECK((size_t)ls.ld * n + m <= (size_t)m * (m + 1), BSP_E_INVALID, "restart staging too small");
    BSP_HIP(hipMemcpyAsync(down, ls.H.p, sizeof(double) * ls.ld * n, hipMemcpyDeviceToHost, ctx->stream));
    BSP_HIP(hipMemcpyAsync(down + (size_t)ls.ld * n, ls.resv.p, sizeof(double) * m, hipMemcpyDeviceToHost,
                           ctx->stream));
    BSP_HIP(hipEventRecord(ls.ev, ctx->stream));
    return BSP_OK;
}

int ls_deflate_end(bsp_ctx* ctx, DeviceLS& ls, int n, int k, bool fast, int& kout) {
    BSP_HIP(hipEventSynchronize(ls.ev));
    kout = -1;
    if (ls.defl_device) {
        std::memcpy(&kout, ls.hpin, sizeof(int));
        if (kout >= 0) return BSP_OK;
        // declined (singular harmonic matrix, QR steps without convergence or nearly dependent
        // eigenvectors): the complex Schur route on the host
        fast = false;
    }
    // host route: P, P_k, H_new, c_r, kout up in one copy through the pinned staging buffer
    const int m = n + 1;
    double* up = ls.hpin;                         // rbuf_doubles
    const double* down = ls.hpin + ls.rbuf_doubles;
    std::vector<double> H((size_t)m * n), P, Hnew, cr;
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < n; ++j) H[(size_t)i * n + j] = down[(size_t)j * ls.ld + i];
    kout = dense::deflate(n, H.data(), down + (size_t)ls.ld * n, k, P, Hnew, cr, fast);
    BSP_CHECK(kout >= 0, BSP_E_NUMERIC, "deflated restart: small eigenproblem failed");
    const int kp = kout + 1;
    // the device block's layout, filled at its actual sizes: offsets as in ls_alloc
    const size_t oPk = (size_t)ls.Pk - (size_t)ls.P, oH = (size_t)ls.Hnew - (size_t)ls.P,
                 oc = (size_t)ls.cr - (size_t)ls.P, ok = (size_t)ls.kout - (size_t)ls.P;
    char* base = reinterpret_cast<char*>(up);
    std::copy(P.begin(), P.end(), reinterpret_cast<double*>(base));
    double* pk = reinterpret_cast<double*>(base + oPk);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < kout; ++j) pk[(size_t)i * kout + j] = P[(size_t)i * kp + j];
    std::copy(Hnew.begin(), Hnew.end(), reinterpret_cast<double*>(base + oH));
    std::copy(cr.begin(), cr.end(), reinterpret_cast<double*>(base + oc));
    std::memcpy(base + ok, &kout, sizeof(int));
    BSP_HIP(hipMemcpyAsync(ls.P, up, ok + sizeof(int), hipMemcpyHostToDevice, ctx->stream));
    // the staging buffer is rewritten only at the next restart, after that restart's event
    // (recorded after this copy) has completed
    ++ls.host_deflations;
    return BSP_OK;
}

int ls_restart(bsp_ctx* ctx, DeviceLS& ls, int kout) {
    const int64_t w = sizeof(double) * (int64_t)rs::ls_restart_ws(kout);
    const bool in_lds = w <= ctx->lds_per_block;
    hipLaunchKernelGGL(k_ls_restart, dim3(1), dim3(kTeam), in_lds ? (size_t)w : 0, ctx->stream, state_of(ls), ls.kout,
                       ls.Hnew, ls.cr, in_lds ? nullptr : ls.wsg.as<double>());
    ls.restarted = true;
    return launched(ctx, "ls_restart");
}

int ls_download_h(bsp_ctx* ctx, const DeviceLS& ls, int nc, std::vector<double>& H) {
    const int m = nc + 1;
    std::vector<double> Hc((size_t)ls.ld * std::max(nc, 1));
    BSP_HIP(hipStreamSynchronize(ctx->stream));
    if (nc > 0) BSP_HIP(hipMemcpy(Hc.data(), ls.H.p, sizeof(double) * ls.ld * nc, hipMemcpyDeviceToHost));
    H.assign((size_t)m * nc, 0.0);
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < nc; ++j) H[(size_t)i * nc + j] = Hc[(size_t)j * ls.ld + i];
    return BSP_OK;
}

}  // namespace bsp
