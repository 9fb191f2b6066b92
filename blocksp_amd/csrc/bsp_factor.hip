// bsp_factor.hip -- block ILU(0) and IChol(0) factorisation on the device (gfx950).
//
// The reference factors on the host, right-looking (blocksp_preconditioners.hpp:274-291 for
// ILU0, 374-429 for IChol0): at step i every later row k with a block (k, i) gets
// L_ki = A_ki U_ii^{-1} (IChol0: A_ki L_ii^{-T}) and then A_kj -= L_ki A_ij (IChol0:
// L_kj -= L_ki L_ji^T) for the j in both patterns.  Each block (k, j) receives its updates in
// elimination order of i, and L_ki only needs the updates of steps before i, so the same
// arithmetic runs row by row: row k takes its strict-lower blocks in elimination order, turns
// each into L_ki and applies that block's updates to the rest of row k, then factors (IChol0)
// and inverts its diagonal block.  Row k depends exactly on the rows of its strict-lower
// blocks, i.e. the forward-sweep level sets: one launch per level, and within a level one
// group of bs*bs lanes per row (lane = entry (r, c) of a block), several groups per 64-wide
// wavefront for small blocks.  The symbolic part (which block updates which, in what order)
// is integer work done once on the host from the structure.
#include <algorithm>
#include <climits>
#include <cstring>
#include <numeric>

#include "bsp_factor.hpp"

namespace bsp {

namespace {

constexpr int kWave = 64;

// Plan arrays on the device (see bsp_factor.hpp for their meaning).
struct PlanArgs {
    const int* rows;        // rows sorted by level
    const int* row_task;    // [n + 1] strict-lower tasks of each row (by row id)
    const int* task_p;      // block (k, i) of the task
    const int* task_i;      // its column i
    const int* task_u;      // [ntasks + 1] update range of each task
    const int* upd_src;     // ILU0: block (i, j); IChol0: block (j, i)
    const int* upd_dst;     // block (k, j)
    const int* diag;        // [n] position of the diagonal block
};

template <int BS>
struct Groups {
    static constexpr int G = BS * BS;
    static constexpr int R = kWave / G;   // rows per wavefront
};

// Inverse of the group's block W[0..G) into W[G..2G) by Gauss-Jordan with partial pivoting
// (the textbook method of the host's dense::invert: first maximal pivot, row scaling, then
// elimination of the other rows).  All lanes of the group run the same control flow; false
// when a pivot is zero or not finite.
template <int BS>
__device__ bool group_invert(double* W, int e) {
    constexpr int G = BS * BS;
    const int r = e / BS, c = e % BS;
    double* m = W;
    double* inv = W + G;
    inv[e] = r == c ? 1.0 : 0.0;
    __builtin_amdgcn_wave_barrier();
    for (int col = 0; col < BS; ++col) {
        int piv = col;
        double best = fabs(m[col * BS + col]);
        for (int rr = col + 1; rr < BS; ++rr) {
            const double v = fabs(m[rr * BS + col]);
            if (v > best) {
                best = v;
                piv = rr;
            }
        }
        if (best == 0.0 || !isfinite(best)) return false;
        if (piv != col) {
            double tm = 0.0, ti = 0.0;
            if (r == col) {
                tm = m[piv * BS + c];
                ti = inv[piv * BS + c];
            } else if (r == piv) {
                tm = m[col * BS + c];
                ti = inv[col * BS + c];
            }
            __builtin_amdgcn_wave_barrier();
            if (r == col || r == piv) {
                m[e] = tm;
                inv[e] = ti;
            }
            __builtin_amdgcn_wave_barrier();
        }
        const double d = m[col * BS + col];
        double mv = m[e], iv = inv[e];
        __builtin_amdgcn_wave_barrier();
        if (r == col) {
            m[e] = mv / d;
            inv[e] = iv / d;
        }
        __builtin_amdgcn_wave_barrier();
        if (r != col) {
            const double f = m[r * BS + col];
            const double pm = m[col * BS + c], pi = inv[col * BS + c];
            mv = m[e];
            iv = inv[e];
            if (f != 0.0) {
                mv -= f * pm;
                iv -= f * pi;
            }
        }
        __builtin_amdgcn_wave_barrier();
        if (r != col) {
            m[e] = mv;
            inv[e] = iv;
        }
        __builtin_amdgcn_wave_barrier();
    }
    return true;
}

// In-place lower Cholesky of the group's block W[0..G) (upper part zeroed), the host's
// dense::cholesky_lower column by column; false if not positive definite.
template <int BS>
__device__ bool group_cholesky(double* W, int e) {
    const int r = e / BS, c = e % BS;
    double* a = W;
    for (int j = 0; j < BS; ++j) {
        double t = 0.0;
        if (c == j && r >= j) {
            t = a[r * BS + j];
            for (int p = 0; p < j; ++p) t -= a[r * BS + p] * a[j * BS + p];
        }
        __builtin_amdgcn_wave_barrier();
        if (c == j && r == j) a[e] = t;
        __builtin_amdgcn_wave_barrier();
        const double s = a[j * BS + j];
        if (!(s > 0.0)) return false;
        const double d = sqrt(s);
        __builtin_amdgcn_wave_barrier();
        if (c == j && r > j) a[e] = t / d;
        if (c == j && r == j) a[e] = d;
        __builtin_amdgcn_wave_barrier();
    }
    if (c > r) a[e] = 0.0;
    __builtin_amdgcn_wave_barrier();
    return true;
}

// One level of the factorisation: `cnt` rows starting at rows[s0].  F: the factor in A's
// block layout (A's values on entry), dinv: [n][G] inverse of each finished row's pivot block
// (ILU0: U_kk^{-1}; IChol0: L_kk^{-1}).  err: smallest failing row + 1 (atomicMin).
template <int BS, bool CHOL>
__global__ __launch_bounds__(kWave) void k_factor_level(PlanArgs pa, int s0, int cnt, double* __restrict__ F,
                                                        double* __restrict__ dinv, int* __restrict__ err) {
    constexpr int G = Groups<BS>::G, R = Groups<BS>::R;
    __shared__ double X[R][G];
    __shared__ double Y[R][G];
    __shared__ double W[R][2 * G];
    const int lane = threadIdx.x;
    const int g = lane / G, e = lane % G;
    const int idx = blockIdx.x * R + g;
    if (g >= R || idx >= cnt) return;
    const int r = e / BS, c = e % BS;
    const int k = pa.rows[s0 + idx];
    for (int t = pa.row_task[k]; t < pa.row_task[k + 1]; ++t) {
        const int pki = pa.task_p[t], i = pa.task_i[t];
        // updates of earlier tasks of this row (other lanes' stores) must be visible
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        X[g][e] = F[(size_t)pki * G + e];
        Y[g][e] = dinv[(size_t)i * G + e];
        __builtin_amdgcn_wave_barrier();
        double s = 0.0;
        if constexpr (CHOL) {
            for (int m = 0; m < BS; ++m) s += X[g][r * BS + m] * Y[g][c * BS + m];   // L_ki L_ii^{-T}
        } else {
            for (int m = 0; m < BS; ++m) s += X[g][r * BS + m] * Y[g][m * BS + c];   // L_ki U_ii^{-1}
        }
        __builtin_amdgcn_wave_barrier();
        X[g][e] = s;
        F[(size_t)pki * G + e] = s;
        for (int u = pa.task_u[t]; u < pa.task_u[t + 1]; ++u) {
            const int src = pa.upd_src[u], dst = pa.upd_dst[u];
            __builtin_amdgcn_wave_barrier();
            // IChol0's diagonal update reads the block this task has just written
            Y[g][e] = src == pki ? s : F[(size_t)src * G + e];
            __builtin_amdgcn_wave_barrier();
            double prod = 0.0;
            if constexpr (CHOL) {
                for (int m = 0; m < BS; ++m) prod += X[g][r * BS + m] * Y[g][c * BS + m];   // L_ki L_ji^T
            } else {
                for (int m = 0; m < BS; ++m) prod += X[g][r * BS + m] * Y[g][m * BS + c];   // L_ki U_ij
            }
            double* d = F + (size_t)dst * G + e;
            *d = *d - prod;
        }
        __builtin_amdgcn_wave_barrier();
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    const int pd = pa.diag[k];
    W[g][e] = F[(size_t)pd * G + e];
    __builtin_amdgcn_wave_barrier();
    bool ok = true;
    if constexpr (CHOL) {
        ok = group_cholesky<BS>(W[g], e);
        if (ok) F[(size_t)pd * G + e] = W[g][e];
        __builtin_amdgcn_wave_barrier();
    }
    if (ok) ok = group_invert<BS>(W[g], e);
    if (!ok) {
        if (e == 0) atomicMin(err, k + 1);
        return;
    }
    dinv[(size_t)k * G + e] = W[g][G + e];
}

// Inverse (and optionally its transpose) of the diagonal blocks of F, in the output type.
__device__ inline uint16_t f64_to_bf16(double v) {
    const float f = (float)v;
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

template <typename OT>
__device__ inline OT convert_out(double v) {
    if constexpr (std::is_same_v<OT, uint16_t>) return f64_to_bf16(v);
    else return (OT)v;
}

template <int BS, typename OT>
__global__ __launch_bounds__(kWave) void k_invert_diag(int n, const int* __restrict__ diag,
                                                       const double* __restrict__ F, OT* __restrict__ out,
                                                       OT* __restrict__ out_t, int* __restrict__ err) {
    constexpr int G = Groups<BS>::G, R = Groups<BS>::R;
    __shared__ double W[R][2 * G];
    const int lane = threadIdx.x;
    const int g = lane / G, e = lane % G;
    const int k = blockIdx.x * R + g;
    if (g >= R || k >= n) return;
    W[g][e] = F[(size_t)diag[k] * G + e];
    __builtin_amdgcn_wave_barrier();
    if (!group_invert<BS>(W[g], e)) {
        if (e == 0) atomicMin(err, k + 1);
        return;
    }
    const double v = W[g][G + e];
    out[(size_t)k * G + e] = convert_out<OT>(v);
    if (out_t) {
        const int r = e / BS, c = e % BS;
        out_t[(size_t)k * G + c * BS + r] = convert_out<OT>(v);
    }
}

// out[q] = F[map[q]] (or its transpose), converted to the output type.
template <typename OT>
__global__ void k_gather_blocks(int64_t cnt, int bs, const int* __restrict__ map, const double* __restrict__ F,
                                OT* __restrict__ out, int transpose) {
    const int b2 = bs * bs;
    const int64_t total = cnt * b2;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t q = t / b2;
        const int e = (int)(t % b2);
        const int src = transpose ? (e % bs) * bs + e / bs : e;
        out[t] = convert_out<OT>(F[(size_t)map[q] * b2 + src]);
    }
}

__global__ void k_set_int(int* p, int v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

template <typename T>
hipError_t upload_vec(DevMem& m, const std::vector<T>& v) {
    hipError_t e = m.alloc(sizeof(T) * std::max<size_t>(v.size(), 1));
    if (e != hipSuccess || v.empty()) return e;
    return hipMemcpy(m.p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice);
}

int find_col(const std::vector<int32_t>& rp, const std::vector<int32_t>& ci, int i, int j) {
    for (int p = rp[i]; p < rp[i + 1]; ++p)
        if (ci[p] == j) return p;
    return -1;
}

template <int BS, bool CHOL>
void launch_level(hipStream_t st, const PlanArgs& pa, int s0, int cnt, double* F, double* dinv, int* err) {
    constexpr int R = Groups<BS>::R;
    hipLaunchKernelGGL((k_factor_level<BS, CHOL>), dim3((cnt + R - 1) / R), dim3(kWave), 0, st, pa, s0, cnt, F, dinv,
                       err);
}

template <bool CHOL>
bool launch_level_bs(int bs, hipStream_t st, const PlanArgs& pa, int s0, int cnt, double* F, double* dinv, int* err) {
    switch (bs) {
        case 1: launch_level<1, CHOL>(st, pa, s0, cnt, F, dinv, err); return true;
        case 2: launch_level<2, CHOL>(st, pa, s0, cnt, F, dinv, err); return true;
        case 3: launch_level<3, CHOL>(st, pa, s0, cnt, F, dinv, err); return true;
        case 4: launch_level<4, CHOL>(st, pa, s0, cnt, F, dinv, err); return true;
        case 5: launch_level<5, CHOL>(st, pa, s0, cnt, F, dinv, err); return true;
        case 6: launch_level<6, CHOL>(st, pa, s0, cnt, F, dinv, err); return true;
        case 7: launch_level<7, CHOL>(st, pa, s0, cnt, F, dinv, err); return true;
        case 8: launch_level<8, CHOL>(st, pa, s0, cnt, F, dinv, err); return true;
        default: return false;
    }
}

template <int BS, typename OT>
void launch_inv(hipStream_t st, int n, const int* diag, const double* F, OT* out, OT* out_t, int* err) {
    constexpr int R = Groups<BS>::R;
    hipLaunchKernelGGL((k_invert_diag<BS, OT>), dim3((n + R - 1) / R), dim3(kWave), 0, st, n, diag, F, out, out_t,
                       err);
}

template <typename OT>
bool launch_inv_bs(int bs, hipStream_t st, int n, const int* diag, const double* F, OT* out, OT* out_t, int* err) {
    switch (bs) {
        case 1: launch_inv<1, OT>(st, n, diag, F, out, out_t, err); return true;
        case 2: launch_inv<2, OT>(st, n, diag, F, out, out_t, err); return true;
        case 3: launch_inv<3, OT>(st, n, diag, F, out, out_t, err); return true;
        case 4: launch_inv<4, OT>(st, n, diag, F, out, out_t, err); return true;
        case 5: launch_inv<5, OT>(st, n, diag, F, out, out_t, err); return true;
        case 6: launch_inv<6, OT>(st, n, diag, F, out, out_t, err); return true;
        case 7: launch_inv<7, OT>(st, n, diag, F, out, out_t, err); return true;
        case 8: launch_inv<8, OT>(st, n, diag, F, out, out_t, err); return true;
        default: return false;
    }
}

int check_err(bsp_ctx* ctx, const DevMem& err, const char* what) {
    int h = INT_MAX;
    BSP_HIP(hipMemcpyAsync(&h, err.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    BSP_HIP(hipStreamSynchronize(ctx->stream));
    if (h != INT_MAX) return fail(ctx, BSP_E_NUMERIC, std::string(what) + " in row " + std::to_string(h - 1));
    return BSP_OK;
}

}  // namespace

bool factor_device_supported(int bs) { return bs >= 1 && bs <= 8; }

int plan_factor(bsp_ctx* ctx, const bsp_bsr* A, const std::vector<int32_t>& pos, bool chol, FactorPlan& P) {
    const int n = A->n_rows;
    const auto& rp = A->h_rp;
    const auto& ci = A->h_ci;
    std::vector<int32_t> elim(n);
    for (int i = 0; i < n; ++i) elim[pos[i]] = i;
    std::vector<int32_t> row_task(n + 1, 0), task_p, task_i, task_u, src, dst, diag(n, -1), level(n, 0);
    task_p.reserve(A->nnzb);
    task_i.reserve(A->nnzb);
    task_u.reserve(A->nnzb + 1);
    src.reserve(3 * A->nnzb);
    dst.reserve(3 * A->nnzb);
    // tasks are stored by row id; a row's strict-lower blocks in elimination order
    std::vector<std::pair<int32_t, int32_t>> lo;
    std::vector<int32_t> cnt(n, 0);
    int nlev = 0;
    // levels in elimination order (a row's strict-lower rows are eliminated before it)
    for (int t = 0; t < n; ++t) {
        const int k = elim[t];
        int l = 0;
        for (int p = rp[k]; p < rp[k + 1]; ++p)
            if (pos[ci[p]] < t) l = std::max(l, level[ci[p]] + 1);
        level[k] = l;
        nlev = std::max(nlev, l + 1);
    }
    for (int k = 0; k < n; ++k) {
        lo.clear();
        for (int p = rp[k]; p < rp[k + 1]; ++p) {
            const int j = ci[p];
            if (j == k && diag[k] < 0) diag[k] = p;
            if (pos[j] < pos[k]) lo.emplace_back(pos[j], p);
        }
        BSP_CHECK(diag[k] >= 0, BSP_E_NUMERIC,
                  std::string(chol ? "IChol0" : "ILU0") + ": missing diagonal block in row " + std::to_string(k));
        std::sort(lo.begin(), lo.end());
        for (auto [pi, pki] : lo) {
            const int i = ci[pki];
            task_p.push_back(pki);
            task_i.push_back(i);
            task_u.push_back((int32_t)src.size());
            for (int p = rp[k]; p < rp[k + 1]; ++p) {
                const int j = ci[p];
                if (!chol) {
                    // A_kj -= L_ki U_ij for pos[j] > pos[i] and (i, j) in the pattern
                    if (pos[j] <= pi) continue;
                    const int pij = find_col(rp, ci, i, j);
                    if (pij < 0) continue;
                    src.push_back(pij);
                    dst.push_back(p);
                } else {
                    // L_kj -= L_ki L_ji^T for pos[i] < pos[j] <= pos[k], A(i, j) and L(j, i) present
                    if (pos[j] <= pi || pos[j] > pos[k]) continue;
                    if (find_col(rp, ci, i, j) < 0) continue;
                    const int pji = find_col(rp, ci, j, i);
                    if (pji < 0) continue;
                    src.push_back(pji);
                    dst.push_back(p);
                }
            }
        }
        row_task[k + 1] = (int32_t)task_p.size();
    }
    task_u.push_back((int32_t)src.size());
    // rows sorted by level, ascending row within a level
    std::vector<int32_t> lev_cnt(nlev + 1, 0);
    for (int k = 0; k < n; ++k) lev_cnt[level[k] + 1]++;
    for (int l = 0; l < nlev; ++l) lev_cnt[l + 1] += lev_cnt[l];
    P.level_ptr.assign(lev_cnt.begin(), lev_cnt.end());
    std::vector<int32_t> rows(n), fill(lev_cnt.begin(), lev_cnt.end() - 1);
    for (int k = 0; k < n; ++k) rows[fill[level[k]]++] = k;
    P.n = n;
    P.ntasks = (int64_t)task_p.size();
    P.nupd = (int64_t)src.size();
    BSP_HIP(upload_vec(P.rows, rows));
    BSP_HIP(upload_vec(P.row_task, row_task));
    BSP_HIP(upload_vec(P.task_p, task_p));
    BSP_HIP(upload_vec(P.task_i, task_i));
    BSP_HIP(upload_vec(P.task_u, task_u));
    BSP_HIP(upload_vec(P.upd_src, src));
    BSP_HIP(upload_vec(P.upd_dst, dst));
    BSP_HIP(upload_vec(P.diag, diag));
    return BSP_OK;
}

int factor_device(bsp_ctx* ctx, const bsp_bsr* A, const FactorPlan& P, bool chol, DevMem& F, DevMem& dinv) {
    const int n = A->n_rows, bs = A->bs;
    BSP_CHECK(factor_device_supported(bs), BSP_E_UNSUPPORTED, "device factorisation supports block sizes 1..8");
    const size_t b2 = (size_t)bs * bs;
    BSP_HIP(hipSetDevice(ctx->device));
    BSP_HIP(F.alloc(sizeof(double) * std::max<size_t>((size_t)A->nnzb * b2, 1)));
    BSP_HIP(dinv.alloc(sizeof(double) * std::max<size_t>((size_t)n * b2, 1)));
    if (A->nnzb)
        BSP_HIP(hipMemcpyAsync(F.p, A->val.p, sizeof(double) * A->nnzb * b2, hipMemcpyDeviceToDevice, ctx->stream));
    DevMem err;
    BSP_HIP(err.alloc(sizeof(int)));
    hipLaunchKernelGGL(k_set_int, dim3(1), dim3(64), 0, ctx->stream, err.as<int>(), INT_MAX);
    PlanArgs pa{P.rows.as<int>(),   P.row_task.as<int>(), P.task_p.as<int>(),  P.task_i.as<int>(),
                P.task_u.as<int>(), P.upd_src.as<int>(),  P.upd_dst.as<int>(), P.diag.as<int>()};
    const int nlev = (int)P.level_ptr.size() - 1;
    for (int l = 0; l < nlev; ++l) {
        const int s0 = P.level_ptr[l], cnt = P.level_ptr[l + 1] - s0;
        if (cnt <= 0) continue;
        const bool ok = chol ? launch_level_bs<true>(bs, ctx->stream, pa, s0, cnt, F.as<double>(), dinv.as<double>(),
                                                     err.as<int>())
                             : launch_level_bs<false>(bs, ctx->stream, pa, s0, cnt, F.as<double>(),
                                                      dinv.as<double>(), err.as<int>());
        if (!ok) return fail(ctx, BSP_E_UNSUPPORTED, "device factorisation: block size");
    }
    BSP_HIP(hipGetLastError());
    return check_err(ctx, err, chol ? "IChol0: pivot block is not positive definite"
                                    : "ILU0: singular pivot block");
}

int invert_diag_device(bsp_ctx* ctx, int n, int bs, const int* diag, const double* F, void* out, void* out_t,
                       int precision) {
    DevMem err;
    BSP_HIP(err.alloc(sizeof(int)));
    hipLaunchKernelGGL(k_set_int, dim3(1), dim3(64), 0, ctx->stream, err.as<int>(), INT_MAX);
    bool ok;
    if (n > 0) {
        if (precision == BSP_PREC_BF16)
            ok = launch_inv_bs<uint16_t>(bs, ctx->stream, n, diag, F, (uint16_t*)out, (uint16_t*)out_t, err.as<int>());
        else if (precision == BSP_PREC_F32)
            ok = launch_inv_bs<float>(bs, ctx->stream, n, diag, F, (float*)out, (float*)out_t, err.as<int>());
        else
            ok = launch_inv_bs<double>(bs, ctx->stream, n, diag, F, (double*)out, (double*)out_t, err.as<int>());
        if (!ok) return fail(ctx, BSP_E_UNSUPPORTED, "diagonal inverses: block size");
        BSP_HIP(hipGetLastError());
    }
    return check_err(ctx, err, "singular diagonal block");
}

int gather_blocks(bsp_ctx* ctx, int64_t cnt, int bs, const int* map, const double* F, void* out, int precision,
                  bool transpose) {
    if (cnt <= 0) return BSP_OK;
    const int64_t total = cnt * bs * bs;
    const int grid = (int)std::min<int64_t>((total + 255) / 256, 65536);
    if (precision == BSP_PREC_BF16)
        hipLaunchKernelGGL(k_gather_blocks<uint16_t>, dim3(grid), dim3(256), 0, ctx->stream, cnt, bs, map, F,
                           (uint16_t*)out, transpose ? 1 : 0);
    else if (precision == BSP_PREC_F32)
        hipLaunchKernelGGL(k_gather_blocks<float>, dim3(grid), dim3(256), 0, ctx->stream, cnt, bs, map, F,
                           (float*)out, transpose ? 1 : 0);
    else
        hipLaunchKernelGGL(k_gather_blocks<double>, dim3(grid), dim3(256), 0, ctx->stream, cnt, bs, map, F,
                           (double*)out, transpose ? 1 : 0);
    BSP_HIP(hipGetLastError());
    return BSP_OK;
}

}  // namespace bsp
