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
// is done by each row's group from A's structure as it goes, so nothing is planned on the host
// beyond the level schedule, which the forward sweep needs anyway.
#include <algorithm>
#include <climits>
#include <cstring>
#include <numeric>

#include "bsp_factor.hpp"

namespace bsp {

namespace {

constexpr int kWave = 64;

// What a row's group reads: A's structure, the elimination position of every row, and the
// schedule (rows sorted by level, -1 padding slots skipped).
struct FactorArgs {
    const int* rp;
    const int* ci;
    const int* pos;
    const int* order;
    int* diag;          // out: position of each row's (first) diagonal block
    int* err;           // [0]: smallest failing pivot row + 1; [1]: smallest row without diagonal + 1
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

__device__ inline int find_in_row(const int* __restrict__ rp, const int* __restrict__ ci, int i, int j) {
    for (int p = rp[i]; p < rp[i + 1]; ++p)
        if (ci[p] == j) return p;
    return -1;
}

// One level of the factorisation: the rows in schedule slots [s0, s0 + cnt).  F: the factor in
// A's block layout (A's values on entry), dinv: [n][G] inverse of each finished row's pivot
// block (ILU0: U_kk^{-1}; IChol0: L_kk^{-1}).  Row k finds its own tasks: its strict-lower
// blocks (pos[i] < pos[k]) in increasing (pos[i], position) order, and for each task the
// blocks it updates, from A's structure -- the symbolic part of the reference's loops.
template <int BS, bool CHOL>
__global__ __launch_bounds__(kWave) void k_factor_level(FactorArgs fa, int s0, int cnt, double* __restrict__ F,
                                                        double* __restrict__ dinv) {
    constexpr int G = Groups<BS>::G, R = Groups<BS>::R;
    __shared__ double X[R][G];
    __shared__ double Y[R][G];
    __shared__ double W[R][2 * G];
    const int lane = threadIdx.x;
    const int g = lane / G, e = lane % G;
    const int idx = blockIdx.x * R + g;
    if (g >= R || idx >= cnt) return;
    const int k = fa.order[s0 + idx];
    if (k < 0) return;
    const int r = e / BS, c = e % BS;
    const int rs = fa.rp[k], re = fa.rp[k + 1];
    const int pk = fa.pos[k];
    int pd = -1;
    for (int p = rs; p < re && pd < 0; ++p)
        if (fa.ci[p] == k) pd = p;
    if (pd < 0) {
        if (e == 0) atomicMin(fa.err + 1, k + 1);
        return;
    }
    int last_q = -1, last_p = -1;
    for (;;) {
        // next strict-lower block in (elimination position, position) order
        int pki = -1, bq = INT_MAX;
        for (int p = rs; p < re; ++p) {
            const int q = fa.pos[fa.ci[p]];
            if (q >= pk) continue;
            if (q < last_q || (q == last_q && p <= last_p)) continue;
            if (q < bq) {
                bq = q;
                pki = p;
            }
        }
        if (pki < 0) break;
        last_q = bq;
        last_p = pki;
        const int i = fa.ci[pki];
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
        for (int p = rs; p < re; ++p) {
            const int j = fa.ci[p];
            const int qj = fa.pos[j];
            int src;
            if constexpr (CHOL) {
                // L_kj -= L_ki L_ji^T for pos[i] < pos[j] <= pos[k], A(i, j) and L(j, i) present
                if (qj <= bq || qj > pk) continue;
                if (find_in_row(fa.rp, fa.ci, i, j) < 0) continue;
                src = find_in_row(fa.rp, fa.ci, j, i);
            } else {
                // A_kj -= L_ki U_ij for pos[j] > pos[i] and (i, j) in the pattern
                if (qj <= bq) continue;
                src = find_in_row(fa.rp, fa.ci, i, j);
            }
            if (src < 0) continue;
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
            double* d = F + (size_t)p * G + e;
            *d = *d - prod;
        }
        __builtin_amdgcn_wave_barrier();
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    if (e == 0) fa.diag[k] = pd;
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
        if (e == 0) atomicMin(fa.err, k + 1);
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

__global__ void k_set_int(int* p, int v, int cnt) {
    if (blockIdx.x == 0 && (int)threadIdx.x < cnt) p[threadIdx.x] = v;
}

template <int BS, bool CHOL>
void launch_level(hipStream_t st, const FactorArgs& fa, int s0, int cnt, double* F, double* dinv) {
    constexpr int R = Groups<BS>::R;
    hipLaunchKernelGGL((k_factor_level<BS, CHOL>), dim3((cnt + R - 1) / R), dim3(kWave), 0, st, fa, s0, cnt, F, dinv);
}

template <bool CHOL>
bool launch_level_bs(int bs, hipStream_t st, const FactorArgs& fa, int s0, int cnt, double* F, double* dinv) {
    switch (bs) {
        case 1: launch_level<1, CHOL>(st, fa, s0, cnt, F, dinv); return true;
        case 2: launch_level<2, CHOL>(st, fa, s0, cnt, F, dinv); return true;
        case 3: launch_level<3, CHOL>(st, fa, s0, cnt, F, dinv); return true;
        case 4: launch_level<4, CHOL>(st, fa, s0, cnt, F, dinv); return true;
        case 5: launch_level<5, CHOL>(st, fa, s0, cnt, F, dinv); return true;
        case 6: launch_level<6, CHOL>(st, fa, s0, cnt, F, dinv); return true;
        case 7: launch_level<7, CHOL>(st, fa, s0, cnt, F, dinv); return true;
        case 8: launch_level<8, CHOL>(st, fa, s0, cnt, F, dinv); return true;
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

int check_err(bsp_ctx* ctx, const DevMem& err, const char* what, const char* what2 = nullptr) {
    int h[2] = {INT_MAX, INT_MAX};
    BSP_HIP(hipMemcpyAsync(h, err.p, sizeof(int) * (what2 ? 2 : 1), hipMemcpyDeviceToHost, ctx->stream));
    BSP_HIP(hipStreamSynchronize(ctx->stream));
    if (what2 && h[1] != INT_MAX) return fail(ctx, BSP_E_NUMERIC, std::string(what2) + " in row " + std::to_string(h[1] - 1));
    if (h[0] != INT_MAX) return fail(ctx, BSP_E_NUMERIC, std::string(what) + " in row " + std::to_string(h[0] - 1));
    return BSP_OK;
}

}  // namespace

bool factor_device_supported(int bs) { return bs >= 1 && bs <= 8; }

int factor_device(bsp_ctx* ctx, const bsp_bsr* A, const int* pos_dev, const Sweep& lower, bool chol, DevMem& F,
                  DevMem& dinv, DevMem& diag) {
    const int n = A->n_rows, bs = A->bs;
    BSP_CHECK(factor_device_supported(bs), BSP_E_UNSUPPORTED, "device factorisation supports block sizes 1..8");
    BSP_CHECK(lower.n == n && (int)lower.level_ptr.size() == lower.nlevels + 1, BSP_E_INVALID,
              "device factorisation needs the forward sweep's schedule");
    const size_t b2 = (size_t)bs * bs;
    BSP_HIP(hipSetDevice(ctx->device));
    BSP_HIP(F.alloc(sizeof(double) * std::max<size_t>((size_t)A->nnzb * b2, 1)));
    BSP_HIP(dinv.alloc(sizeof(double) * std::max<size_t>((size_t)n * b2, 1)));
    BSP_HIP(diag.alloc(sizeof(int) * std::max(n, 1)));
    if (A->nnzb)
        BSP_HIP(hipMemcpyAsync(F.p, A->val.p, sizeof(double) * A->nnzb * b2, hipMemcpyDeviceToDevice, ctx->stream));
    DevMem err;
    BSP_HIP(err.alloc(2 * sizeof(int)));
    hipLaunchKernelGGL(k_set_int, dim3(1), dim3(64), 0, ctx->stream, err.as<int>(), INT_MAX, 2);
    FactorArgs fa{A->rp.as<int>(), A->ci.as<int>(), pos_dev, lower.order.as<int>(), diag.as<int>(), err.as<int>()};
    for (int l = 0; l < lower.nlevels; ++l) {
        const int s0 = lower.level_ptr[l], cnt = lower.level_ptr[l + 1] - s0;
        if (cnt <= 0) continue;
        const bool ok = chol ? launch_level_bs<true>(bs, ctx->stream, fa, s0, cnt, F.as<double>(), dinv.as<double>())
                             : launch_level_bs<false>(bs, ctx->stream, fa, s0, cnt, F.as<double>(), dinv.as<double>());
        if (!ok) return fail(ctx, BSP_E_UNSUPPORTED, "device factorisation: block size");
    }
    BSP_HIP(hipGetLastError());
    return check_err(ctx, err, chol ? "IChol0: pivot block is not positive definite" : "ILU0: singular pivot block",
                     chol ? "IChol0: missing diagonal block" : "ILU0: missing diagonal block");
}

int invert_diag_device(bsp_ctx* ctx, int n, int bs, const int* diag, const double* F, void* out, void* out_t,
                       int precision) {
    DevMem err;
    BSP_HIP(err.alloc(sizeof(int)));
    hipLaunchKernelGGL(k_set_int, dim3(1), dim3(64), 0, ctx->stream, err.as<int>(), INT_MAX, 1);
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
