// bsp_mixed.hip -- reduced-precision kernels of the mixed-precision solve path (gfx950).
//
//  * k_bsmv_lp<BS, AT, XT, YT>: y = A x with A in AT (float), x in XT, FP32
//    accumulation, y stored as YT (float, or double for the FP64 Krylov basis);
//    same lane mapping as the FP64 bsmv (BS lanes per block row, lane r owns row r).
//  * k_bsmv_lp_perm<BS>: the FP32-value, FP64-result form over chunk-major FP32 values;
//  * conversions double <-> float, and the FP64 combination x += sum_j c_j Z_j with
//    FP32 basis vectors Z_j (the solution update of flexible GMRES).
#include "bsp_internal.hpp"

#include <utility>

#include <algorithm>

namespace bsp {

namespace {

constexpr int kWave = 64;
constexpr int kBlock = 256;

template <int BS, typename T>
__device__ __forceinline__ void load_vec(const T* __restrict__ src, T (&a)[BS]) {
    if constexpr (std::is_same_v<T, float> && BS % 4 == 0) {
        const float4* s4 = reinterpret_cast<const float4*>(src);
#pragma unroll
        for (int c = 0; c < BS / 4; ++c) {
            const float4 t = s4[c];
            a[4 * c] = t.x;
            a[4 * c + 1] = t.y;
            a[4 * c + 2] = t.z;
            a[4 * c + 3] = t.w;
        }
    } else if constexpr (std::is_same_v<T, double> && BS % 2 == 0) {
        const double2* s2 = reinterpret_cast<const double2*>(src);
#pragma unroll
        for (int c = 0; c < BS / 2; ++c) {
            const double2 t = s2[c];
            a[2 * c] = t.x;
            a[2 * c + 1] = t.y;
        }
    } else {
#pragma unroll
        for (int c = 0; c < BS; ++c) a[c] = src[c];
    }
}

template <int BS, typename AT, typename XT, typename YT>
__global__ __launch_bounds__(kBlock) void k_bsmv_lp(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                                                    const AT* __restrict__ val, const XT* __restrict__ x,
                                                    YT* __restrict__ y) {
    constexpr int RPW = kWave / BS;
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = (blockIdx.x * kBlock + threadIdx.x) / kWave;
    const int g = lane / BS;
    const int r = lane - g * BS;
    if (g >= RPW) return;
    const int i = wave * RPW + g;
    if (i >= n) return;
    float acc = 0.0f;
    const int p0 = rp[i], p1 = rp[i + 1];
    int p = p0;
    for (; p + 1 < p1; p += 2) {
        const int c0 = ci[p], c1 = ci[p + 1];
        AT a0[BS], a1[BS];
        XT x0[BS], x1[BS];
        load_vec<BS, AT>(val + ((size_t)p * BS + r) * BS, a0);
        load_vec<BS, AT>(val + ((size_t)(p + 1) * BS + r) * BS, a1);
        load_vec<BS, XT>(x + (size_t)c0 * BS, x0);
        load_vec<BS, XT>(x + (size_t)c1 * BS, x1);
        float s0 = 0.0f, s1 = 0.0f;
#pragma unroll
        for (int c = 0; c < BS; ++c) {
            s0 = fmaf((float)a0[c], (float)x0[c], s0);
            s1 = fmaf((float)a1[c], (float)x1[c], s1);
        }
        acc += s0;
        acc += s1;
    }
    if (p < p1) {
        const int c0 = ci[p];
        AT a0[BS];
        XT x0[BS];
        load_vec<BS, AT>(val + ((size_t)p * BS + r) * BS, a0);
        load_vec<BS, XT>(x + (size_t)c0 * BS, x0);
        float s0 = 0.0f;
#pragma unroll
        for (int c = 0; c < BS; ++c) s0 = fmaf((float)a0[c], (float)x0[c], s0);
        acc += s0;
    }
    y[(size_t)i * BS + r] = (YT)acc;
}

// Same product with the FP32 values in chunk-major order: in block p, lane r's row is stored as
// BS/4 chunks of 4 floats, chunk k of row r at valp[p*BS*BS + (k*BS + r)*4], so for each k the BS
// lanes of a block row read 16*BS contiguous bytes.  Each XCD walks a contiguous range of rows
// (hardware workgroup w runs on XCD w % 8) and the values, read once, are streamed.  Per lane the
// values and the order of the arithmetic are those of k_bsmv_lp: the results are the same bits.
typedef float f4v __attribute__((ext_vector_type(4)));

template <int BS>
__global__ __launch_bounds__(kBlock) void k_bsmv_lp_perm(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                                                         const float* __restrict__ valp, const float* __restrict__ x,
                                                         double* __restrict__ y, int nwg) {
    static_assert(BS % 4 == 0, "chunk-major FP32 layout needs a block size divisible by 4");
    constexpr int RPW = kWave / BS;
    const int lane = threadIdx.x & (kWave - 1);
    const int per = (nwg + 7) / 8;
    const int wg = (blockIdx.x % 8) * per + blockIdx.x / 8;
    if (wg >= nwg) return;
    const int wave = (wg * kBlock + threadIdx.x) / kWave;
    const int g = lane / BS;
    const int r = lane - g * BS;
    if (g >= RPW) return;
    const int i = wave * RPW + g;
    if (i >= n) return;
    auto load_row = [&](int p, float (&a)[BS]) {
        const f4v* s4 = reinterpret_cast<const f4v*>(valp + (size_t)p * BS * BS) + r;
#pragma unroll
        for (int k = 0; k < BS / 4; ++k) {
            const f4v t = __builtin_nontemporal_load(s4 + k * BS);
            a[4 * k] = t.x;
            a[4 * k + 1] = t.y;
            a[4 * k + 2] = t.z;
            a[4 * k + 3] = t.w;
        }
    };
    float acc = 0.0f;
    const int p0 = rp[i], p1 = rp[i + 1];
    int p = p0;
    for (; p + 1 < p1; p += 2) {
        const int c0 = ci[p], c1 = ci[p + 1];
        float a0[BS], a1[BS], x0[BS], x1[BS];
        load_row(p, a0);
        load_row(p + 1, a1);
        load_vec<BS, float>(x + (size_t)c0 * BS, x0);
        load_vec<BS, float>(x + (size_t)c1 * BS, x1);
        float s0 = 0.0f, s1 = 0.0f;
#pragma unroll
        for (int c = 0; c < BS; ++c) {
            s0 = fmaf(a0[c], x0[c], s0);
            s1 = fmaf(a1[c], x1[c], s1);
        }
        acc += s0;
        acc += s1;
    }
    if (p < p1) {
        const int c0 = ci[p];
        float a0[BS], x0[BS];
        load_row(p, a0);
        load_vec<BS, float>(x + (size_t)c0 * BS, x0);
        float s0 = 0.0f;
#pragma unroll
        for (int c = 0; c < BS; ++c) s0 = fmaf(a0[c], x0[c], s0);
        acc += s0;
    }
    y[(size_t)i * BS + r] = (double)acc;
}

// The same product over an ELL copy of the chunk-major FP32 values (W <= 8 slots per row, column
// -1 and zero blocks on padding): the column indices load from the row number, all at once, so
// the blocks and x entries wait on one load.  Real blocks in stored order, padding skipped: the
// results are the bits of k_bsmv_lp_perm.
template <int BS>
__global__ __launch_bounds__(kBlock) void k_bsmv_lp_ell(int n, int W, const int* __restrict__ cie,
                                                        const float* __restrict__ vale, const float* __restrict__ x,
                                                        double* __restrict__ y, int nwg) {
    static_assert(BS % 4 == 0, "chunk-major FP32 layout needs a block size divisible by 4");
    constexpr int RPW = kWave / BS;
    const int lane = threadIdx.x & (kWave - 1);
    const int per = (nwg + 7) / 8;
    const int wg = (blockIdx.x % 8) * per + blockIdx.x / 8;
    if (wg >= nwg) return;
    const int wave = (wg * kBlock + threadIdx.x) / kWave;
    const int g = lane / BS;
    const int r = lane - g * BS;
    if (g >= RPW) return;
    const int i = wave * RPW + g;
    if (i >= n) return;
    int col[8];
#pragma unroll
    for (int s = 0; s < 8; ++s) col[s] = s < W ? cie[(size_t)i * W + s] : -1;
    auto load_row = [&](size_t q, float (&a)[BS]) {
        const f4v* s4 = reinterpret_cast<const f4v*>(vale + q * BS * BS) + r;
#pragma unroll
        for (int k = 0; k < BS / 4; ++k) {
            const f4v t = __builtin_nontemporal_load(s4 + k * BS);
            a[4 * k] = t.x;
            a[4 * k + 1] = t.y;
            a[4 * k + 2] = t.z;
            a[4 * k + 3] = t.w;
        }
    };
    float acc = 0.0f;
    const size_t q0 = (size_t)i * W;
#pragma unroll
    for (int s = 0; s < 8; s += 2) {
        if (col[s] < 0) break;   // padding is at the end of a row
        float a0[BS], x0[BS];
        load_row(q0 + s, a0);
        load_vec<BS, float>(x + (size_t)col[s] * BS, x0);
        if (col[s + 1] >= 0) {
            float a1[BS], x1[BS];
            load_row(q0 + s + 1, a1);
            load_vec<BS, float>(x + (size_t)col[s + 1] * BS, x1);
            float s0 = 0.0f, s1 = 0.0f;
#pragma unroll
            for (int c = 0; c < BS; ++c) {
                s0 = fmaf(a0[c], x0[c], s0);
                s1 = fmaf(a1[c], x1[c], s1);
            }
            acc += s0;
            acc += s1;
        } else {
            float s0 = 0.0f;
#pragma unroll
            for (int c = 0; c < BS; ++c) s0 = fmaf(a0[c], x0[c], s0);
            acc += s0;
            break;
        }
    }
    y[(size_t)i * BS + r] = (double)acc;
}

// FP64 row-major blocks -> FP32 chunk-major blocks (the layout of k_bsmv_lp_perm)
__global__ __launch_bounds__(kBlock) void k_d2f_chunked(int64_t total, int bs, const double* __restrict__ val,
                                                        float* __restrict__ valp) {
    const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (t >= total) return;
    const int b2 = bs * bs;
    const int64_t p = t / b2;
    const int e = (int)(t - p * b2);
    const int r = e / bs, c = e - r * bs;
    valp[p * b2 + ((c / 4) * bs + r) * 4 + (c & 3)] = (float)val[t];
}

__global__ __launch_bounds__(kBlock) void k_d2f(int64_t n, const double* __restrict__ a, float* __restrict__ b) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) b[i] = (float)a[i];
}

__global__ __launch_bounds__(kBlock) void k_f2d(int64_t n, const float* __restrict__ a, double* __restrict__ b) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) b[i] = (double)a[i];
}

constexpr int kMaxBasis = 64;
struct FPtrs {
    const float* q[kMaxBasis];
};
struct FOut {
    float* q[kMaxBasis];
};

// x += sum_j c_j Z_j (FP64 accumulation of FP32 vectors), two entries per thread, four
// vectors' loads issued before their FMAs.
__global__ __launch_bounds__(kBlock) void k_combine_add_f32(int64_t len, int m, FPtrs Z, const double* __restrict__ c,
                                                            double* __restrict__ x, int vec) {
    const int64_t e = ((int64_t)blockIdx.x * kBlock + threadIdx.x) * 2;
    if (vec && e + 1 < len) {
        double2 acc = *reinterpret_cast<const double2*>(x + e);
        int j = 0;
        for (; j + 4 <= m; j += 4) {
            float2 v[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) v[u] = *reinterpret_cast<const float2*>(Z.q[j + u] + e);
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                acc.x += c[j + u] * (double)v[u].x;
                acc.y += c[j + u] * (double)v[u].y;
            }
        }
        for (; j < m; ++j) {
            const float2 v = *reinterpret_cast<const float2*>(Z.q[j] + e);
            acc.x += c[j] * (double)v.x;
            acc.y += c[j] * (double)v.y;
        }
        *reinterpret_cast<double2*>(x + e) = acc;
    } else {
        for (int64_t f = e; f < e + 2 && f < len; ++f) {
            double acc = x[f];
            for (int j = 0; j < m; ++j) acc += c[j] * (double)Z.q[j][f];
            x[f] = acc;
        }
    }
}

// The same for kp = JN <= 16 outputs: eight input vectors' loads issued before their FMAs, P
// read with uniform (scalar) loads, the sums in k_basis_rotate_f32's order; len even, vectors
// 8-byte aligned.
template <int JN>
__global__ __launch_bounds__(kBlock) void k_basis_rotate_f32_jn(int64_t len, int m, FPtrs Z,
                                                                const double* __restrict__ P, FOut Out) {
    const int64_t e = ((int64_t)blockIdx.x * kBlock + threadIdx.x) * 2;
    if (e >= len) return;
    double a0[JN], a1[JN];
#pragma unroll
    for (int j = 0; j < JN; ++j) a0[j] = a1[j] = 0.0;
    int i = 0;
    for (; i + 8 <= m; i += 8) {
        float2 v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = *reinterpret_cast<const float2*>(Z.q[i + u] + e);
#pragma unroll
        for (int u = 0; u < 8; ++u)
#pragma unroll
            for (int j = 0; j < JN; ++j) {
                const double pv = P[(i + u) * JN + j];
                a0[j] = fma((double)v[u].x, pv, a0[j]);
                a1[j] = fma((double)v[u].y, pv, a1[j]);
            }
    }
    for (; i < m; ++i) {
        const float2 v = *reinterpret_cast<const float2*>(Z.q[i] + e);
#pragma unroll
        for (int j = 0; j < JN; ++j) {
            const double pv = P[i * JN + j];
            a0[j] = fma((double)v.x, pv, a0[j]);
            a1[j] = fma((double)v.y, pv, a1[j]);
        }
    }
#pragma unroll
    for (int j = 0; j < JN; ++j) *reinterpret_cast<float2*>(Out.q[j] + e) = make_float2((float)a0[j], (float)a1[j]);
}

// Zout_j = sum_i Z_i P_ij (FP64 accumulation, FP32 storage), m, kp <= 64; two entries per
// thread, four input vectors' loads issued before their FMAs.
__global__ __launch_bounds__(kBlock) void k_basis_rotate_f32(int64_t len, int m, int kp, FPtrs Z,
                                                             const double* __restrict__ P, FOut Out, int vec) {
    __shared__ double sP[kMaxBasis * 16];
    const int64_t e = ((int64_t)blockIdx.x * kBlock + threadIdx.x) * 2;
    for (int j0 = 0; j0 < kp; j0 += 16) {
        const int jn = min(16, kp - j0);
        __syncthreads();
        for (int t = threadIdx.x; t < m * jn; t += kBlock) sP[t] = P[(t / jn) * kp + j0 + (t % jn)];
        __syncthreads();
        if (vec && e + 1 < len) {
            double a0[16], a1[16];
#pragma unroll
            for (int j = 0; j < 16; ++j) a0[j] = a1[j] = 0.0;
            int i = 0;
            for (; i + 4 <= m; i += 4) {
                float2 v[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) v[u] = *reinterpret_cast<const float2*>(Z.q[i + u] + e);
#pragma unroll
                for (int u = 0; u < 4; ++u)
#pragma unroll
                    for (int j = 0; j < 16; ++j)
                        if (j < jn) {
                            a0[j] = fma((double)v[u].x, sP[(i + u) * jn + j], a0[j]);
                            a1[j] = fma((double)v[u].y, sP[(i + u) * jn + j], a1[j]);
                        }
            }
            for (; i < m; ++i) {
                const float2 v = *reinterpret_cast<const float2*>(Z.q[i] + e);
#pragma unroll
                for (int j = 0; j < 16; ++j)
                    if (j < jn) {
                        a0[j] = fma((double)v.x, sP[i * jn + j], a0[j]);
                        a1[j] = fma((double)v.y, sP[i * jn + j], a1[j]);
                    }
            }
#pragma unroll
            for (int j = 0; j < 16; ++j)
                if (j < jn) *reinterpret_cast<float2*>(Out.q[j0 + j] + e) = make_float2((float)a0[j], (float)a1[j]);
        } else {
            for (int64_t f = e; f < e + 2 && f < len; ++f) {
                double acc[16];
#pragma unroll
                for (int j = 0; j < 16; ++j) acc[j] = 0.0;
                for (int i = 0; i < m; ++i) {
                    const double zv = (double)Z.q[i][f];
#pragma unroll
                    for (int j = 0; j < 16; ++j)
                        if (j < jn) acc[j] = fma(zv, sP[i * jn + j], acc[j]);
                }
#pragma unroll
                for (int j = 0; j < 16; ++j)
                    if (j < jn) Out.q[j0 + j][f] = (float)acc[j];
            }
        }
    }
}

inline unsigned grid_for(int64_t threads) { return (unsigned)((threads + kBlock - 1) / kBlock); }

template <typename AT, typename XT, typename YT>
int bsmv_lp_dispatch(const bsp_ctx* ctx, int n, int bs, const int* rp, const int* ci, const AT* val, const XT* x, YT* y) {
    auto launch = [&](auto bsc) {
        constexpr int BS = decltype(bsc)::value;
        constexpr int RPW = kWave / BS;
        const int64_t waves = (n + RPW - 1) / RPW;
        hipLaunchKernelGGL((k_bsmv_lp<BS, AT, XT, YT>), dim3(grid_for(waves * kWave)), dim3(kBlock), 0, ctx->stream, n,
                           rp, ci, val, x, y);
    };
    switch (bs) {
        case 1: launch(std::integral_constant<int, 1>{}); break;
        case 2: launch(std::integral_constant<int, 2>{}); break;
        case 3: launch(std::integral_constant<int, 3>{}); break;
        case 4: launch(std::integral_constant<int, 4>{}); break;
        case 5: launch(std::integral_constant<int, 5>{}); break;
        case 6: launch(std::integral_constant<int, 6>{}); break;
        case 7: launch(std::integral_constant<int, 7>{}); break;
        case 8: launch(std::integral_constant<int, 8>{}); break;
        case 16: launch(std::integral_constant<int, 16>{}); break;
        case 32: launch(std::integral_constant<int, 32>{}); break;
        default: return fail(ctx, BSP_E_UNSUPPORTED, "reduced-precision bsmv: unsupported block size");
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(ctx, BSP_E_HIP, std::string("bsmv_lp launch: ") + hipGetErrorString(e));
    return BSP_OK;
}

}  // namespace

int launch_bsmv_f32(const bsp_ctx* ctx, int n, int bs, const int* rp, const int* ci, const float* val, const float* x,
                    double* y) {
    if (n <= 0) return BSP_OK;
    return bsmv_lp_dispatch<float, float, double>(ctx, n, bs, rp, ci, val, x, y);
}

bool bsmv_f32_chunked_supported(int bs) { return bs == 4 || bs == 8 || bs == 16; }

int launch_d2f_chunked(const bsp_ctx* ctx, int64_t nnzb, int bs, const double* val, float* valp) {
    const int64_t total = nnzb * bs * bs;
    if (total <= 0) return BSP_OK;
    hipLaunchKernelGGL(k_d2f_chunked, dim3(grid_for(total)), dim3(kBlock), 0, ctx->stream, total, bs, val, valp);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? BSP_OK : fail(ctx, BSP_E_HIP, hipGetErrorString(e));
}

int launch_bsmv_f32_ell(const bsp_ctx* ctx, int n, int bs, int W, const int* cie, const float* vale, const float* x,
                        double* y) {
    if (n <= 0) return BSP_OK;
    auto launch = [&](auto bsc) {
        constexpr int BS = decltype(bsc)::value;
        constexpr int RPW = kWave / BS;
        const int64_t waves = (n + RPW - 1) / RPW;
        const int nwg = (int)grid_for(waves * kWave);
        const int grid = (nwg + 7) / 8 * 8;
        hipLaunchKernelGGL((k_bsmv_lp_ell<BS>), dim3(grid), dim3(kBlock), 0, ctx->stream, n, W, cie, vale, x, y, nwg);
    };
    switch (bs) {
        case 4: launch(std::integral_constant<int, 4>{}); break;
        case 8: launch(std::integral_constant<int, 8>{}); break;
        case 16: launch(std::integral_constant<int, 16>{}); break;
        default: return fail(ctx, BSP_E_UNSUPPORTED, "ELL FP32 bsmv: block size 4, 8 or 16");
    }
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? BSP_OK : fail(ctx, BSP_E_HIP, std::string("bsmv_lp_ell launch: ") + hipGetErrorString(e));
}

int launch_bsmv_f32_chunked(const bsp_ctx* ctx, int n, int bs, const int* rp, const int* ci, const float* valp,
                            const float* x, double* y) {
    if (n <= 0) return BSP_OK;
    auto launch = [&](auto bsc) {
        constexpr int BS = decltype(bsc)::value;
        constexpr int RPW = kWave / BS;
        const int64_t waves = (n + RPW - 1) / RPW;
        const int nwg = (int)grid_for(waves * kWave);
        const int grid = (nwg + 7) / 8 * 8;
        hipLaunchKernelGGL((k_bsmv_lp_perm<BS>), dim3(grid), dim3(kBlock), 0, ctx->stream, n, rp, ci, valp, x, y, nwg);
    };
    switch (bs) {
        case 4: launch(std::integral_constant<int, 4>{}); break;
        case 8: launch(std::integral_constant<int, 8>{}); break;
        case 16: launch(std::integral_constant<int, 16>{}); break;
        default: return fail(ctx, BSP_E_UNSUPPORTED, "chunked FP32 bsmv: block size must be 4, 8 or 16");
    }
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? BSP_OK : fail(ctx, BSP_E_HIP, std::string("bsmv_lp_perm launch: ") + hipGetErrorString(e));
}

int launch_bsmv_f32f(const bsp_ctx* ctx, int n, int bs, const int* rp, const int* ci, const float* val, const float* x,
                     float* y) {
    if (n <= 0) return BSP_OK;
    return bsmv_lp_dispatch<float, float, float>(ctx, n, bs, rp, ci, val, x, y);
}

int launch_d2f(const bsp_ctx* ctx, int64_t n, const double* a, float* b) {
    if (n <= 0) return BSP_OK;
    hipLaunchKernelGGL(k_d2f, dim3(grid_for(n)), dim3(kBlock), 0, ctx->stream, n, a, b);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? BSP_OK : fail(ctx, BSP_E_HIP, hipGetErrorString(e));
}

int launch_f2d(const bsp_ctx* ctx, int64_t n, const float* a, double* b) {
    if (n <= 0) return BSP_OK;
    hipLaunchKernelGGL(k_f2d, dim3(grid_for(n)), dim3(kBlock), 0, ctx->stream, n, a, b);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? BSP_OK : fail(ctx, BSP_E_HIP, hipGetErrorString(e));
}

int launch_combine_add_f32(const bsp_ctx* ctx, int64_t len, int m, const float* const* Z, const double* c, double* x) {
    for (int j0 = 0; j0 < m; j0 += kMaxBasis) {
        const int mm = std::min(kMaxBasis, m - j0);
        FPtrs t{};
        bool vec = (uintptr_t)x % 16 == 0;
        for (int j = 0; j < mm; ++j) {
            t.q[j] = Z[j0 + j];
            vec = vec && (uintptr_t)Z[j0 + j] % 8 == 0;
        }
        hipLaunchKernelGGL(k_combine_add_f32, dim3(grid_for((len + 1) / 2)), dim3(kBlock), 0, ctx->stream, len, mm, t,
                           c + j0, x, vec ? 1 : 0);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return fail(ctx, BSP_E_HIP, hipGetErrorString(e));
    }
    return BSP_OK;
}

int launch_basis_rotate_f32(const bsp_ctx* ctx, int64_t len, int m, int kp, const float* const* Z, const double* P,
                            float* const* Zout) {
    if (m > kMaxBasis || kp > kMaxBasis) return fail(ctx, BSP_E_UNSUPPORTED, "basis larger than 64 vectors");
    FPtrs t{};
    FOut o{};
    bool vec = true;
    for (int i = 0; i < m; ++i) {
        t.q[i] = Z[i];
        vec = vec && (uintptr_t)Z[i] % 8 == 0;
    }
    for (int j = 0; j < kp; ++j) {
        o.q[j] = Zout[j];
        vec = vec && (uintptr_t)Zout[j] % 8 == 0;
    }
    if (vec && len % 2 == 0 && kp >= 1 && kp <= 16) {
        bool done = false;
        auto go = [&](auto jn) {
            constexpr int JN = decltype(jn)::value;
            if (JN != kp) return;
            hipLaunchKernelGGL(k_basis_rotate_f32_jn<JN>, dim3(grid_for(len / 2)), dim3(kBlock), 0, ctx->stream, len,
                               m, t, P, o);
            done = true;
        };
        [&]<int... J>(std::integer_sequence<int, J...>) { (go(std::integral_constant<int, J + 1>{}), ...); }(
            std::make_integer_sequence<int, 16>{});
        if (done) {
            hipError_t e = hipGetLastError();
            return e == hipSuccess ? BSP_OK : fail(ctx, BSP_E_HIP, hipGetErrorString(e));
        }
    }
    hipLaunchKernelGGL(k_basis_rotate_f32, dim3(grid_for((len + 1) / 2)), dim3(kBlock), 0, ctx->stream, len, m, kp, t,
                       P, o, vec ? 1 : 0);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? BSP_OK : fail(ctx, BSP_E_HIP, hipGetErrorString(e));
}

}  // namespace bsp
