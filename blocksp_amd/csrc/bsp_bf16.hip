// bsp_bf16.hip -- BF16 block SpMM on the matrix cores (gfx950): Y = A X for many
// right-hand sides, 8 x 8 blocks, BF16 operands, FP32 accumulation.
//
// Layouts ("block-major"): Xb[(col * nrhs + k) * 8 + c] (BF16) and Yb[(row * nrhs + k) * 8 + r]
// (FP32): for each block row, the right-hand sides one after another, each with its 8
// components contiguous.  With that layout every MFMA operand a lane needs -- the 8
// components of one right-hand side, or one row of a block -- is one 16-byte load, and every
// result a lane holds (4 rows for one right-hand side) is one 16-byte store.
//
// The matrix is read from a BF16 copy of its values made on the device at first use: an ELL copy
// (W slots per row, zero blocks and column -1 on padding) when no row holds more than 8 blocks and
// padding stays within a quarter of the blocks, the values in CSR order otherwise.  Measured on
// the 64^3 DG Laplacian (262144 block rows, 7 blocks per row) with 32 right-hand sides:
//   * one wavefront per row pair, v_mfma_f32_16x16x16_bf16 with one slot of each row per MFMA,
//     8-byte operand loads, CSR: 175-178 us;
//   * the same with v_mfma_f32_16x16x32_bf16 (two slots of each row, 16-byte loads): 178-180 us;
//   * 16x16x32 over the ELL copy (no row-pointer load before the column indices): 146 us;
//   * the ELL kernel persistent with the next pair's column indices loaded ahead: 136-140 us
//     (3 workgroups of 4 wavefronts per CU; 142 us at 5, 156 us at 2).
// Measured and dropped: chunks of rows taken in turn from slabs of an XCD's range (keeps the
// z-neighbour right-hand sides in L2; 143-149 us), two pairs in flight per wavefront with the
// next pair's operands loaded before the current pair's MFMAs (130 VGPRs, 157-188 us), two
// wavefronts per pair splitting the right-hand sides (161-177 us), and cached instead of streaming
// matrix loads or result stores (141-153 us).
#include "bsp_internal.hpp"

#include <algorithm>

namespace bsp {

namespace {

typedef float v4f __attribute__((ext_vector_type(4)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef unsigned v4u __attribute__((ext_vector_type(4)));

constexpr int kBlock = 256;
constexpr int kXcds = 8;
constexpr int kWgPerCu = 3;       // persistent kernel: workgroups per CU

__device__ __forceinline__ uint16_t f32_to_bf16(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) return (uint16_t)((u >> 16) | 0x40);   // quiet NaN
    u += 0x7fffu + ((u >> 16) & 1u);   // round to nearest, ties to even
    return (uint16_t)(u >> 16);
}

__device__ __forceinline__ void store_y(float* p, v4f v) { __builtin_nontemporal_store(v, reinterpret_cast<v4f*>(p)); }

__device__ __forceinline__ v8bf load16(const uint16_t* p) { return __builtin_bit_cast(v8bf, *reinterpret_cast<const v4u*>(p)); }
__device__ __forceinline__ v8bf load16_nt(const uint16_t* p) {
    return __builtin_bit_cast(v8bf, __builtin_nontemporal_load(reinterpret_cast<const v4u*>(p)));
}

// Yb = A Xb with v_mfma_f32_16x16x32_bf16: one wavefront per pair of block rows (i0, i1 = i0 + 1),
// M = 16 output rows (8 of each), N = 16 right-hand sides, K = 32 = two slots of each row.  Lane
// (m = lane % 16, g = lane / 16) supplies row m of the A operand and column m of the B operand over
// K entries 8g .. 8g + 7; those entries belong to row i0 (g < 2) or i1 (g >= 2), slot s + g % 2:
//     A operand [ A(i0, s)  A(i0, s+1)      0           0      ]     B operand  X(col(i0, s))
//               [    0          0       A(i1, s)  A(i1, s+1)   ]                X(col(i0, s+1))
//                                                                               X(col(i1, s))
//                                                                               X(col(i1, s+1))
// so every operand a lane holds is one row of a block or one right-hand side of a block of X:
// 8 contiguous BF16, one 16-byte load.  The sum over K does not depend on how the hardware
// orders K inside a lane group, as long as the A and B operands use the same order, which they do.
// Result: lane (m, g) holds rows 4g .. 4g + 3 of the 16 for right-hand side m, one 16-byte store.
//
// ELL = true reads the ELL copy: the column indices and blocks of a row are found from the row
// number alone, so the loads of a row pair form a chain of two (columns, then blocks and
// right-hand sides) instead of three (row pointers first).  All slots of a row (up to 2 NP per
// round) are loaded before the first MFMA.  Matrix blocks are read and results written with the
// streaming (non-temporal) policy, leaving the L2 to the reused right-hand sides; workgroups are
// remapped so that each XCD processes a contiguous range of block rows.
template <int NP, bool ELL>
__global__ __launch_bounds__(kBlock) void k_bsmm_bf16_k32(int n, int nrhs, int W, const int* __restrict__ rp,
                                                          const int* __restrict__ ci, const uint16_t* __restrict__ A,
                                                          const uint16_t* __restrict__ Xb, float* __restrict__ Yb,
                                                          int nwg) {
    const int w = blockIdx.x;
    const int per = (nwg + kXcds - 1) / kXcds;
    const int lw = (w % kXcds) * per + w / kXcds;
    if (lw >= nwg) return;
    const int lane = threadIdx.x & 63;
    const int pair = lw * (kBlock / 64) + (threadIdx.x >> 6);
    const int i0 = 2 * pair;
    if (i0 >= n) return;
    const int m = lane & 15, g = lane >> 4;
    const bool top = g < 2;
    const int row = top ? i0 : i0 + 1;
    const bool row_ok = row < n;
    const bool a_live = top == (m < 8);
    const int arow = m & 7, par = g & 1;
    int64_t base;
    int len, len_pair;
    if (ELL) {
        base = (int64_t)row * W;
        len = row_ok ? W : 0;
        len_pair = W;
    } else {
        const int p0 = rp[i0], p1 = rp[i0 + 1], p2 = i0 + 1 < n ? rp[i0 + 2] : p1;
        base = top ? p0 : p1;
        len = top ? p1 - p0 : p2 - p1;
        len_pair = max(p1 - p0, p2 - p1);
    }
    const int ntiles = nrhs / 16;
    for (int t0 = 0; t0 < ntiles; t0 += 2) {
        const bool two = t0 + 1 < ntiles;
        v4f acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = {0.f, 0.f, 0.f, 0.f};
        for (int s0 = 0; s0 < len_pair; s0 += 2 * NP) {
            int col[NP];
#pragma unroll
            for (int q = 0; q < NP; ++q) {
                const int slot = s0 + 2 * q + par;
                col[q] = slot < len ? ci[base + slot] : -1;
            }
            v8bf a[NP], b0[NP], b1[NP];
#pragma unroll
            for (int q = 0; q < NP; ++q) {
                const int slot = s0 + 2 * q + par;
                a[q] = v8bf{};
                b0[q] = v8bf{};
                b1[q] = v8bf{};
                if (col[q] >= 0) {
                    if (a_live) a[q] = load16_nt(A + (size_t)(base + slot) * 64 + arow * 8);
                    const uint16_t* xb = Xb + ((size_t)col[q] * nrhs + t0 * 16 + m) * 8;
                    b0[q] = load16(xb);
                    if (two) b1[q] = load16(xb + 16 * 8);
                }
            }
            const int qn = min(NP, (len_pair - s0 + 1) / 2);
#pragma unroll
            for (int q = 0; q < NP; ++q) {
                if (q < qn) {
                    acc0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[q], b0[q], acc0, 0, 0, 0);
                    if (two) acc1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[q], b1[q], acc1, 0, 0, 0);
                }
            }
        }
        if (row_ok) {
            float* y = Yb + ((size_t)row * nrhs + t0 * 16 + m) * 8 + 4 * par;
            store_y(y, acc0);
            if (two) store_y(y + 16 * 8, acc1);
        }
    }
}

// Persistent form of the ELL kernel (nrhs <= 32, W <= 8): a fixed grid of wavefronts, each walking
// pairs of block rows of its XCD's contiguous range (pair = range start + wave + k * wavefronts of
// the XCD, so the XCD's wavefronts move through the range together and share the right-hand-side
// blocks in its L2).  The column indices of the next pair are loaded while the current pair's
// blocks and right-hand sides are in flight, so each pair waits on one load latency instead of two.
__global__ __launch_bounds__(kBlock) void k_bsmm_bf16_ell_persist(int n, int nrhs, int W, const int* __restrict__ ci,
                                                                  const uint16_t* __restrict__ A,
                                                                  const uint16_t* __restrict__ Xb,
                                                                  float* __restrict__ Yb, int wg_per_xcd) {
    constexpr int NP = 4;
    const int xcd = blockIdx.x % kXcds;
    const int nwx = wg_per_xcd * (kBlock / 64);                      // wavefronts of this XCD
    const int lwave = (blockIdx.x / kXcds) * (kBlock / 64) + (threadIdx.x >> 6);
    const int pairs = (n + 1) / 2;
    const int per = (pairs + kXcds - 1) / kXcds;
    const int pend = min(pairs, (xcd + 1) * per);
    int pair = xcd * per + lwave;
    if (pair >= pend) return;
    const int lane = threadIdx.x & 63;
    const int m = lane & 15, g = lane >> 4;
    const bool top = g < 2;
    const bool a_live = top == (m < 8);
    const int arow = m & 7, par = g & 1;
    const bool two = nrhs > 16;
    const int nslot = min(W, 2 * NP);
    int col[NP];
    {
        const int row = 2 * pair + (top ? 0 : 1);
#pragma unroll
        for (int q = 0; q < NP; ++q) {
            const int slot = 2 * q + par;
            col[q] = (row < n && slot < nslot) ? ci[(size_t)row * W + slot] : -1;
        }
    }
    while (true) {
        const int row = 2 * pair + (top ? 0 : 1);
        const int64_t base = (int64_t)row * W;
        v8bf a[NP], b0[NP], b1[NP];
#pragma unroll
        for (int q = 0; q < NP; ++q) {
            const int slot = 2 * q + par;
            a[q] = v8bf{};
            b0[q] = v8bf{};
            b1[q] = v8bf{};
            if (col[q] >= 0) {
                if (a_live) a[q] = load16_nt(A + (size_t)(base + slot) * 64 + arow * 8);
                const uint16_t* xb = Xb + ((size_t)col[q] * nrhs + m) * 8;
                b0[q] = load16(xb);
                if (two) b1[q] = load16(xb + 16 * 8);
            }
        }
        const int next = pair + nwx;
        const bool more = next < pend;
        if (more) {
            const int nrow = 2 * next + (top ? 0 : 1);
#pragma unroll
            for (int q = 0; q < NP; ++q) {
                const int slot = 2 * q + par;
                col[q] = (nrow < n && slot < nslot) ? ci[(size_t)nrow * W + slot] : -1;
            }
        }
        v4f acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int q = 0; q < NP; ++q) {
            if (2 * q < nslot) {
                acc0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[q], b0[q], acc0, 0, 0, 0);
                if (two) acc1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[q], b1[q], acc1, 0, 0, 0);
            }
        }
        if (row < n) {
            float* y = Yb + ((size_t)row * nrhs + m) * 8 + 4 * par;
            store_y(y, acc0);
            if (two) store_y(y + 16 * 8, acc1);
        }
        if (!more) break;
        pair = next;
    }
}

// ELL copy of the BF16 values: W slots per row, zero blocks and column -1 on padding.
__global__ void k_ell_bf16(int64_t total, int W, const int* __restrict__ rp, const int* __restrict__ ci,
                           const double* __restrict__ v, uint16_t* __restrict__ out, int* __restrict__ ci_ell) {
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;   // (row, slot, entry)
    if (t >= total) return;
    const int64_t e = t & 63, rs = t >> 6;
    const int64_t row = rs / W, slot = rs % W;
    const int p0 = rp[row], len = rp[row + 1] - p0;
    const bool live = slot < len;
    out[t] = live ? f32_to_bf16((float)v[(size_t)(p0 + slot) * 64 + e]) : (uint16_t)0;
    if (e == 0) ci_ell[rs] = live ? ci[p0 + slot] : -1;
}

__global__ void k_values_bf16(int64_t cnt, const double* __restrict__ v, uint16_t* __restrict__ out) {
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t < cnt) out[t] = f32_to_bf16((float)v[t]);
}

// RHS-minor X[(row * bs + c) * nrhs + k] (double) -> block-major BF16 Xb[(row * nrhs + k) * bs + c]
__global__ void k_pack_bf16(int64_t total, int bs, int nrhs, const double* __restrict__ X, uint16_t* __restrict__ Xb) {
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= total) return;
    const int64_t c = t % bs, rk = t / bs;
    const int64_t k = rk % nrhs, row = rk / nrhs;
    Xb[t] = f32_to_bf16((float)X[(row * bs + c) * nrhs + k]);
}

// block-major FP32 Yb -> RHS-minor double Y
__global__ void k_unpack_f32(int64_t total, int bs, int nrhs, const float* __restrict__ Yb, double* __restrict__ Y) {
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= total) return;
    const int64_t c = t % bs, rk = t / bs;
    const int64_t k = rk % nrhs, row = rk / nrhs;
    Y[(row * bs + c) * nrhs + k] = (double)Yb[t];
}

inline unsigned blocks_for(int64_t work) { return (unsigned)((work + kBlock - 1) / kBlock); }

}  // namespace

int launch_bsmm_bf16(const bsp_ctx* ctx, int n, int nrhs, int W, const int* rp, const int* ci, const uint16_t* A,
                     const uint16_t* Xb, float* Yb) {
    if (n <= 0) return BSP_OK;
    const int pairs = (n + 1) / 2;
    const int nwg = (pairs + kBlock / 64 - 1) / (kBlock / 64);
    const int grid = (nwg + kXcds - 1) / kXcds * kXcds;
    if (W == 0)
        hipLaunchKernelGGL((k_bsmm_bf16_k32<4, false>), dim3(grid), dim3(kBlock), 0, ctx->stream, n, nrhs, 0, rp, ci,
                           A, Xb, Yb, nwg);
    else if (nrhs > 32 || W > 8)
        hipLaunchKernelGGL((k_bsmm_bf16_k32<4, true>), dim3(grid), dim3(kBlock), 0, ctx->stream, n, nrhs, W, rp, ci,
                           A, Xb, Yb, nwg);
    else {
        const int wpx = std::max(1, std::min(ctx->num_cus / kXcds * kWgPerCu, (nwg + kXcds - 1) / kXcds));
        hipLaunchKernelGGL(k_bsmm_bf16_ell_persist, dim3(wpx * kXcds), dim3(kBlock), 0, ctx->stream, n, nrhs, W, ci, A,
                           Xb, Yb, wpx);
    }
    BSP_HIP(hipGetLastError());
    return BSP_OK;
}

int launch_ell_bf16(const bsp_ctx* ctx, int n, int W, const int* rp, const int* ci, const double* v, uint16_t* out,
                    int* ci_ell) {
    const int64_t total = (int64_t)n * W * 64;
    if (total <= 0) return BSP_OK;
    hipLaunchKernelGGL(k_ell_bf16, dim3(blocks_for(total)), dim3(kBlock), 0, ctx->stream, total, W, rp, ci, v, out,
                       ci_ell);
    BSP_HIP(hipGetLastError());
    return BSP_OK;
}

int launch_values_bf16(const bsp_ctx* ctx, int64_t cnt, const double* v, uint16_t* out) {
    if (cnt <= 0) return BSP_OK;
    hipLaunchKernelGGL(k_values_bf16, dim3(blocks_for(cnt)), dim3(kBlock), 0, ctx->stream, cnt, v, out);
    BSP_HIP(hipGetLastError());
    return BSP_OK;
}

int launch_pack_bf16(const bsp_ctx* ctx, int64_t rows, int bs, int nrhs, const double* X, uint16_t* Xb) {
    const int64_t total = rows * bs * nrhs;
    if (total <= 0) return BSP_OK;
    hipLaunchKernelGGL(k_pack_bf16, dim3(blocks_for(total)), dim3(kBlock), 0, ctx->stream, total, bs, nrhs, X, Xb);
    BSP_HIP(hipGetLastError());
    return BSP_OK;
}

int launch_unpack_f32(const bsp_ctx* ctx, int64_t rows, int bs, int nrhs, const float* Yb, double* Y) {
    const int64_t total = rows * bs * nrhs;
    if (total <= 0) return BSP_OK;
    hipLaunchKernelGGL(k_unpack_f32, dim3(blocks_for(total)), dim3(kBlock), 0, ctx->stream, total, bs, nrhs, Yb, Y);
    BSP_HIP(hipGetLastError());
    return BSP_OK;
}

}  // namespace bsp

using namespace bsp;

extern "C" {

int bsp_rhs_pack_bf16(bsp_ctx* ctx, int64_t rows, int bs, int nrhs, const double* X_dev, uint16_t* Xb_dev) {
    if (!ctx || (!X_dev && rows > 0) || (!Xb_dev && rows > 0) || rows < 0 || bs <= 0 || nrhs <= 0)
        return BSP_E_INVALID;
    BSP_HIP(hipSetDevice(ctx->device));
    return launch_pack_bf16(ctx, rows, bs, nrhs, X_dev, Xb_dev);
}

int bsp_rhs_unpack_f32(bsp_ctx* ctx, int64_t rows, int bs, int nrhs, const float* Yb_dev, double* Y_dev) {
    if (!ctx || (!Yb_dev && rows > 0) || (!Y_dev && rows > 0) || rows < 0 || bs <= 0 || nrhs <= 0)
        return BSP_E_INVALID;
    BSP_HIP(hipSetDevice(ctx->device));
    return launch_unpack_f32(ctx, rows, bs, nrhs, Yb_dev, Y_dev);
}

int bsp_bsmm_rhs_bf16(bsp_ctx* ctx, const bsp_bsr* A, int nrhs, const uint16_t* Xb_dev, float* Yb_dev) {
    if (!ctx || !A || !Xb_dev || !Yb_dev) return BSP_E_INVALID;
    BSP_CHECK(A->cbs == 0, BSP_E_UNSUPPORTED, "the BF16 block SpMM does not take complex matrices");
    BSP_CHECK(A->bs == 8, BSP_E_UNSUPPORTED, "the BF16 MFMA block SpMM needs 8 x 8 blocks");
    BSP_CHECK(nrhs > 0 && nrhs % 16 == 0, BSP_E_INVALID, "the BF16 MFMA block SpMM needs a multiple of 16 RHS");
    BSP_HIP(hipSetDevice(ctx->device));
    if (A->val16_w < 0) {
        // ELL copy when no row holds more than 8 blocks and padding stays within a quarter of the blocks
        int W = 0;
        for (int i = 0; i < A->n_rows; ++i) W = std::max(W, A->h_rp[i + 1] - A->h_rp[i]);
        W = std::max(2, (W + 1) / 2 * 2);
        A->val16_w = W <= 8 && (int64_t)W * A->n_rows <= A->nnzb + A->nnzb / 4 ? W : 0;
    }
    const bool ell = ctx->bf16_ell && A->val16_w > 0;
    if (A->val16.p && ell != A->val16_ell) {   // made under the other setting of the bf16_ell option
        A->val16.release();
        A->ci16.release();
    }
    if (!A->val16.p) {
        if (ell) {
            const int64_t slots = (int64_t)A->n_rows * A->val16_w;
            BSP_HIP(A->val16.alloc(sizeof(uint16_t) * 64 * std::max<int64_t>(slots, 1)));
            BSP_HIP(A->ci16.alloc(sizeof(int) * std::max<int64_t>(slots, 1)));
            int rc = launch_ell_bf16(ctx, A->n_rows, A->val16_w, A->rp.as<int>(), A->ci.as<int>(), A->val.as<double>(),
                                     A->val16.as<uint16_t>(), A->ci16.as<int>());
            if (rc) return rc;
        } else {
            const int64_t cnt = A->nnzb * 64;
            BSP_HIP(A->val16.alloc(sizeof(uint16_t) * std::max<int64_t>(cnt, 1)));
            int rc = launch_values_bf16(ctx, cnt, A->val.as<double>(), A->val16.as<uint16_t>());
            if (rc) return rc;
        }
        A->val16_ell = ell;
    }
    if (ell)
        return launch_bsmm_bf16(ctx, A->n_rows, nrhs, A->val16_w, nullptr, A->ci16.as<int>(), A->val16.as<uint16_t>(),
                                Xb_dev, Yb_dev);
    return launch_bsmm_bf16(ctx, A->n_rows, nrhs, 0, A->rp.as<int>(), A->ci.as<int>(), A->val16.as<uint16_t>(), Xb_dev,
                            Yb_dev);
}

}  // extern "C"
