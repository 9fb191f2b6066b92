// bsp_bf16.hip -- BF16 block SpMM on the matrix cores (gfx950): Y = A X for many
// right-hand sides, 8 x 8 blocks, BF16 operands, FP32 accumulation.
//
// Layouts ("block-major"): Xb[(col * nrhs + k) * 8 + c] (BF16) and Yb[(row * nrhs + k) * 8 + r]
// (FP32): for each block row, the right-hand sides one after another, each with its 8
// components contiguous.  With that layout every MFMA operand a lane needs -- four
// consecutive components of one right-hand side, or four consecutive entries of one
// row of a block -- is one 8-byte load, and every result a lane holds is one 16-byte
// store.
//
// Mapping: one wavefront computes two block rows (i0, i1) for 16 right-hand sides at a
// time with v_mfma_f32_16x16x16_bf16 (M = 16 output rows = 8 of i0 + 8 of i1,
// N = 16 right-hand sides, K = 16 = the 8 columns of i0's block at slot q and the 8
// columns of i1's block at slot q).  The A operand is block diagonal
//     [ A(i0, q)    0      ]
//     [   0      A(i1, q)  ]
// and the B operand stacks X(col(i0, q)) over X(col(i1, q)).  Half the A operand is
// zero, which costs MFMA issue slots only: the kernel is bound by HBM (BF16 blocks,
// BF16 right-hand sides, FP32 results), not by the matrix cores.  Workgroups are
// remapped so that each XCD processes a contiguous range of block rows, which keeps
// the neighbouring right-hand-side blocks in that XCD's L2.
#include "bsp_internal.hpp"

#include <algorithm>

namespace bsp {

namespace {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int kBlock = 256;
constexpr int kXcds = 8;

__device__ __forceinline__ uint16_t f32_to_bf16(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu)) return (uint16_t)((u >> 16) | 0x40);   // quiet NaN
    u += 0x7fffu + ((u >> 16) & 1u);   // round to nearest, ties to even
    return (uint16_t)(u >> 16);
}

__device__ __forceinline__ v4s load_a(const uint16_t* p) {
    const uint64_t u = __builtin_nontemporal_load(reinterpret_cast<const uint64_t*>(p));
    return __builtin_bit_cast(v4s, u);
}

__device__ __forceinline__ void store_y(float* p, v4f v) { __builtin_nontemporal_store(v, reinterpret_cast<v4f*>(p)); }

// Yb = A Xb, nrhs a multiple of 16, with every load of a row pair issued before the first
// MFMA: the block indices of up to Q slots, then all their matrix and right-hand-side operands,
// so a wavefront keeps ~3Q loads in flight instead of waiting on one slot at a time (rows with
// more than Q blocks continue Q slots at a time).  The matrix blocks are read and the results
// written with the streaming (non-temporal) policy, leaving the L2 to the reused right-hand
// sides.  Measured on the 64^3 DG matrix with 32 right-hand sides: Q = 4 176 us; Q = 8 182 us;
// slot by slot 193 us (streaming) and 211 us (cached).
template <int Q>
__global__ __launch_bounds__(kBlock) void k_bsmm_bf16_pf(int n, int nrhs, const int* __restrict__ rp,
                                                         const int* __restrict__ ci, const uint16_t* __restrict__ A,
                                                         const uint16_t* __restrict__ Xb, float* __restrict__ Yb,
                                                         int nwg) {
    const int w = blockIdx.x;
    const int per = (nwg + kXcds - 1) / kXcds;
    const int lw = (w % kXcds) * per + w / kXcds;
    if (lw >= nwg) return;
    const int lane = threadIdx.x & 63;
    const int pair = lw * (kBlock / 64) + (threadIdx.x >> 6);
    const int i0 = 2 * pair, i1 = i0 + 1;
    if (i0 >= n) return;
    const int p00 = rp[i0], len0 = rp[i0 + 1] - p00;
    const int p10 = i1 < n ? rp[i1] : 0, len1 = i1 < n ? rp[i1 + 1] - p10 : 0;
    const int m = lane & 15, kg = lane >> 4;
    const bool top = kg < 2;
    const int kk = (kg & 1) * 4;
    const bool a_live = top ? (m < 8) : (m >= 8);
    const int arow = m & 7;
    const int mylen = top ? len0 : len1;
    const int pbase = top ? p00 : p10;
    const int len = max(len0, len1);
    const int ntiles = nrhs / 16;
    for (int t0 = 0; t0 < ntiles; t0 += 2) {
        const bool two = t0 + 1 < ntiles;
        v4f acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = {0.f, 0.f, 0.f, 0.f};
        for (int q0 = 0; q0 < len; q0 += Q) {
            int col[Q];
            v4s a[Q], b0[Q], b1[Q];
#pragma unroll
            for (int q = 0; q < Q; ++q) col[q] = (q0 + q < mylen) ? ci[pbase + q0 + q] : -1;
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                a[q] = v4s{0, 0, 0, 0};
                b0[q] = v4s{0, 0, 0, 0};
                b1[q] = v4s{0, 0, 0, 0};
                if (col[q] >= 0) {
                    if (a_live) a[q] = load_a(A + (size_t)(pbase + q0 + q) * 64 + arow * 8 + kk);
                    const uint16_t* xb = Xb + ((size_t)col[q] * nrhs + t0 * 16 + m) * 8 + kk;
                    b0[q] = *reinterpret_cast<const v4s*>(xb);
                    if (two) b1[q] = *reinterpret_cast<const v4s*>(xb + 16 * 8);
                }
            }
            const int qn = min(Q, len - q0);
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                if (q < qn) {
                    acc0 = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[q], b0[q], acc0, 0, 0, 0);
                    if (two) acc1 = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[q], b1[q], acc1, 0, 0, 0);
                }
            }
        }
        const int row_blk = top ? i0 : i1;
        if (row_blk < n) {
            float* y = Yb + ((size_t)row_blk * nrhs + t0 * 16 + m) * 8 + kk;
            store_y(y, acc0);
            if (two) store_y(y + 16 * 8, acc1);
        }
    }
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

int launch_bsmm_bf16(const bsp_ctx* ctx, int n, int nrhs, const int* rp, const int* ci, const uint16_t* A,
                     const uint16_t* Xb, float* Yb) {
    if (n <= 0) return BSP_OK;
    const int pairs = (n + 1) / 2;
    const int nwg = (pairs + kBlock / 64 - 1) / (kBlock / 64);
    const int grid = (nwg + kXcds - 1) / kXcds * kXcds;
    hipLaunchKernelGGL((k_bsmm_bf16_pf<4>), dim3(grid), dim3(kBlock), 0, ctx->stream, n, nrhs, rp, ci, A, Xb, Yb, nwg);
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
    if (!A->val16.p) {
        const int64_t cnt = A->nnzb * 64;
        BSP_HIP(A->val16.alloc(sizeof(uint16_t) * std::max<int64_t>(cnt, 1)));
        int rc = launch_values_bf16(ctx, cnt, A->val.as<double>(), A->val16.as<uint16_t>());
        if (rc) return rc;
    }
    return launch_bsmm_bf16(ctx, A->n_rows, nrhs, A->rp.as<int>(), A->ci.as<int>(), A->val16.as<uint16_t>(), Xb_dev,
                            Yb_dev);
}

}  // extern "C"
