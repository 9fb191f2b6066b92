// bsp_mpnn.hip -- kernels of the learned ILU0-correction network (gfx950).
//
// The network is a 3-layer message-passing net (hidden width 64) over the block
// graph of a BSR matrix.  Every dense layer is a row-wise product
//     Y[r] = act(W [X0[idx0 r] ; X1[idx1 r]] + b)
// over node rows or edge rows with W of at most 64 outputs and two input parts
// of at most 64 features each.  One wavefront handles one row at a time: lane o
// keeps output o's weights in registers, the row's inputs are loaded once
// (coalesced, one feature per lane) and broadcast with v_readlane, so the
// product needs no LDS and no synchronisation.  The backward pass uses the same
// mapping with W transposed (lane k keeps input k's weights), and the weight
// gradients are reduced per workgroup into partial sums that a second kernel adds
// in a fixed order, so training is deterministic run to run.
//
// The loss kernels build the corrected factor F and evaluate L'U'z for a batch
// of sample vectors (FP32, block size <= 8), and their adjoints.
#include "bsp_internal.hpp"
#include "bsp_mpnn.hpp"

#include <algorithm>

namespace bsp {
namespace mpnn {

namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ float bcast(float v, int k) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), k));
}

__device__ __forceinline__ int row_of(const Part& p, int r) { return p.idx ? p.idx[r] : r; }

template <int NP, bool RELU>
__global__ __launch_bounds__(kBlock) void k_linear_fwd(int R, int out, Part p0, Part p1, int in_total,
                                                       const float* __restrict__ W, const float* __restrict__ b,
                                                       float* __restrict__ Y, int ystride,
                                                       const float* __restrict__ resid, float* __restrict__ Y2) {
    const int lane = threadIdx.x & 63;
    const int wave = (int)((blockIdx.x * (size_t)blockDim.x + threadIdx.x) >> 6);
    const int nwaves = (int)((gridDim.x * (size_t)blockDim.x) >> 6);
    const bool act = lane < out;
    float w0[64];
    float w1[NP > 1 ? 64 : 1];
#pragma unroll
    for (int k = 0; k < 64; ++k) w0[k] = (act && k < p0.dim) ? W[(size_t)lane * in_total + k] : 0.0f;
    if constexpr (NP > 1) {
#pragma unroll
        for (int k = 0; k < 64; ++k)
            w1[k] = (act && k < p1.dim) ? W[(size_t)lane * in_total + p0.dim + k] : 0.0f;
    }
    const float bias = act ? b[lane] : 0.0f;
    for (int r = wave; r < R; r += nwaves) {
        const float v0 = lane < p0.dim ? p0.x[(size_t)row_of(p0, r) * p0.stride + lane] : 0.0f;
        float acc = bias;
#pragma unroll
        for (int k = 0; k < 64; ++k) acc = fmaf(w0[k], bcast(v0, k), acc);
        if constexpr (NP > 1) {
            const float v1 = lane < p1.dim ? p1.x[(size_t)row_of(p1, r) * p1.stride + lane] : 0.0f;
#pragma unroll
            for (int k = 0; k < 64; ++k) acc = fmaf(w1[k], bcast(v1, k), acc);
        }
        if (RELU) acc = fmaxf(acc, 0.0f);
        if (act) {
            Y[(size_t)r * ystride + lane] = acc;
            if (Y2) Y2[(size_t)r * ystride + lane] = resid[(size_t)r * ystride + lane] + acc;
        }
    }
}

// gradient of the layer output for row r, with the ReLU mask applied
__device__ __forceinline__ float grad_out(const GradIn& g, int out, int lane, int r) {
    if (lane >= out) return 0.0f;
    const int gr = g.idx ? g.idx[r] : r;
    float v = g.gy[(size_t)gr * g.stride + lane];
    if (g.mask && !(g.mask[(size_t)r * g.mstride + lane] > 0.0f)) v = 0.0f;
    return v;
}

template <int NP>
__global__ __launch_bounds__(kBlock) void k_linear_bwd_in(int R, int out, int dim0, int dim1, int in_total,
                                                          const float* __restrict__ W, GradIn g, float* gx0, int s0,
                                                          int acc0, float* gx1, int s1, int acc1) {
    const int lane = threadIdx.x & 63;
    const int wave = (int)((blockIdx.x * (size_t)blockDim.x + threadIdx.x) >> 6);
    const int nwaves = (int)((gridDim.x * (size_t)blockDim.x) >> 6);
    float wt0[64];
    float wt1[NP > 1 ? 64 : 1];
#pragma unroll
    for (int o = 0; o < 64; ++o) wt0[o] = (o < out && lane < dim0) ? W[(size_t)o * in_total + lane] : 0.0f;
    if constexpr (NP > 1) {
#pragma unroll
        for (int o = 0; o < 64; ++o)
            wt1[o] = (o < out && lane < dim1) ? W[(size_t)o * in_total + dim0 + lane] : 0.0f;
    }
    for (int r = wave; r < R; r += nwaves) {
        const float gy = grad_out(g, out, lane, r);
        float a0 = 0.0f;
#pragma unroll
        for (int o = 0; o < 64; ++o) a0 = fmaf(wt0[o], bcast(gy, o), a0);
        if (gx0 && lane < dim0) {
            float* p = gx0 + (size_t)r * s0 + lane;
            *p = acc0 ? *p + a0 : a0;
        }
        if constexpr (NP > 1) {
            float a1 = 0.0f;
#pragma unroll
            for (int o = 0; o < 64; ++o) a1 = fmaf(wt1[o], bcast(gy, o), a1);
            if (gx1 && lane < dim1) {
                float* p = gx1 + (size_t)r * s1 + lane;
                *p = acc1 ? *p + a1 : a1;
            }
        }
    }
}

// Per-workgroup partial weight gradients: partial[blk][o][c] = sum_r gY[r][o] X[r][c]
// (c == in_total: the bias).  Wave w of the 4 covers input columns [32w, 32w + 32).
template <int NP>
__global__ __launch_bounds__(kBlock) void k_linear_bwd_w(int R, int out, Part p0, Part p1, int in_total, GradIn g,
                                                         int rows_per_block, float* __restrict__ partial) {
    const int lane = threadIdx.x & 63;
    const int w = threadIdx.x >> 6;
    const int r0 = blockIdx.x * rows_per_block;
    const int r1 = min(R, r0 + rows_per_block);
    float acc[32];
#pragma unroll
    for (int j = 0; j < 32; ++j) acc[j] = 0.0f;
    float bacc = 0.0f;
    const int col = w * 32 + lane;
    for (int r = r0; r < r1; ++r) {
        const float gy = grad_out(g, out, lane, r);
        float xv = 0.0f;
        if (lane < 32) {
            if (col < p0.dim) xv = p0.x[(size_t)row_of(p0, r) * p0.stride + col];
            else if (NP > 1 && col - p0.dim < p1.dim)
                xv = p1.x[(size_t)row_of(p1, r) * p1.stride + (col - p0.dim)];
        }
#pragma unroll
        for (int j = 0; j < 32; ++j) acc[j] = fmaf(gy, bcast(xv, j), acc[j]);
        bacc += gy;
    }
    if (lane < out) {
        float* dst = partial + ((size_t)blockIdx.x * out + lane) * (in_total + 1);
#pragma unroll
        for (int j = 0; j < 32; ++j)
            if (w * 32 + j < in_total) dst[w * 32 + j] = acc[j];
        if (w == 0) dst[in_total] = bacc;
    }
}

__global__ void k_reduce_wgrad(int nblk, int out, int in_total, const float* __restrict__ partial,
                               float* __restrict__ gW, float* __restrict__ gb) {
    const int total = out * (in_total + 1);
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    // four independent partial sums (loads in flight), combined in a fixed order
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    int b = 0;
    for (; b + 4 <= nblk; b += 4) {
        s0 += partial[(size_t)b * total + t];
        s1 += partial[(size_t)(b + 1) * total + t];
        s2 += partial[(size_t)(b + 2) * total + t];
        s3 += partial[(size_t)(b + 3) * total + t];
    }
    for (; b < nblk; ++b) s0 += partial[(size_t)b * total + t];
    const double s = (s0 + s1) + (s2 + s3);
    const int o = t / (in_total + 1), c = t - o * (in_total + 1);
    if (c < in_total) gW[(size_t)o * in_total + c] = (float)s;
    else gb[o] = (float)s;
}

// dst[i][f] (+)= sum_{e in [ptr[i], ptr[i+1])} src[idx ? idx[e] : e][f], 64 features
__global__ void k_segsum(int n, const int* __restrict__ ptr, const int* __restrict__ idx,
                         const float* __restrict__ src, float* __restrict__ dst, int accumulate) {
    const size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    const int i = (int)(t >> 6), f = (int)(t & 63);
    if (i >= n) return;
    float s = 0.0f;
    for (int e = ptr[i]; e < ptr[i + 1]; ++e) {
        const int se = idx ? idx[e] : e;
        s += src[(size_t)se * 64 + f];
    }
    float* p = dst + (size_t)i * 64 + f;
    *p = accumulate ? *p + s : s;
}

__global__ void k_node_feat(int n, int b2, const int* __restrict__ diag_slot, const double* __restrict__ val,
                            double* __restrict__ sd, float* __restrict__ s32, float* __restrict__ x) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double* a = val + (size_t)diag_slot[i] * b2;
    double s2 = 0.0;
    for (int q = 0; q < b2; ++q) s2 += a[q] * a[q];
    const double s = sqrt(s2);
    sd[i] = s;
    s32[i] = (float)s;
    for (int q = 0; q < b2; ++q) x[(size_t)i * b2 + q] = (float)(a[q] / s);
}

__global__ void k_edge_feat(int E, int b2, const int* __restrict__ esrc, const int* __restrict__ edst,
                            const int* __restrict__ eslot, const double* __restrict__ val,
                            const double* __restrict__ sd, float* __restrict__ ef, float* __restrict__ ef2) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    const double si = sd[esrc[e]], sj = sd[edst[e]];
    const double inv = 1.0 / sqrt(si * sj);
    const double* a = val + (size_t)eslot[e] * b2;
    for (int q = 0; q < b2; ++q) ef[(size_t)e * b2 + q] = (float)(a[q] * inv);
    ef2[e] = (float)(log(sj) - log(si));
}

// F_ii = LU_ii (I + alpha tanh(D_i)),  F_e = LU_e (1 + beta tanh(g_e))
__global__ void k_assemble(int64_t total, int bs, const int* __restrict__ slot_row, const int* __restrict__ slot_edge,
                           const float* __restrict__ lu, const float* __restrict__ dD, const float* __restrict__ gl,
                           float alpha, float beta, float* __restrict__ F) {
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= total) return;
    const int b2 = bs * bs;
    const int64_t p = t / b2;
    const int q = (int)(t - p * b2);
    const int e = slot_edge[p];
    if (e < 0) {
        const int i = slot_row[p];
        const int r = q / bs, c = q - r * bs;
        const float* L = lu + (size_t)p * b2 + r * bs;
        const float* D = dD + (size_t)i * b2 + c;
        float acc = 0.0f;
        for (int k = 0; k < bs; ++k) acc = fmaf(L[k], tanhf(D[k * bs]), acc);
        F[t] = fmaf(alpha, acc, lu[t]);
    } else {
        F[t] = lu[t] * (1.0f + beta * tanhf(gl[e]));
    }
}

// Y[k][i][r] = add[k][i][r] + sum over the row's blocks of the chosen side of F_p X[k][col p]
// (UPPER: blocks with lower[p] == 0, including the diagonal; else lower[p] == 1).
template <bool UPPER>
__global__ void k_factor_prod(int n, int bs, int K, const int* __restrict__ rp, const int* __restrict__ ci,
                              const unsigned char* __restrict__ lower, const float* __restrict__ F,
                              const float* __restrict__ X, const float* __restrict__ add, float* __restrict__ Y) {
    const int64_t len = (int64_t)n * bs;
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= len * K) return;
    const int k = (int)(t / len);
    const int64_t rem = t - (int64_t)k * len;
    const int i = (int)(rem / bs), r = (int)(rem - (int64_t)i * bs);
    const float* Xk = X + (size_t)k * len;
    float acc = add ? add[t] : 0.0f;
    const int b2 = bs * bs;
    for (int p = rp[i]; p < rp[i + 1]; ++p) {
        if ((lower[p] != 0) == UPPER) continue;
        const float* f = F + (size_t)p * b2 + r * bs;
        const float* x = Xk + (size_t)ci[p] * bs;
        for (int c = 0; c < bs; ++c) acc = fmaf(f[c], x[c], acc);
    }
    Y[t] = acc;
}

// per (sample, block) partial sums of (y - r)^2 and r^2, in double
__global__ __launch_bounds__(kBlock) void k_loss_sums(int64_t len, const float* __restrict__ y,
                                                      const float* __restrict__ r, double* __restrict__ out) {
    const int k = blockIdx.y;
    const float* yk = y + (size_t)k * len;
    const float* rk = r + (size_t)k * len;
    double a = 0.0, b = 0.0;
    for (int64_t t = blockIdx.x * (int64_t)kBlock + threadIdx.x; t < len; t += (int64_t)gridDim.x * kBlock) {
        const double d = (double)yk[t] - (double)rk[t];
        a += d * d;
        b += (double)rk[t] * rk[t];
    }
    __shared__ double sa[kBlock], sb[kBlock];
    sa[threadIdx.x] = a;
    sb[threadIdx.x] = b;
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            sa[threadIdx.x] += sa[threadIdx.x + s];
            sb[threadIdx.x] += sb[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[((size_t)k * gridDim.x + blockIdx.x) * 2] = sa[0];
        out[((size_t)k * gridDim.x + blockIdx.x) * 2 + 1] = sb[0];
    }
}

__global__ void k_grad_y(int64_t len, int K, const float* __restrict__ coef, const float* __restrict__ y,
                         const float* __restrict__ r, float* __restrict__ gy) {
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= len * K) return;
    gy[t] = coef[t / len] * (y[t] - r[t]);
}

// gu[k][j][c] = gy[k][j][c] + sum over lower blocks (i, j) of (F_ij^T gy[k][i])[c]; the
// block (i, j) is found through row j's block (j, i) and the transposed-slot map.
__global__ void k_grad_u(int n, int bs, int K, const int* __restrict__ rp, const int* __restrict__ ci,
                         const int* __restrict__ tslot, const unsigned char* __restrict__ lower,
                         const float* __restrict__ F, const float* __restrict__ gy, float* __restrict__ gu) {
    const int64_t len = (int64_t)n * bs;
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= len * K) return;
    const int k = (int)(t / len);
    const int64_t rem = t - (int64_t)k * len;
    const int j = (int)(rem / bs), c = (int)(rem - (int64_t)j * bs);
    const float* g = gy + (size_t)k * len;
    const int b2 = bs * bs;
    float acc = gy[t];
    for (int p = rp[j]; p < rp[j + 1]; ++p) {
        const int q = tslot[p];
        if (!lower[q]) continue;
        const float* f = F + (size_t)q * b2 + c;
        const float* gi = g + (size_t)ci[p] * bs;
        for (int r = 0; r < bs; ++r) acc = fmaf(f[r * bs], gi[r], acc);
    }
    gu[t] = acc;
}

// gF[p][r][c] = sum_k (lower ? gy[k][row][r] u[k][col][c] : gu[k][row][r] z[k][col][c])
__global__ void k_grad_F(int64_t total, int bs, int K, int64_t len, const int* __restrict__ slot_row,
                         const int* __restrict__ ci, const unsigned char* __restrict__ lower,
                         const float* __restrict__ gy, const float* __restrict__ u, const float* __restrict__ gu,
                         const float* __restrict__ z, float* __restrict__ gF) {
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= total) return;
    const int b2 = bs * bs;
    const int64_t p = t / b2;
    const int q = (int)(t - p * b2);
    const int r = q / bs, c = q - r * bs;
    const int64_t ro = (int64_t)slot_row[p] * bs + r, co = (int64_t)ci[p] * bs + c;
    const bool lo = lower[p] != 0;
    const float* A = lo ? gy : gu;
    const float* B = lo ? u : z;
    float acc = 0.0f;
    for (int k = 0; k < K; ++k) acc = fmaf(A[(size_t)k * len + ro], B[(size_t)k * len + co], acc);
    gF[t] = acc;
}

// g_dD[i][r][c] = alpha (1 - tanh(D_i[r][c])^2) (LU_ii^T gF_ii)[r][c]
__global__ void k_grad_heads_d(int n, int bs, const int* __restrict__ diag_slot, const float* __restrict__ lu,
                               const float* __restrict__ dD, const float* __restrict__ gF, float alpha,
                               float* __restrict__ g_dD) {
    const int b2 = bs * bs;
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= (int64_t)n * b2) return;
    const int i = (int)(t / b2), q = (int)(t - (int64_t)i * b2);
    const int r = q / bs, c = q - r * bs;
    const size_t base = (size_t)diag_slot[i] * b2;
    float acc = 0.0f;
    for (int k = 0; k < bs; ++k) acc = fmaf(lu[base + k * bs + r], gF[base + k * bs + c], acc);
    const float th = tanhf(dD[t]);
    g_dD[t] = alpha * (1.0f - th * th) * acc;
}

__global__ void k_grad_heads_g(int E, int b2, const int* __restrict__ eslot, const float* __restrict__ lu,
                               const float* __restrict__ gF, const float* __restrict__ gl, float beta,
                               float* __restrict__ g_gl) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    const size_t base = (size_t)eslot[e] * b2;
    float s = 0.0f;
    for (int q = 0; q < b2; ++q) s = fmaf(lu[base + q], gF[base + q], s);
    const float th = tanhf(gl[e]);
    g_gl[e] = beta * (1.0f - th * th) * s;
}

__global__ void k_adam(int64_t n, float* __restrict__ theta, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2) {
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= n) return;
    const float gt = g[t];
    const float mt = b1 * m[t] + (1.0f - b1) * gt;
    const float vt = b2 * v[t] + (1.0f - b2) * gt * gt;
    m[t] = mt;
    v[t] = vt;
    theta[t] -= lr * (mt * c1) / (sqrtf(vt * c2) + eps);
}

inline unsigned blocks_for(int64_t work, int per = kBlock) { return (unsigned)((work + per - 1) / per); }

// waves for a grid-stride row loop: enough to fill the chip, never more than rows
inline unsigned row_grid(const bsp_ctx* ctx, int R) {
    const int64_t waves = std::min<int64_t>(std::max(R, 1), (int64_t)ctx->num_cus * 16);
    return (unsigned)((waves + 3) / 4);
}

}  // namespace

int linear_fwd(const bsp_ctx* ctx, int R, int out, const Part& p0, const Part* p1, int in_total, const float* W,
               const float* b, float* Y, int ystride, bool relu, const float* resid, float* Y2) {
    if (R <= 0) return BSP_OK;
    const unsigned grid = row_grid(ctx, R);
    const Part q1 = p1 ? *p1 : Part{};
    if (p1) {
        if (relu) hipLaunchKernelGGL((k_linear_fwd<2, true>), dim3(grid), dim3(kBlock), 0, ctx->stream, R, out, p0, q1, in_total, W, b, Y, ystride, resid, Y2);
        else hipLaunchKernelGGL((k_linear_fwd<2, false>), dim3(grid), dim3(kBlock), 0, ctx->stream, R, out, p0, q1, in_total, W, b, Y, ystride, resid, Y2);
    } else {
        if (relu) hipLaunchKernelGGL((k_linear_fwd<1, true>), dim3(grid), dim3(kBlock), 0, ctx->stream, R, out, p0, q1, in_total, W, b, Y, ystride, resid, Y2);
        else hipLaunchKernelGGL((k_linear_fwd<1, false>), dim3(grid), dim3(kBlock), 0, ctx->stream, R, out, p0, q1, in_total, W, b, Y, ystride, resid, Y2);
    }
    BSP_HIP(hipGetLastError());
    return BSP_OK;
}

int linear_bwd_in(const bsp_ctx* ctx, int R, int out, int dim0, int dim1, int in_total, const float* W,
                  const GradIn& g, float* gx0, int s0, bool acc0, float* gx1, int s1, bool acc1) {
    if (R <= 0) return BSP_OK;
    const unsigned grid = row_grid(ctx, R);
    if (dim1 > 0)
        hipLaunchKernelGGL((k_linear_bwd_in<2>), dim3(grid), dim3(kBlock), 0, ctx->stream, R, out, dim0, dim1, in_total, W, g, gx0, s0, (int)acc0, gx1, s1, (int)acc1);
    else
        hipLaunchKernelGGL((k_linear_bwd_in<1>), dim3(grid), dim3(kBlock), 0, ctx->stream, R, out, dim0, 0, in_total, W, g, gx0, s0, (int)acc0, gx1, s1, (int)acc1);
    BSP_HIP(hipGetLastError());
    return BSP_OK;
}

int linear_bwd_w(const bsp_ctx* ctx, int R, int out, const Part& p0, const Part* p1, int in_total, const GradIn& g,
                 float* partial, size_t partial_floats, float* gW, float* gb) {
    BSP_CHECK(in_total <= 128 && out <= 64, BSP_E_INVALID, "layer too wide for the weight-gradient kernel");
    const size_t per_blk = (size_t)out * (in_total + 1);
    int nblk = std::max(1, std::min<int>((R + 63) / 64, kMaxPartialBlocks));
    nblk = std::min<int>(nblk, (int)(partial_floats / per_blk));
    BSP_CHECK(nblk >= 1, BSP_E_INVALID, "weight-gradient scratch too small");
    const int rpb = (std::max(R, 1) + nblk - 1) / nblk;
    nblk = (std::max(R, 1) + rpb - 1) / rpb;
    const Part q1 = p1 ? *p1 : Part{};
    if (p1)
        hipLaunchKernelGGL((k_linear_bwd_w<2>), dim3(nblk), dim3(kBlock), 0, ctx->stream, R, out, p0, q1, in_total, g, rpb, partial);
    else
        hipLaunchKernelGGL((k_linear_bwd_w<1>), dim3(nblk), dim3(kBlock), 0, ctx->stream, R, out, p0, q1, in_total, g, rpb, partial);
    BSP_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_reduce_wgrad, dim3(blocks_for((int64_t)per_blk)), dim3(kBlock), 0, ctx->stream, nblk, out,
                       in_total, partial, gW, gb);
    BSP_HIP(hipGetLastError());
    return BSP_OK;
}

int segsum(const bsp_ctx* ctx, int n, const int* ptr, const int* idx, const float* src, float* dst, bool accumulate) {
    if (n <= 0) return BSP_OK;
    hipLaunchKernelGGL(k_segsum, dim3(blocks_for((int64_t)n * 64)), dim3(kBlock), 0, ctx->stream, n, ptr, idx, src,
                       dst, (int)accumulate);
    BSP_HIP(hipGetLastError());
    return BSP_OK;
}

int features(const bsp_ctx* ctx, int n, int E, int b2, const int* diag_slot, const int* esrc, const int* edst,
             const int* eslot, const double* val, double* sd, float* s32, float* x, float* ef, float* ef2) {
    if (n > 0)
        hipLaunchKernelGGL(k_node_feat, dim3(blocks_for(n)), dim3(kBlock), 0, ctx->stream, n, b2, diag_slot, val, sd,
                           s32, x);
    if (E > 0)
        hipLaunchKernelGGL(k_edge_feat, dim3(blocks_for(E)), dim3(kBlock), 0, ctx->stream, E, b2, esrc, edst, eslot,
                           val, sd, ef, ef2);
    BSP_HIP(hipGetLastError());
    return BSP_OK;
}

int assemble(const bsp_ctx* ctx, int64_t nnzb, int bs, const int* slot_row, const int* slot_edge, const float* lu,
             const float* dD, const float* gl, float alpha, float beta, float* F) {
    const int64_t total = nnzb * bs * bs;
    if (total <= 0) return BSP_OK;
    hipLaunchKernelGGL(k_assemble, dim3(blocks_for(total)), dim3(kBlock), 0, ctx->stream, total, bs, slot_row,
                       slot_edge, lu, dD, gl, alpha, beta, F);
    BSP_HIP(hipGetLastError());
    return BSP_OK;
}

int factor_prod(const bsp_ctx* ctx, bool upper, int n, int bs, int K, const int* rp, const int* ci,
                const unsigned char* lower, const float* F, const float* X, const float* add, float* Y) {
    const int64_t total = (int64_t)n * bs * K;
    if (total <= 0) return BSP_OK;
    if (upper)
        hipLaunchKernelGGL((k_factor_prod<true>), dim3(blocks_for(total)), dim3(kBlock), 0, ctx->stream, n, bs, K, rp, ci, lower, F, X, add, Y);
    else
        hipLaunchKernelGGL((k_factor_prod<false>), dim3(blocks_for(total)), dim3(kBlock), 0, ctx->stream, n, bs, K, rp, ci, lower, F, X, add, Y);
    BSP_HIP(hipGetLastError());
    return BSP_OK;
}

int loss_sums(const bsp_ctx* ctx, int64_t len, int K, int nblk, const float* y, const float* r, double* out) {
    hipLaunchKernelGGL(k_loss_sums, dim3(nblk, K), dim3(kBlock), 0, ctx->stream, len, y, r, out);
    BSP_HIP(hipGetLastError());
    return BSP_OK;
}

int grad_y(const bsp_ctx* ctx, int64_t len, int K, const float* coef, const float* y, const float* r, float* gy) {
    hipLaunchKernelGGL(k_grad_y, dim3(blocks_for(len * K)), dim3(kBlock), 0, ctx->stream, len, K, coef, y, r, gy);
    BSP_HIP(hipGetLastError());
    return BSP_OK;
}

int grad_u(const bsp_ctx* ctx, int n, int bs, int K, const int* rp, const int* ci, const int* tslot,
           const unsigned char* lower, const float* F, const float* gy, float* gu) {
    hipLaunchKernelGGL(k_grad_u, dim3(blocks_for((int64_t)n * bs * K)), dim3(kBlock), 0, ctx->stream, n, bs, K, rp,
                       ci, tslot, lower, F, gy, gu);
    BSP_HIP(hipGetLastError());
    return BSP_OK;
}

int grad_F(const bsp_ctx* ctx, int64_t nnzb, int bs, int K, int64_t len, const int* slot_row, const int* ci,
           const unsigned char* lower, const float* gy, const float* u, const float* gu, const float* z, float* gF) {
    const int64_t total = nnzb * bs * bs;
    hipLaunchKernelGGL(k_grad_F, dim3(blocks_for(total)), dim3(kBlock), 0, ctx->stream, total, bs, K, len, slot_row,
                       ci, lower, gy, u, gu, z, gF);
    BSP_HIP(hipGetLastError());
    return BSP_OK;
}

int grad_heads(const bsp_ctx* ctx, int n, int E, int bs, const int* diag_slot, const int* eslot, const float* lu,
               const float* dD, const float* gF, const float* gl, float alpha, float beta, float* g_dD, float* g_gl) {
    const int b2 = bs * bs;
    hipLaunchKernelGGL(k_grad_heads_d, dim3(blocks_for((int64_t)n * b2)), dim3(kBlock), 0, ctx->stream, n, bs,
                       diag_slot, lu, dD, gF, alpha, g_dD);
    if (E > 0)
        hipLaunchKernelGGL(k_grad_heads_g, dim3(blocks_for(E)), dim3(kBlock), 0, ctx->stream, E, b2, eslot, lu, gF, gl,
                           beta, g_gl);
    BSP_HIP(hipGetLastError());
    return BSP_OK;
}

int adam(const bsp_ctx* ctx, int64_t n, float* theta, const float* g, float* m, float* v, float lr, float b1,
         float b2, float eps, float c1, float c2) {
    hipLaunchKernelGGL(k_adam, dim3(blocks_for(n)), dim3(kBlock), 0, ctx->stream, n, theta, g, m, v, lr, b1, b2, eps,
                       c1, c2);
    BSP_HIP(hipGetLastError());
    return BSP_OK;
}

}  // namespace mpnn
}  // namespace bsp
