// bsp_mpnn.hpp -- launchers of the learned ILU0-correction network (bsp_mpnn.hip).
#pragma once

#include "bsp_internal.hpp"

namespace bsp {
namespace mpnn {

constexpr int kHidden = 64;
constexpr int kLayers = 3;
// diagonal correction F_ii = LU_ii (I + kAlpha tanh(D_i)); with |entries| <= kAlpha and
// block size <= 8, ||kAlpha tanh(D_i)||_2 <= 0.4, so F_ii stays invertible
constexpr float kAlpha = 0.05f;
constexpr float kBeta = 0.25f;   // off-diagonal scaling: F_ij = LU_ij (1 + kBeta tanh g_ij)
constexpr int kMaxPartialBlocks = 256;   // workgroups (partials) of a weight-gradient reduction

// One input part of a dense layer: row r reads x[(idx ? idx[r] : r) * stride + 0 .. dim-1].
struct Part {
    const float* x = nullptr;
    const int* idx = nullptr;
    int dim = 0;      // <= 64
    int stride = 0;
};

// Gradient with respect to a layer's output for row r: gy[(idx ? idx[r] : r) * stride + o],
// zeroed where mask[r * mstride + o] <= 0 (the ReLU of the forward pass) when mask != nullptr.
struct GradIn {
    const float* gy = nullptr;
    const int* idx = nullptr;
    int stride = 0;
    const float* mask = nullptr;
    int mstride = 0;
};

// Y[r] = act(W [p0 ; p1] + b) for r < R; W is [out][in_total] row-major (out <= 64).
// With Y2 != nullptr also Y2[r] = resid[r] + Y[r] (same stride).
int linear_fwd(const bsp_ctx* ctx, int R, int out, const Part& p0, const Part* p1, int in_total, const float* W,
               const float* b, float* Y, int ystride, bool relu, const float* resid, float* Y2);
// gx_p[r] (+)= (W_p^T g(r)) for the two input parts (dim1 == 0: one part).
int linear_bwd_in(const bsp_ctx* ctx, int R, int out, int dim0, int dim1, int in_total, const float* W,
                  const GradIn& g, float* gx0, int s0, bool acc0, float* gx1, int s1, bool acc1);
// gW = sum_r g(r) [p0 ; p1](r)^T, gb = sum_r g(r) (overwritten; deterministic order).
int linear_bwd_w(const bsp_ctx* ctx, int R, int out, const Part& p0, const Part* p1, int in_total, const GradIn& g,
                 float* partial, size_t partial_floats, float* gW, float* gb);
// dst[i] (+)= sum_{e in [ptr[i], ptr[i+1])} src[idx ? idx[e] : e]  (64 features per row)
int segsum(const bsp_ctx* ctx, int n, const int* ptr, const int* idx, const float* src, float* dst, bool accumulate);
int features(const bsp_ctx* ctx, int n, int E, int b2, const int* diag_slot, const int* esrc, const int* edst,
             const int* eslot, const double* val, double* sd, float* s32, float* x, float* ef, float* ef2);
int assemble(const bsp_ctx* ctx, int64_t nnzb, int bs, const int* slot_row, const int* slot_edge, const float* lu,
             const float* dD, const float* gl, float alpha, float beta, float* F);
int factor_prod(const bsp_ctx* ctx, bool upper, int n, int bs, int K, const int* rp, const int* ci,
                const unsigned char* lower, const float* F, const float* X, const float* add, float* Y);
int loss_sums(const bsp_ctx* ctx, int64_t len, int K, int nblk, const float* y, const float* r, double* out);
int grad_y(const bsp_ctx* ctx, int64_t len, int K, const float* coef, const float* y, const float* r, float* gy);
int grad_u(const bsp_ctx* ctx, int n, int bs, int K, const int* rp, const int* ci, const int* tslot,
           const unsigned char* lower, const float* F, const float* gy, float* gu);
int grad_F(const bsp_ctx* ctx, int64_t nnzb, int bs, int K, int64_t len, const int* slot_row, const int* ci,
           const unsigned char* lower, const float* gy, const float* u, const float* gu, const float* z, float* gF);
int grad_heads(const bsp_ctx* ctx, int n, int E, int bs, const int* diag_slot, const int* eslot, const float* lu,
               const float* dD, const float* gF, const float* gl, float alpha, float beta, float* g_dD, float* g_gl);
int adam(const bsp_ctx* ctx, int64_t n, float* theta, const float* g, float* m, float* v, float lr, float b1,
         float b2, float eps, float c1, float c2);

}  // namespace mpnn
}  // namespace bsp
