// bsp_mgs.hip -- modified Gram-Schmidt of one Arnoldi step without grid-wide fences.
//
// For w against q_0..q_j (one_arnoldi, blocksp_solvers.hpp:127-135):
//   step s = 0..j+1 (one kernel each):
//     if s > 0: h_{s-1} = sum of the previous step's block partials;  w -= h_{s-1} q_{s-1}
//     partial_s[block] = sum over the block's elements of q_s w   (s <= j)   or   w w   (s = j+1)
//   scale kernel: h_{j+1} = sqrt(|sum partial_{j+1}|);  q_{j+1} = w / h_{j+1}
// Short vectors (at most 4 partials) run all steps in one kernel of one workgroup, with the
// same arithmetic.  Every block of a kernel sums the previous partials itself, in one fixed order
// (thread t takes partials t, t+256, ..., then a fixed tree), so all blocks use the
// bitwise-identical h, and block 0 records it in hcol for the host.  The kernel
// boundary orders the partials; no device-scope fence or atomic is needed.
#include "bsp_internal.hpp"

#include <algorithm>

namespace bsp {

namespace {

constexpr int kWave = 64;
constexpr int kBlock = 256;
constexpr int kTile = kBlock * 4;            // elements per block per tile (2 x double2 per thread)
constexpr int kMaxParts = 1024;

__device__ __forceinline__ double block_sum(double v, double* sh) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    const int w = threadIdx.x / kWave;
    if ((threadIdx.x & (kWave - 1)) == 0) sh[w] = v;
    __syncthreads();
    const double r = sh[0] + sh[1] + sh[2] + sh[3];
    __syncthreads();
    return r;
}

// sum of nparts partials, identical in every block
__device__ __forceinline__ double sum_partials(const double* __restrict__ part, int nparts, double* sh) {
    double s = 0.0;
    for (int b = threadIdx.x; b < nparts; b += kBlock) s += part[b];
    return block_sum(s, sh);
}

template <bool VEC>
__global__ __launch_bounds__(kBlock) void k_mgs2(int64_t len, double* __restrict__ v, const double* __restrict__ qprev,
                                                 const double* __restrict__ qnext, const double* __restrict__ pin,
                                                 double* __restrict__ pout, int nparts, double* __restrict__ hprev_out) {
    __shared__ double sh[kBlock / kWave];
    double mh = 0.0;
    if (qprev) {
        const double h = sum_partials(pin, nparts, sh);
        if (blockIdx.x == 0 && threadIdx.x == 0) *hprev_out = h;
        mh = -h;
    }
    double s0 = 0.0, s1 = 0.0;
    for (int64_t base = (int64_t)blockIdx.x * kTile; base < len; base += (int64_t)gridDim.x * kTile) {
        if (VEC && base + kTile <= len) {
            double2* v2 = reinterpret_cast<double2*>(v + base);
            double2 a0 = v2[threadIdx.x], a1 = v2[threadIdx.x + kBlock];
            double2 p0{0.0, 0.0}, p1{0.0, 0.0}, q0{0.0, 0.0}, q1{0.0, 0.0};
            if (qprev) {
                const double2* p2 = reinterpret_cast<const double2*>(qprev + base);
                p0 = p2[threadIdx.x];
                p1 = p2[threadIdx.x + kBlock];
            }
            if (qnext) {
                const double2* q2 = reinterpret_cast<const double2*>(qnext + base);
                q0 = q2[threadIdx.x];
                q1 = q2[threadIdx.x + kBlock];
            }
            if (qprev) {
                a0.x = a0.x + mh * p0.x;
                a0.y = a0.y + mh * p0.y;
                a1.x = a1.x + mh * p1.x;
                a1.y = a1.y + mh * p1.y;
                v2[threadIdx.x] = a0;
                v2[threadIdx.x + kBlock] = a1;
            }
            if (qnext) {
                s0 = fma(q0.x, a0.x, s0);
                s1 = fma(q0.y, a0.y, s1);
                s0 = fma(q1.x, a1.x, s0);
                s1 = fma(q1.y, a1.y, s1);
            } else {
                s0 = fma(a0.x, a0.x, s0);
                s1 = fma(a0.y, a0.y, s1);
                s0 = fma(a1.x, a1.x, s0);
                s1 = fma(a1.y, a1.y, s1);
            }
        } else {
            for (int64_t i = base + threadIdx.x; i < len && i < base + kTile; i += kBlock) {
                double vi = v[i];
                if (qprev) {
                    vi = vi + mh * qprev[i];
                    v[i] = vi;
                }
                s0 = qnext ? fma(qnext[i], vi, s0) : fma(vi, vi, s0);
            }
        }
    }
    const double s = block_sum(s0 + s1, sh);
    if (threadIdx.x == 0) pout[blockIdx.x] = s;
}

__global__ __launch_bounds__(kBlock) void k_mgs_scale(int64_t len, const double* __restrict__ v,
                                                      const double* __restrict__ pin, int nparts,
                                                      double* __restrict__ q, double* __restrict__ hout) {
    __shared__ double sh[kBlock / kWave];
    const double h = sqrt(fabs(sum_partials(pin, nparts, sh)));
    if (blockIdx.x == 0 && threadIdx.x == 0) *hout = h;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < len; i += (int64_t)gridDim.x * kBlock)
        q[i] = v[i] / h;
}

// One-workgroup form for short vectors (at most 4 partials): thread group g (256 threads) does
// what block g of k_mgs2 / k_mgs_scale does, over all steps of the Arnoldi step in one kernel,
// the partials passing through LDS; every group re-sums them in k_mgs2's order, so h and the
// vectors are bitwise those of the launch per step.  Each thread rewrites only elements it owns.
constexpr int kOneGroups = 4;
constexpr int kOneMaxQ = 64;

struct QList {
    const double* q[kOneMaxQ];
};

__device__ __forceinline__ double group_sum(double v, double* sh, int tid) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    if ((tid & (kWave - 1)) == 0) sh[tid / kWave] = v;
    __syncthreads();
    const double r = sh[0] + sh[1] + sh[2] + sh[3];
    __syncthreads();
    return r;
}

template <bool VEC>
__global__ __launch_bounds__(kBlock * kOneGroups) void k_mgs_onewg(int64_t len, double* __restrict__ v, QList Q,
                                                                   int j, int nparts, double* __restrict__ hcol,
                                                                   double* __restrict__ qnew) {
    __shared__ double sh[kOneGroups][kBlock / kWave];
    __shared__ double part[2][kOneGroups];
    const int grp = threadIdx.x / kBlock, tid = threadIdx.x % kBlock;
    auto sum_parts = [&](const double* pin) {
        double s = 0.0;
        for (int b = tid; b < nparts; b += kBlock) s += pin[b];
        return group_sum(s, sh[grp], tid);
    };
    for (int st = 0; st <= j + 1; ++st) {
        const double* qprev = st ? Q.q[st - 1] : nullptr;
        const double* qnext = st <= j ? Q.q[st] : nullptr;
        double mh = 0.0;
        if (qprev) {
            const double h = sum_parts(part[(st + 1) & 1]);
            if (grp == 0 && tid == 0) hcol[st - 1] = h;
            mh = -h;
        }
        double s0 = 0.0, s1 = 0.0;
        for (int64_t base = (int64_t)grp * kTile; base < len; base += (int64_t)nparts * kTile) {
            if (VEC && base + kTile <= len) {
                double2* v2 = reinterpret_cast<double2*>(v + base);
                double2 a0 = v2[tid], a1 = v2[tid + kBlock];
                double2 p0{0.0, 0.0}, p1{0.0, 0.0}, q0{0.0, 0.0}, q1{0.0, 0.0};
                if (qprev) {
                    const double2* p2 = reinterpret_cast<const double2*>(qprev + base);
                    p0 = p2[tid];
                    p1 = p2[tid + kBlock];
                }
                if (qnext) {
                    const double2* q2 = reinterpret_cast<const double2*>(qnext + base);
                    q0 = q2[tid];
                    q1 = q2[tid + kBlock];
                }
                if (qprev) {
                    a0.x = a0.x + mh * p0.x;
                    a0.y = a0.y + mh * p0.y;
                    a1.x = a1.x + mh * p1.x;
                    a1.y = a1.y + mh * p1.y;
                    v2[tid] = a0;
                    v2[tid + kBlock] = a1;
                }
                if (qnext) {
                    s0 = fma(q0.x, a0.x, s0);
                    s1 = fma(q0.y, a0.y, s1);
                    s0 = fma(q1.x, a1.x, s0);
                    s1 = fma(q1.y, a1.y, s1);
                } else {
                    s0 = fma(a0.x, a0.x, s0);
                    s1 = fma(a0.y, a0.y, s1);
                    s0 = fma(a1.x, a1.x, s0);
                    s1 = fma(a1.y, a1.y, s1);
                }
            } else {
                for (int64_t i = base + tid; i < len && i < base + kTile; i += kBlock) {
                    double vi = v[i];
                    if (qprev) {
                        vi = vi + mh * qprev[i];
                        v[i] = vi;
                    }
                    s0 = qnext ? fma(qnext[i], vi, s0) : fma(vi, vi, s0);
                }
            }
        }
        const double ps = group_sum(s0 + s1, sh[grp], tid);
        if (tid == 0) part[st & 1][grp] = ps;
        __syncthreads();
    }
    const double h = sqrt(fabs(sum_parts(part[(j + 1) & 1])));
    if (grp == 0 && tid == 0) hcol[j + 1] = h;
    for (int64_t base = (int64_t)grp * kTile; base < len; base += (int64_t)nparts * kTile) {
        if (VEC && base + kTile <= len) {
            for (int e = 0; e < 2; ++e) {
                const int64_t i = base + 2 * (tid + e * kBlock);
                qnew[i] = v[i] / h;
                qnew[i + 1] = v[i + 1] / h;
            }
        } else {
            for (int64_t i = base + tid; i < len && i < base + kTile; i += kBlock) qnew[i] = v[i] / h;
        }
    }
}

// The same for vectors of whole tiles (at most 2 per group): each thread keeps its elements of w
// and of the previous basis vector in registers across the steps and loads the next basis vector
// one step ahead, so a step's loads overlap the previous step's reductions; w is written back
// once at the end.  Same operations in the same order as k_mgs_onewg.
struct Tile2 {
    double2 h0, h1;   // elements 2 tid, 2 tid + 1 and 512 + 2 tid, 513 + 2 tid of the tile
};

__global__ __launch_bounds__(kBlock * kOneGroups) void k_mgs_onewg_reg(int64_t len, double* __restrict__ v, QList Q,
                                                                       int j, int nparts, double* __restrict__ hcol,
                                                                       double* __restrict__ qnew) {
    // per wavefront sums of a step, double-buffered: one barrier per step
    __shared__ double shw[2][kOneGroups * (kBlock / kWave)];
    const int grp = threadIdx.x / kBlock, tid = threadIdx.x % kBlock;
    int64_t base[2];
    bool own[2];
    Tile2 a[2], prev[2], cur[2], nxt[2];
    auto ld = [&](const double* q, int tt) {
        const double2* q2 = reinterpret_cast<const double2*>(q + base[tt]);
        return Tile2{q2[tid], q2[tid + kBlock]};
    };
#pragma unroll
    for (int tt = 0; tt < 2; ++tt) {
        base[tt] = ((int64_t)grp + (int64_t)tt * nparts) * kTile;
        own[tt] = base[tt] < len;
        if (own[tt]) {
            a[tt] = ld(v, tt);
            cur[tt] = ld(Q.q[0], tt);
        }
    }
    // The step's sum from the wavefront sums, in the order k_mgs2 / k_mgs_scale take: each group's
    // four wavefront sums in order (block_sum), then the group sums as the 64-lane xor butterfly of
    // block_sum adds them -- values first plus +0 (the zeros of the idle lanes), then (p0 + p2) +
    // (p1 + p3) -- so the result has the same bits as the launch per step.
    auto sum_parts = [&](const double* w) {
        double p[kOneGroups];
#pragma unroll
        for (int g = 0; g < kOneGroups; ++g) {
            const double* x = w + g * (kBlock / kWave);
            p[g] = g < nparts ? ((x[0] + x[1]) + x[2]) + x[3] : 0.0;
            p[g] = p[g] + 0.0;
        }
        return (p[0] + p[2]) + (p[1] + p[3]);
    };
    for (int st = 0; st <= j + 1; ++st) {
        if (st + 1 <= j) {
#pragma unroll
            for (int tt = 0; tt < 2; ++tt)
                if (own[tt]) nxt[tt] = ld(Q.q[st + 1], tt);
        }
        double mh = 0.0;
        if (st > 0) {
            const double h = sum_parts(shw[(st + 1) & 1]);
            if (grp == 0 && tid == 0) hcol[st - 1] = h;
            mh = -h;
        }
        double s0 = 0.0, s1 = 0.0;
#pragma unroll
        for (int tt = 0; tt < 2; ++tt) {
            if (!own[tt]) continue;
            double2 a0 = a[tt].h0, a1 = a[tt].h1;
            if (st > 0) {
                a0.x = a0.x + mh * prev[tt].h0.x;
                a0.y = a0.y + mh * prev[tt].h0.y;
                a1.x = a1.x + mh * prev[tt].h1.x;
                a1.y = a1.y + mh * prev[tt].h1.y;
                a[tt].h0 = a0;
                a[tt].h1 = a1;
            }
            if (st <= j) {
                const double2 q0 = cur[tt].h0, q1 = cur[tt].h1;
                s0 = fma(q0.x, a0.x, s0);
                s1 = fma(q0.y, a0.y, s1);
                s0 = fma(q1.x, a1.x, s0);
                s1 = fma(q1.y, a1.y, s1);
            } else {
                s0 = fma(a0.x, a0.x, s0);
                s1 = fma(a0.y, a0.y, s1);
                s0 = fma(a1.x, a1.x, s0);
                s1 = fma(a1.y, a1.y, s1);
            }
        }
        double ps = s0 + s1;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) ps += __shfl_xor(ps, o, kWave);
        if ((threadIdx.x & (kWave - 1)) == 0) shw[st & 1][threadIdx.x / kWave] = ps;
        __syncthreads();
#pragma unroll
        for (int tt = 0; tt < 2; ++tt) {
            prev[tt] = cur[tt];
            cur[tt] = nxt[tt];
        }
    }
    const double h = sqrt(fabs(sum_parts(shw[(j + 1) & 1])));
    if (grp == 0 && tid == 0) hcol[j + 1] = h;
#pragma unroll
    for (int tt = 0; tt < 2; ++tt) {
        if (!own[tt]) continue;
        double2* v2 = reinterpret_cast<double2*>(v + base[tt]);
        double2* q2 = reinterpret_cast<double2*>(qnew + base[tt]);
        v2[tid] = a[tt].h0;
        v2[tid + kBlock] = a[tt].h1;
        q2[tid] = double2{a[tt].h0.x / h, a[tt].h0.y / h};
        q2[tid + kBlock] = double2{a[tt].h1.x / h, a[tt].h1.y / h};
    }
}

inline bool aligned16(const void* p) { return ((uintptr_t)p & 15u) == 0; }

}  // namespace

int mgs_parts(int64_t len) {
    const int64_t b = (len + 2 * kTile - 1) / (2 * kTile);
    return (int)std::max<int64_t>(1, std::min<int64_t>(b, kMaxParts));
}

int launch_arnoldi_mgs(const bsp_ctx* ctx, int64_t len, double* w, const double* const* Q, int j, double* hcol,
                       double* qnew, double* parts) {
    const int nb = mgs_parts(len);
    double* P[2] = {parts, parts + kMaxParts};
    bool vec = aligned16(w);
    for (int s = 0; s <= j; ++s) vec = vec && aligned16(Q[s]);
    if (ctx->mgs_onewg && nb <= kOneGroups && j + 1 <= kOneMaxQ) {
        QList ql{};
        for (int s = 0; s <= j; ++s) ql.q[s] = Q[s];
        // whole tiles, at most 2 per group (nb >= tiles / 2 by construction), 16-byte aligned qnew
        if (vec && len % kTile == 0 && (len / kTile) <= 2 * nb && aligned16(qnew))
            hipLaunchKernelGGL(k_mgs_onewg_reg, dim3(1), dim3(nb * kBlock), 0, ctx->stream, len, w, ql, j, nb, hcol,
                               qnew);
        else if (vec)
            hipLaunchKernelGGL(k_mgs_onewg<true>, dim3(1), dim3(nb * kBlock), 0, ctx->stream, len, w, ql, j, nb, hcol,
                               qnew);
        else
            hipLaunchKernelGGL(k_mgs_onewg<false>, dim3(1), dim3(nb * kBlock), 0, ctx->stream, len, w, ql, j, nb, hcol,
                               qnew);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return fail(ctx, BSP_E_HIP, std::string("mgs launch: ") + hipGetErrorString(e));
        return BSP_OK;
    }
    for (int s = 0; s <= j + 1; ++s) {
        const double* qprev = s ? Q[s - 1] : nullptr;
        const double* qnext = s <= j ? Q[s] : nullptr;
        const double* pin = P[(s + 1) & 1];
        double* pout = P[s & 1];
        double* hprev = s ? hcol + (s - 1) : nullptr;
        if (vec)
            hipLaunchKernelGGL(k_mgs2<true>, dim3(nb), dim3(kBlock), 0, ctx->stream, len, w, qprev, qnext, pin, pout, nb,
                               hprev);
        else
            hipLaunchKernelGGL(k_mgs2<false>, dim3(nb), dim3(kBlock), 0, ctx->stream, len, w, qprev, qnext, pin, pout,
                               nb, hprev);
    }
    const int gs = (int)std::min<int64_t>((len + kBlock - 1) / kBlock, 4096);
    hipLaunchKernelGGL(k_mgs_scale, dim3(gs), dim3(kBlock), 0, ctx->stream, len, w, P[(j + 1) & 1], nb, qnew,
                       hcol + j + 1);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(ctx, BSP_E_HIP, std::string("mgs launch: ") + hipGetErrorString(e));
    return BSP_OK;
}

}  // namespace bsp
