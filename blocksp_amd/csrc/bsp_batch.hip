// bsp_batch.hip -- batched IChol0-pCG over many independent systems in one persistent
// launch (gfx950).
//
// The systems are the diagonal blocks of one block-diagonal BSR matrix.  A work item
// is (system, group of R right-hand sides); items are ordered largest system first and
// handed out by an atomic counter, so the persistent grid load-balances uneven sizes
// (longest-processing-time-first).  A workgroup runs the whole pCG of its item: vectors
// are interleaved by right-hand side (entry (row, c, k) at (row * BS + c) * nrhs + k),
// so the R right-hand sides of an item share every matrix and factor load.  The two
// triangular sweeps of the IChol0 apply walk the system's level sets with a workgroup
// barrier between levels -- no kernel launch per level -- and thread (row, c) forms
// component c of the row's residual, the BS threads of a row exchange them with lane
// shuffles and each applies its row of the diagonal-block inverse.  Reductions are
// wave shuffles plus a fixed-order sum over the waves (deterministic for a given
// workgroup size).
#include "bsp_internal.hpp"
#include "bsp_batch.hpp"

namespace bsp {
namespace batch {

namespace {

constexpr int kMaxWaves = 16;

template <int N>
__device__ __forceinline__ void block_sum(double (&v)[N], double* sh) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
    for (int j = 0; j < N; ++j)
        for (int off = 32; off > 0; off >>= 1) v[j] += __shfl_xor(v[j], off);
    if (lane == 0)
#pragma unroll
        for (int j = 0; j < N; ++j) sh[wave * N + j] = v[j];
    __syncthreads();
#pragma unroll
    for (int j = 0; j < N; ++j) {
        double s = 0.0;
        for (int w = 0; w < nw; ++w) s += sh[w * N + j];
        v[j] = s;
    }
    __syncthreads();
}

// One triangular sweep over system s's level sets:
//   out_i = Dinv_i (in_i - sum_p T_p out_{col p})   for the rows of each level in turn.
template <int BS, int R>
__device__ void sweep(const int* __restrict__ rows, const int* __restrict__ lev, int lb, int le,
                      const int* __restrict__ trp, const int* __restrict__ tci, const double* __restrict__ tval,
                      const double* __restrict__ dinv, const double* in, double* out, int nrhs, int k0) {
    constexpr int B2 = BS * BS;
    const int T = blockDim.x, tid = threadIdx.x, lane = tid & 63;
    const int gbase = lane & ~(BS - 1);
    for (int l = lb; l < le; ++l) {
        const int a0 = lev[l], cnt = (lev[l + 1] - a0) * BS;
        for (int idx = tid; idx < cnt; idx += T) {
            const int row = rows[a0 + idx / BS];
            const int c = idx % BS;
            const size_t self = ((size_t)row * BS + c) * nrhs + k0;
            double t[R];
#pragma unroll
            for (int k = 0; k < R; ++k) t[k] = in[self + k];
            for (int p = trp[row]; p < trp[row + 1]; ++p) {
                const double* blk = tval + (size_t)p * B2 + c * BS;
                const double* xc = out + (size_t)tci[p] * BS * nrhs + k0;
#pragma unroll
                for (int j = 0; j < BS; ++j) {
                    const double aj = blk[j];
#pragma unroll
                    for (int k = 0; k < R; ++k) t[k] = fma(-aj, xc[j * nrhs + k], t[k]);
                }
            }
            const double* dr = dinv + (size_t)row * B2 + c * BS;
            double y[R];
#pragma unroll
            for (int k = 0; k < R; ++k) y[k] = 0.0;
#pragma unroll
            for (int j = 0; j < BS; ++j) {
                const double dj = dr[j];
#pragma unroll
                for (int k = 0; k < R; ++k) y[k] = fma(dj, __shfl(t[k], gbase + j), y[k]);
            }
#pragma unroll
            for (int k = 0; k < R; ++k) out[self + k] = y[k];
        }
        __syncthreads();
    }
}

template <int BS, int R>
__global__ __launch_bounds__(1024) void k_batch_pcg(Args a) {
    __shared__ double red[kMaxWaves * 2 * R];
    __shared__ int item_sh;
    constexpr int B2 = BS * BS;
    const int T = blockDim.x, tid = threadIdx.x;
    const int nrhs = a.nrhs;
    for (;;) {
        if (tid == 0) item_sh = (int)atomicAdd(a.work, 1u);
        __syncthreads();
        const int item = item_sh;
        __syncthreads();
        if (item >= a.nitems) break;
        const int s = a.item_sys[item], k0 = a.item_rhs0[item];
        const int r0 = a.sys_row0[s], r1 = a.sys_row0[s + 1];
        const int nslot = (r1 - r0) * BS;
        const size_t base = (size_t)r0 * BS * nrhs + k0;
        // x = 0, r = b, ||b||^2
        double acc[2 * R];
#pragma unroll
        for (int k = 0; k < 2 * R; ++k) acc[k] = 0.0;
        for (int idx = tid; idx < nslot; idx += T) {
            const size_t g = base + (size_t)idx * nrhs;
#pragma unroll
            for (int k = 0; k < R; ++k) {
                const double bv = a.B[g + k];
                a.X[g + k] = 0.0;
                a.Rv[g + k] = bv;
                acc[k] += bv * bv;
            }
        }
        __syncthreads();
        // z = M^{-1} r  (forward sweep with L, then backward with L^T)
        sweep<BS, R>(a.frows, a.flev, a.flev_base[s], a.flev_base[s + 1] - 1, a.lrp, a.lci, a.lval, a.dinvL, a.Rv,
                     a.Z, nrhs, k0);
        sweep<BS, R>(a.brows, a.blev, a.blev_base[s], a.blev_base[s + 1] - 1, a.urp, a.uci, a.uval, a.dinvU, a.Z, a.Z,
                     nrhs, k0);
        // p = z; rz = r.z; ||z||^2
        double rzz[2 * R];
#pragma unroll
        for (int k = 0; k < 2 * R; ++k) rzz[k] = 0.0;
        for (int idx = tid; idx < nslot; idx += T) {
            const size_t g = base + (size_t)idx * nrhs;
#pragma unroll
            for (int k = 0; k < R; ++k) {
                const double zv = a.Z[g + k];
                a.P[g + k] = zv;
                rzz[k] += a.Rv[g + k] * zv;
                rzz[R + k] += zv * zv;
            }
        }
        block_sum<2 * R>(rzz, red);
        block_sum<2 * R>(acc, red);
        double rz[R], ref[R], res[R];
        bool active[R];
        int its[R];
#pragma unroll
        for (int k = 0; k < R; ++k) {
            rz[k] = rzz[k];
            // reference rule: ||r|| <= tol ||M^{-1} b|| (blocksp_solvers.hpp:58-66); else ||r|| <= tol ||b||
            ref[k] = a.mode == 0 ? sqrt(rzz[R + k]) : sqrt(acc[k]);
            res[k] = sqrt(acc[k]);
            active[k] = res[k] > a.tol * ref[k];
            its[k] = 0;
        }
        for (int it = 0; it < a.maxIters; ++it) {
            bool any = false;
#pragma unroll
            for (int k = 0; k < R; ++k) any |= active[k];
            if (!any) break;
            // Ap = A p, pAp
            double pap[R];
#pragma unroll
            for (int k = 0; k < R; ++k) pap[k] = 0.0;
            for (int idx = tid; idx < nslot; idx += T) {
                const int row = r0 + idx / BS, c = idx % BS;
                double y[R];
#pragma unroll
                for (int k = 0; k < R; ++k) y[k] = 0.0;
                for (int p = a.rp[row]; p < a.rp[row + 1]; ++p) {
                    const double* blk = a.val + (size_t)p * B2 + c * BS;
                    const double* xc = a.P + (size_t)a.ci[p] * BS * nrhs + k0;
#pragma unroll
                    for (int j = 0; j < BS; ++j) {
                        const double aj = blk[j];
#pragma unroll
                        for (int k = 0; k < R; ++k) y[k] = fma(aj, xc[j * nrhs + k], y[k]);
                    }
                }
                const size_t g = base + (size_t)idx * nrhs;
#pragma unroll
                for (int k = 0; k < R; ++k) {
                    a.AP[g + k] = y[k];
                    pap[k] += a.P[g + k] * y[k];
                }
            }
            block_sum<R>(pap, red);
            double alpha[R], rr[R];
#pragma unroll
            for (int k = 0; k < R; ++k) {
                alpha[k] = active[k] ? rz[k] / pap[k] : 0.0;
                rr[k] = 0.0;
            }
            for (int idx = tid; idx < nslot; idx += T) {
                const size_t g = base + (size_t)idx * nrhs;
#pragma unroll
                for (int k = 0; k < R; ++k) {
                    if (!active[k]) continue;
                    a.X[g + k] = fma(alpha[k], a.P[g + k], a.X[g + k]);
                    const double rn = fma(-alpha[k], a.AP[g + k], a.Rv[g + k]);
                    a.Rv[g + k] = rn;
                    rr[k] += rn * rn;
                }
            }
            block_sum<R>(rr, red);
            bool cont = false;
#pragma unroll
            for (int k = 0; k < R; ++k) {
                if (!active[k]) continue;
                ++its[k];
                res[k] = sqrt(rr[k]);
                active[k] = res[k] > a.tol * ref[k] && isfinite(res[k]);
                cont |= active[k];
            }
            if (!cont) break;
            sweep<BS, R>(a.frows, a.flev, a.flev_base[s], a.flev_base[s + 1] - 1, a.lrp, a.lci, a.lval, a.dinvL,
                         a.Rv, a.Z, nrhs, k0);
            sweep<BS, R>(a.brows, a.blev, a.blev_base[s], a.blev_base[s + 1] - 1, a.urp, a.uci, a.uval, a.dinvU,
                         a.Z, a.Z, nrhs, k0);
            double rzn[R];
#pragma unroll
            for (int k = 0; k < R; ++k) rzn[k] = 0.0;
            for (int idx = tid; idx < nslot; idx += T) {
                const size_t g = base + (size_t)idx * nrhs;
#pragma unroll
                for (int k = 0; k < R; ++k) rzn[k] += a.Rv[g + k] * a.Z[g + k];
            }
            block_sum<R>(rzn, red);
            double beta[R];
#pragma unroll
            for (int k = 0; k < R; ++k) {
                beta[k] = rzn[k] / rz[k];
                if (active[k]) rz[k] = rzn[k];
            }
            for (int idx = tid; idx < nslot; idx += T) {
                const size_t g = base + (size_t)idx * nrhs;
#pragma unroll
                for (int k = 0; k < R; ++k)
                    if (active[k]) a.P[g + k] = fma(beta[k], a.P[g + k], a.Z[g + k]);
            }
            __syncthreads();
        }
        if (tid == 0) {
#pragma unroll
            for (int k = 0; k < R; ++k) {
                const size_t o = (size_t)s * nrhs + k0 + k;
                a.iters[o] = its[k];
                a.res[2 * o] = res[k];
                a.res[2 * o + 1] = ref[k];
            }
        }
        __syncthreads();
    }
}

// RHS-major (k, row*bs + c) <-> interleaved ((row*bs + c) * nrhs + k)
__global__ void k_interleave(int64_t len, int nrhs, const double* __restrict__ src, double* __restrict__ dst,
                             int to_interleaved) {
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= len * nrhs) return;
    const int64_t e = t / nrhs;
    const int k = (int)(t - e * nrhs);
    if (to_interleaved) dst[t] = src[(size_t)k * len + e];
    else dst[(size_t)k * len + e] = src[t];
}

template <int BS, int R>
const void* kernel_ptr() {
    return reinterpret_cast<const void*>(&k_batch_pcg<BS, R>);
}

template <int BS>
const void* pick_r(int R) {
    switch (R) {
        case 1: return kernel_ptr<BS, 1>();
        case 2: return kernel_ptr<BS, 2>();
        case 4: return kernel_ptr<BS, 4>();
    }
    return nullptr;
}

const void* pick(int bs, int R) {
    switch (bs) {
        case 1: return pick_r<1>(R);
        case 2: return pick_r<2>(R);
        case 4: return pick_r<4>(R);
        case 8: return pick_r<8>(R);
    }
    return nullptr;
}

}  // namespace

bool supported(int bs, int R) { return pick(bs, R) != nullptr; }

int max_resident_blocks(const bsp_ctx* ctx, int bs, int R, int wg) {
    const void* k = pick(bs, R);
    if (!k) return 0;
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, wg, 0) != hipSuccess || per_cu < 1) per_cu = 1;
    return per_cu * ctx->num_cus;
}

int launch_pcg(const bsp_ctx* ctx, int bs, int R, int wg, int grid, const Args& a) {
    const void* k = pick(bs, R);
    BSP_CHECK(k != nullptr, BSP_E_UNSUPPORTED, "batched pCG supports block sizes 1, 2, 4, 8 and 1, 2, 4 RHS per item");
    BSP_CHECK(wg % 64 == 0 && wg >= 64 && wg <= 1024, BSP_E_INVALID, "workgroup size must be a multiple of 64 <= 1024");
    void* args[] = {const_cast<Args*>(&a)};
    BSP_HIP(hipLaunchKernel(k, dim3(grid), dim3(wg), args, 0, ctx->stream));
    return BSP_OK;
}

int launch_interleave(const bsp_ctx* ctx, int64_t len, int nrhs, const double* src, double* dst, bool to_interleaved) {
    const int64_t total = len * nrhs;
    if (total <= 0) return BSP_OK;
    hipLaunchKernelGGL(k_interleave, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, ctx->stream, len, nrhs, src,
                       dst, (int)to_interleaved);
    BSP_HIP(hipGetLastError());
    return BSP_OK;
}

}  // namespace batch
}  // namespace bsp
