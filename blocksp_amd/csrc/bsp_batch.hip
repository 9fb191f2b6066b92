// bsp_batch.hip -- batched IChol0-pCG over many independent systems in one persistent
// launch (gfx950).
//
// The systems are the diagonal blocks of one block-diagonal BSR matrix.  A work item
// is (system, group of R right-hand sides); items are ordered largest system first and
// handed out by an atomic counter, so the persistent grid load-balances uneven sizes
// (longest-processing-time-first).  A workgroup runs the whole pCG of its item.  Vectors
// are stored per group of R right-hand sides and interleaved inside the group: entry
// (row, c, k) at (k / R) * len * R + (row * BS + c) * R + k % R (len = rows * BS), so an
// item's vectors are contiguous and its R right-hand sides share every matrix and factor
// load.  The two triangular sweeps of the IChol0 apply walk the system's level sets with
// a workgroup barrier between levels -- no kernel launch per level -- and thread (row, c)
// forms component c of the row's residual, the BS threads of a row exchange them with
// lane shuffles and each applies its row of the diagonal-block inverse.  Reductions are
// wave shuffles plus a fixed-order sum over the waves (deterministic for a given
// workgroup size).
//
// One workgroup streams a whole system's vectors, so a vector pass is bound by the loads a
// thread has in flight: each thread takes U slots per round (4 / R, or 1 for R = 4), issues
// all their loads, then computes them in slot order (the order of the one-slot loop, so
// the sums are the same).
//
// A level step is a chain of dependent loads, so its length sets the sweep's speed.  With
// the factor in CSR the chain is level pointer -> row -> row pointer -> column -> vector.
// The ELL form (K > 0) stores each sweep slot-ordered (slot = position in the level-sorted
// row list): K columns per slot (-1 padding), their K blocks and the slot's diagonal-block
// inverse, so the slot's structure and values load at once -- a level ahead, while the
// previous level computes -- and only the vector entries wait on earlier levels.  Entries
// keep the CSR order and padding is skipped, so both forms run the same floating-point
// operations in the same order.
#include "bsp_internal.hpp"
#include "bsp_batch.hpp"

namespace bsp {
namespace batch {

namespace {

constexpr int kMaxWaves = 16;
constexpr int kEllMaxBs = 4;   // larger blocks: a slot a level ahead does not fit the registers

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

// R consecutive doubles of one slot (16-byte aligned for R >= 2: group bases and slot
// offsets are multiples of R doubles).
template <int R>
__device__ __forceinline__ void ld(const double* p, double (&v)[R]) {
    if constexpr (R == 1) {
        v[0] = *p;
    } else {
#pragma unroll
        for (int k = 0; k < R; k += 2) {
            const double2 t = *reinterpret_cast<const double2*>(p + k);
            v[k] = t.x;
            v[k + 1] = t.y;
        }
    }
}

template <int R>
__device__ __forceinline__ void st(double* p, const double (&v)[R]) {
    if constexpr (R == 1) {
        *p = v[0];
    } else {
#pragma unroll
        for (int k = 0; k < R; k += 2) *reinterpret_cast<double2*>(p + k) = make_double2(v[k], v[k + 1]);
    }
}

// One triangular sweep over system s's level sets:
//   out_i = Dinv_i (in_i - sum_p T_p out_{col p})   for the rows of each level in turn.
// in / out: the item's group (entry (row, c, k) at (row * BS + c) * R + k; offsets in a
// group fit in 32 bits).
template <int BS, int R>
__device__ void sweep(const int* __restrict__ rows, const int* __restrict__ lev, int lb, int le,
                      const int* __restrict__ trp, const int* __restrict__ tci, const double* __restrict__ tval,
                      const double* __restrict__ dinv, const double* in, double* out) {
    constexpr int B2 = BS * BS;
    const int T = blockDim.x, tid = threadIdx.x, lane = tid & 63;
    const int gbase = lane & ~(BS - 1);
    for (int l = lb; l < le; ++l) {
        const int a0 = lev[l], cnt = (lev[l + 1] - a0) * BS;
        for (int idx = tid; idx < cnt; idx += T) {
            const int row = rows[a0 + idx / BS];
            const int c = idx % BS;
            const int self = (row * BS + c) * R;
            double t[R];
            ld<R>(in + self, t);
            for (int p = trp[row]; p < trp[row + 1]; ++p) {
                const double* blk = tval + (size_t)p * B2 + c * BS;
                const double* xc = out + tci[p] * BS * R;
#pragma unroll
                for (int j = 0; j < BS; ++j) {
                    const double aj = blk[j];
                    double xv[R];
                    ld<R>(xc + j * R, xv);
#pragma unroll
                    for (int k = 0; k < R; ++k) t[k] = fma(-aj, xv[k], t[k]);
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
            st<R>(out + self, y);
        }
        __syncthreads();
    }
}

// The same sweep over the ELL copy: slot q holds columns col[q*K + k] (-1: padding), blocks
// val[(q*K + k) * BS*BS] and the inverse diagonal block dinv[q * BS*BS]; rows[q] is its row.
template <int BS, int K>
struct EllSlot {
    int q, row;
    int cq[K];
    double av[K][BS];
};

// The slot's structure and its row of the K blocks (what a level ahead can load).
template <int BS, int K>
__device__ __forceinline__ void ell_load(const int* __restrict__ rows, const int* __restrict__ col,
                                         const double* __restrict__ val, int q, int c, EllSlot<BS, K>& e) {
    constexpr int B2 = BS * BS;
    e.q = q;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        e.cq[k] = col[(size_t)q * K + k];
#pragma unroll
        for (int j = 0; j < BS; ++j) e.av[k][j] = val[((size_t)q * K + k) * B2 + c * BS + j];
    }
    e.row = rows[q];
}

// Right-hand side, inverse diagonal block and the gathered vector entries, then the row.
template <int BS, int R, int K>
__device__ __forceinline__ void ell_finish(const double* __restrict__ dinv, const double* in, double* out, int c,
                                           int gbase, const EllSlot<BS, K>& e) {
    constexpr int B2 = BS * BS;
    double t[R], dv[BS];
    ld<R>(in + (e.row * BS + c) * R, t);
#pragma unroll
    for (int j = 0; j < BS; ++j) dv[j] = dinv[(size_t)e.q * B2 + c * BS + j];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        if (e.cq[k] < 0) continue;
        const double* xc = out + e.cq[k] * BS * R;
#pragma unroll
        for (int j = 0; j < BS; ++j) {
            double xv[R];
            ld<R>(xc + j * R, xv);
#pragma unroll
            for (int kk = 0; kk < R; ++kk) t[kk] = fma(-e.av[k][j], xv[kk], t[kk]);
        }
    }
    double y[R];
#pragma unroll
    for (int k = 0; k < R; ++k) y[k] = 0.0;
#pragma unroll
    for (int j = 0; j < BS; ++j)
#pragma unroll
        for (int k = 0; k < R; ++k) y[k] = fma(dv[j], __shfl(t[k], gbase + j), y[k]);
    st<R>(out + (e.row * BS + c) * R, y);
}

template <int BS, int R, int K>
__device__ void sweep_ell(const int* __restrict__ rows, const int* __restrict__ lev, int lb, int le,
                          const int* __restrict__ col, const double* __restrict__ val,
                          const double* __restrict__ dinv, const double* in, double* out) {
    const int T = blockDim.x, tid = threadIdx.x, lane = tid & 63;
    const int gbase = lane & ~(BS - 1);
    const int c = tid % BS;   // T is a multiple of BS: every slot of this thread has component c
    if (lb >= le) return;
    if constexpr (K * BS > 8) {
        // wide slots: loading one a level ahead would spill registers
        for (int l = lb; l < le; ++l) {
            const int a0 = lev[l], cnt = (lev[l + 1] - a0) * BS;
            for (int i0 = tid; i0 < cnt; i0 += T) {
                EllSlot<BS, K> e;
                ell_load<BS, K>(rows, col, val, a0 + i0 / BS, c, e);
                ell_finish<BS, R, K>(dinv, in, out, c, gbase, e);
            }
            __syncthreads();
        }
        return;
    }
    // The structure and blocks of the thread's first slot of the next level are loaded while
    // this level computes (the sweep writes none of them), so a level's chain is the loads
    // of its vector entries, right-hand sides and inverse diagonal blocks.
    int a0 = lev[lb], a1 = lev[lb + 1];
    EllSlot<BS, K> nxt;
    bool have_n = tid < (a1 - a0) * BS;
    if (have_n) ell_load<BS, K>(rows, col, val, a0 + tid / BS, c, nxt);
    for (int l = lb; l < le; ++l) {
        const int a2 = l + 2 <= le ? lev[l + 2] : a1;
        const int cnt = (a1 - a0) * BS;
        EllSlot<BS, K> cur = nxt;
        const bool have_c = have_n;
        // slots beyond the first of a level wider than the workgroup: loaded in place
        for (int i0 = tid + T; i0 < cnt; i0 += T) {
            EllSlot<BS, K> e;
            ell_load<BS, K>(rows, col, val, a0 + i0 / BS, c, e);
            ell_finish<BS, R, K>(dinv, in, out, c, gbase, e);
        }
        have_n = l + 1 < le && tid < (a2 - a1) * BS;
        if (have_n) ell_load<BS, K>(rows, col, val, a1 + tid / BS, c, nxt);
        if (have_c) ell_finish<BS, R, K>(dinv, in, out, c, gbase, cur);
        __syncthreads();
        a0 = a1;
        a1 = a2;
    }
}

// z = M^{-1} r: forward sweep with L, then backward with L^T, in the CSR (K == 0) or ELL form.
template <int BS, int R, int K>
__device__ __forceinline__ void apply_ichol0(const Args& a, int s, size_t vb) {
    if constexpr (K == 0) {
        sweep<BS, R>(a.frows, a.flev, a.flev_base[s], a.flev_base[s + 1] - 1, a.lrp, a.lci, a.lval, a.dinvL,
                     a.Rv + vb, a.Z + vb);
        sweep<BS, R>(a.brows, a.blev, a.blev_base[s], a.blev_base[s + 1] - 1, a.urp, a.uci, a.uval, a.dinvU,
                     a.Z + vb, a.Z + vb);
    } else {
        sweep_ell<BS, R, K>(a.frows, a.flev, a.flev_base[s], a.flev_base[s + 1] - 1, a.fcol, a.fval, a.fdinv,
                            a.Rv + vb, a.Z + vb);
        sweep_ell<BS, R, K>(a.brows, a.blev, a.blev_base[s], a.blev_base[s + 1] - 1, a.bcol, a.bval, a.bdinv,
                            a.Z + vb, a.Z + vb);
    }
}

// The slots idx = tid, tid + T, ... of [0, nslot) in increasing order: rounds of U slots
// whose loads are all issued before any is used, then single slots.  `load(idx, st)` reads
// a slot into st, `use(idx, st)` computes and stores it.
template <int U, typename S, typename Load, typename Use>
__device__ __forceinline__ void for_slots(int nslot, Load&& load, Use&& use) {
    const int T = blockDim.x;
    int i0 = threadIdx.x;
    for (; i0 + (U - 1) * T < nslot; i0 += U * T) {
        S st[U];
#pragma unroll
        for (int u = 0; u < U; ++u) load(i0 + u * T, st[u]);
#pragma unroll
        for (int u = 0; u < U; ++u) use(i0 + u * T, st[u]);
    }
    for (; i0 < nslot; i0 += T) {
        S st;
        load(i0, st);
        use(i0, st);
    }
}

template <int R>
struct V1 {
    double a[R];
};
template <int R>
struct V2 {
    double a[R], b[R];
};
template <int R>
struct V4 {
    double a[R], b[R], c[R], d[R];
};

template <int BS, int R, int K>
__global__ __launch_bounds__(1024) void k_batch_pcg(Args a) {
    __shared__ double red[kMaxWaves * 2 * R];
    __shared__ int item_sh;
    constexpr int B2 = BS * BS;
    constexpr int U = R == 4 ? 1 : 4 / R;   // slots per thread per round of a vector pass (more
                                            // spills registers once the sweeps load a level ahead)
    const int T = blockDim.x, tid = threadIdx.x;
    for (;;) {
        if (tid == 0) item_sh = (int)atomicAdd(a.work, 1u);
        __syncthreads();
        const int item = item_sh;
        __syncthreads();
        if (item >= a.nitems) break;
        const int s = a.item_sys[item], k0 = a.item_rhs0[item];
        const int r0 = a.sys_row0[s], r1 = a.sys_row0[s + 1];
        const int nslot = (r1 - r0) * BS;
        const size_t vb = (size_t)(k0 / R) * (size_t)a.len * R;   // the item's group
        // the item's slots: slot idx of each vector at [idx * R, idx * R + R)
        const size_t base = vb + (size_t)r0 * BS * R;
        const double* Bq = a.B + base;
        double* Xq = a.X + base;
        double* Rq = a.Rv + base;
        double* Pq = a.P + base;
        double* Zq = a.Z + base;
        double* APq = a.AP + base;
        const double* Pg = a.P + vb;   // P of the group, by global row (SpMV gathers)
        // x = 0, r = b, ||b||^2
        double acc[2 * R];
#pragma unroll
        for (int k = 0; k < 2 * R; ++k) acc[k] = 0.0;
        for_slots<U, V1<R>>(
            nslot, [&](int idx, V1<R>& v) { ld<R>(Bq + idx * R, v.a); },
            [&](int idx, V1<R>& v) {
                double zero[R];
#pragma unroll
                for (int k = 0; k < R; ++k) {
                    zero[k] = 0.0;
                    acc[k] += v.a[k] * v.a[k];
                }
                st<R>(Xq + idx * R, zero);
                st<R>(Rq + idx * R, v.a);
            });
        __syncthreads();
        // z = M^{-1} r  (forward sweep with L, then backward with L^T)
        apply_ichol0<BS, R, K>(a, s, vb);
        // p = z; rz = r.z; ||z||^2
        double rzz[2 * R];
#pragma unroll
        for (int k = 0; k < 2 * R; ++k) rzz[k] = 0.0;
        for_slots<U, V2<R>>(
            nslot,
            [&](int idx, V2<R>& v) {
                ld<R>(Zq + idx * R, v.a);
                ld<R>(Rq + idx * R, v.b);
            },
            [&](int idx, V2<R>& v) {
                st<R>(Pq + idx * R, v.a);
#pragma unroll
                for (int k = 0; k < R; ++k) {
                    rzz[k] += v.b[k] * v.a[k];
                    rzz[R + k] += v.a[k] * v.a[k];
                }
            });
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
                    const double* xc = Pg + a.ci[p] * BS * R;
#pragma unroll
                    for (int j = 0; j < BS; ++j) {
                        const double aj = blk[j];
                        double xv[R];
                        ld<R>(xc + j * R, xv);
#pragma unroll
                        for (int k = 0; k < R; ++k) y[k] = fma(aj, xv[k], y[k]);
                    }
                }
                double pv[R];
                ld<R>(Pq + idx * R, pv);
                st<R>(APq + idx * R, y);
#pragma unroll
                for (int k = 0; k < R; ++k) pap[k] += pv[k] * y[k];
            }
            block_sum<R>(pap, red);
            double alpha[R], rr[R];
#pragma unroll
            for (int k = 0; k < R; ++k) {
                alpha[k] = active[k] ? rz[k] / pap[k] : 0.0;
                rr[k] = 0.0;
            }
            for_slots<U, V4<R>>(
                nslot,
                [&](int idx, V4<R>& v) {
                    ld<R>(Xq + idx * R, v.a);
                    ld<R>(Pq + idx * R, v.b);
                    ld<R>(APq + idx * R, v.c);
                    ld<R>(Rq + idx * R, v.d);
                },
                [&](int idx, V4<R>& v) {
#pragma unroll
                    for (int k = 0; k < R; ++k) {
                        if (!active[k]) continue;
                        v.a[k] = fma(alpha[k], v.b[k], v.a[k]);
                        v.d[k] = fma(-alpha[k], v.c[k], v.d[k]);
                        rr[k] += v.d[k] * v.d[k];
                    }
                    st<R>(Xq + idx * R, v.a);
                    st<R>(Rq + idx * R, v.d);
                });
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
            apply_ichol0<BS, R, K>(a, s, vb);
            double rzn[R];
#pragma unroll
            for (int k = 0; k < R; ++k) rzn[k] = 0.0;
            for_slots<U, V2<R>>(
                nslot,
                [&](int idx, V2<R>& v) {
                    ld<R>(Rq + idx * R, v.a);
                    ld<R>(Zq + idx * R, v.b);
                },
                [&](int, V2<R>& v) {
#pragma unroll
                    for (int k = 0; k < R; ++k) rzn[k] += v.a[k] * v.b[k];
                });
            block_sum<R>(rzn, red);
            double beta[R];
#pragma unroll
            for (int k = 0; k < R; ++k) {
                beta[k] = rzn[k] / rz[k];
                if (active[k]) rz[k] = rzn[k];
            }
            for_slots<U, V2<R>>(
                nslot,
                [&](int idx, V2<R>& v) {
                    ld<R>(Pq + idx * R, v.a);
                    ld<R>(Zq + idx * R, v.b);
                },
                [&](int idx, V2<R>& v) {
#pragma unroll
                    for (int k = 0; k < R; ++k)
                        if (active[k]) v.a[k] = fma(beta[k], v.a[k], v.b[k]);
                    st<R>(Pq + idx * R, v.a);
                });
            __syncthreads();
        }
        if (tid == 0) {
#pragma unroll
            for (int k = 0; k < R; ++k) {
                const size_t o = (size_t)s * a.nrhs + k0 + k;
                a.iters[o] = its[k];
                a.res[2 * o] = res[k];
                a.res[2 * o + 1] = ref[k];
            }
        }
        __syncthreads();
    }
}

// RHS-major (k, row*bs + c) <-> grouped ((k / R) * len * R + (row*bs + c) * R + k % R)
__global__ void k_interleave(int64_t len, int nrhs, int R, const double* __restrict__ src, double* __restrict__ dst,
                             int to_interleaved) {
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= len * nrhs) return;
    const int64_t grp = t / (len * R);
    const int64_t rem = t - grp * len * R;
    const int64_t e = rem / R;
    const int k = (int)(grp * R + (rem - e * R));
    if (to_interleaved) dst[t] = src[(size_t)k * len + e];
    else dst[(size_t)k * len + e] = src[t];
}

// ELL copy of one sweep: thread t covers element e of entry k of slot q (k == K: the
// slot's inverse diagonal block).
__global__ void k_ell_gather(int64_t nslots, int K, int b2, const int* __restrict__ rows, const int* __restrict__ rp,
                             const int* __restrict__ ci, const double* __restrict__ val,
                             const double* __restrict__ dinv, int* __restrict__ ecol, double* __restrict__ eval,
                             double* __restrict__ edinv) {
    const int64_t per = (int64_t)(K + 1) * b2;
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= nslots * per) return;
    const int64_t q = t / per;
    const int r = (int)(t - q * per);
    const int k = r / b2, e = r - k * b2;
    const int row = rows[q];
    if (k == K) {
        edinv[q * b2 + e] = dinv[(size_t)row * b2 + e];
        return;
    }
    const int p = rp[row] + k;
    const bool have = p < rp[row + 1];
    eval[(q * K + k) * b2 + e] = have ? val[(size_t)p * b2 + e] : 0.0;
    if (e == 0) ecol[q * K + k] = have ? ci[p] : -1;
}

template <int BS, int R>
const void* pick_k(int K) {
    if constexpr (BS > kEllMaxBs) {
        return K == 0 ? reinterpret_cast<const void*>(&k_batch_pcg<BS, R, 0>) : nullptr;
    } else {
        switch (K) {
            case 0: return reinterpret_cast<const void*>(&k_batch_pcg<BS, R, 0>);
            case 2: return reinterpret_cast<const void*>(&k_batch_pcg<BS, R, 2>);
            case 3: return reinterpret_cast<const void*>(&k_batch_pcg<BS, R, 3>);
            case 4: return reinterpret_cast<const void*>(&k_batch_pcg<BS, R, 4>);
        }
        return nullptr;
    }
}

template <int BS>
const void* pick_r(int R, int K) {
    switch (R) {
        case 1: return pick_k<BS, 1>(K);
        case 2: return pick_k<BS, 2>(K);
        case 4: return pick_k<BS, 4>(K);
    }
    return nullptr;
}

const void* pick(int bs, int R, int K) {
    switch (bs) {
        case 1: return pick_r<1>(R, K);
        case 2: return pick_r<2>(R, K);
        case 4: return pick_r<4>(R, K);
        case 8: return pick_r<8>(R, K);
    }
    return nullptr;
}

}  // namespace

bool supported(int bs, int R) { return pick(bs, R, 0) != nullptr; }

int ell_width(int bs, int max_entries) {
    if (bs > kEllMaxBs) return 0;
    return max_entries <= 2 ? 2 : max_entries <= 4 ? max_entries : 0;
}

int max_resident_blocks(const bsp_ctx* ctx, int bs, int R, int K, int wg) {
    const void* k = pick(bs, R, K);
    if (!k) return 0;
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, wg, 0) != hipSuccess || per_cu < 1) per_cu = 1;
    return per_cu * ctx->num_cus;
}

int launch_pcg(const bsp_ctx* ctx, int bs, int R, int wg, int grid, const Args& a) {
    const void* k = pick(bs, R, a.ell_k);
    BSP_CHECK(k != nullptr, BSP_E_UNSUPPORTED, "batched pCG supports block sizes 1, 2, 4, 8 and 1, 2, 4 RHS per item");
    BSP_CHECK(wg % 64 == 0 && wg >= 64 && wg <= 1024, BSP_E_INVALID, "workgroup size must be a multiple of 64 <= 1024");
    void* args[] = {const_cast<Args*>(&a)};
    BSP_HIP(hipLaunchKernel(k, dim3(grid), dim3(wg), args, 0, ctx->stream));
    return BSP_OK;
}

int launch_ell_gather(const bsp_ctx* ctx, int64_t nslots, int K, int bs, const int* rows, const int* rp, const int* ci,
                      const double* val, const double* dinv, int* ecol, double* eval, double* edinv) {
    const int64_t total = nslots * (K + 1) * bs * bs;
    if (total <= 0) return BSP_OK;
    hipLaunchKernelGGL(k_ell_gather, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, ctx->stream, nslots, K,
                       bs * bs, rows, rp, ci, val, dinv, ecol, eval, edinv);
    BSP_HIP(hipGetLastError());
    return BSP_OK;
}

int launch_interleave(const bsp_ctx* ctx, int64_t len, int nrhs, int R, const double* src, double* dst,
                      bool to_interleaved) {
    const int64_t total = len * nrhs;
    if (total <= 0) return BSP_OK;
    BSP_CHECK(R >= 1 && nrhs % R == 0, BSP_E_INVALID, "RHS groups must divide the number of right-hand sides");
    hipLaunchKernelGGL(k_interleave, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, ctx->stream, len, nrhs, R,
                       src, dst, (int)to_interleaved);
    BSP_HIP(hipGetLastError());
    return BSP_OK;
}

}  // namespace batch
}  // namespace bsp
