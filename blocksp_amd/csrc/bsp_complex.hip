// bsp_complex.hip -- complex (std::complex<double>) vector kernels of the complex pGMRES.
//
// A complex vector is stored interleaved (re, im), the memory of std::complex<double>, so it
// is also the real vector that the real-equivalent matrix (each complex entry a + ib held as
// the 2 x 2 block [[a, -b], [b, a]]) multiplies: the products and preconditioner sweeps of a
// complex system are the real kernels.  What a real kernel cannot do is the complex inner
// product and the complex linear combinations of Arnoldi and of the deflated restart:
//   * k_zdots:   partial sums of (q_j, v) = sum_e conj(q_j[e]) v[e] for m basis vectors at
//                once (grid: blocks x m), one partial per workgroup in a fixed order -- the
//                caller adds the partials in workgroup order, so results do not depend on
//                scheduling;
//   * k_zupdate: v[e] += sign * sum_j h_j q_j[e] (complex h, j in order);
//   * k_zrotate: Qout_i[e] = sum_j Q_j[e] P[j][i] (complex P, row-major m x kp).
// The basis is one contiguous buffer, vector j at Q + j * stride (doubles).
#include "bsp_internal.hpp"

namespace bsp {

namespace {

constexpr int kT = 256;
constexpr int kWaveZ = 64;

__global__ __launch_bounds__(kT) void k_zdots(int64_t nc, int m, const double* __restrict__ Q, int64_t stride,
                                               const double* __restrict__ v, int nblk, double* __restrict__ part) {
    const int j = blockIdx.y;
    const double2* q = reinterpret_cast<const double2*>(Q + (size_t)j * stride);
    const double2* x = reinterpret_cast<const double2*>(v);
    const int64_t per = (nc + nblk - 1) / nblk;
    const int64_t e0 = (int64_t)blockIdx.x * per, e1 = min(nc, e0 + per);
    double re = 0.0, im = 0.0;
    for (int64_t e = e0 + threadIdx.x; e < e1; e += kT) {
        const double2 a = q[e], b = x[e];
        // conj(a) * b
        re = fma(a.x, b.x, re);
        re = fma(a.y, b.y, re);
        im = fma(a.x, b.y, im);
        im = fma(-a.y, b.x, im);
    }
    for (int d = kWaveZ / 2; d > 0; d >>= 1) {
        re += __shfl_down(re, d, kWaveZ);
        im += __shfl_down(im, d, kWaveZ);
    }
    __shared__ double sre[kT / kWaveZ], sim[kT / kWaveZ];
    const int w = threadIdx.x / kWaveZ;
    if ((threadIdx.x % kWaveZ) == 0) {
        sre[w] = re;
        sim[w] = im;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = 0.0, s = 0.0;
        for (int t = 0; t < kT / kWaveZ; ++t) {
            r += sre[t];
            s += sim[t];
        }
        part[((size_t)j * nblk + blockIdx.x) * 2 + 0] = r;
        part[((size_t)j * nblk + blockIdx.x) * 2 + 1] = s;
    }
}

__global__ __launch_bounds__(kT) void k_zupdate(int64_t nc, int m, const double* __restrict__ Q, int64_t stride,
                                                 const double* __restrict__ h, double sign, double* __restrict__ v) {
    const int64_t e = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (e >= nc) return;
    double re = 0.0, im = 0.0;
    for (int j = 0; j < m; ++j) {
        const double2 q = reinterpret_cast<const double2*>(Q + (size_t)j * stride)[e];
        const double hr = h[2 * j], hi = h[2 * j + 1];
        re = fma(hr, q.x, re);
        re = fma(-hi, q.y, re);
        im = fma(hr, q.y, im);
        im = fma(hi, q.x, im);
    }
    double2* x = reinterpret_cast<double2*>(v);
    double2 o = x[e];
    o.x = fma(sign, re, o.x);
    o.y = fma(sign, im, o.y);
    x[e] = o;
}

__global__ __launch_bounds__(kT) void k_zrotate(int64_t nc, int m, int kp, const double* __restrict__ Q,
                                                 int64_t stride, const double* __restrict__ P,
                                                 double* __restrict__ Qout, int64_t ostride) {
    const int64_t e = (int64_t)blockIdx.x * kT + threadIdx.x;
    if (e >= nc) return;
    for (int i = 0; i < kp; ++i) {
        double re = 0.0, im = 0.0;
        for (int j = 0; j < m; ++j) {
            const double2 q = reinterpret_cast<const double2*>(Q + (size_t)j * stride)[e];
            const double pr = P[2 * ((size_t)j * kp + i)], pi = P[2 * ((size_t)j * kp + i) + 1];
            re = fma(q.x, pr, re);
            re = fma(-q.y, pi, re);
            im = fma(q.x, pi, im);
            im = fma(q.y, pr, im);
        }
        reinterpret_cast<double2*>(Qout + (size_t)i * ostride)[e] = make_double2(re, im);
    }
}

inline unsigned zgrid(int64_t nc) { return (unsigned)((nc + kT - 1) / kT); }

}  // namespace

int zdots_blocks(int64_t nc) {
    const int64_t b = (nc + 4 * kT - 1) / (4 * kT);
    return (int)std::max<int64_t>(1, std::min<int64_t>(b, 128));
}

int launch_zdots(const bsp_ctx* ctx, int64_t nc, int m, const double* Q, int64_t stride, const double* v,
                 double* part) {
    if (nc <= 0 || m <= 0) return BSP_OK;
    const int nblk = zdots_blocks(nc);
    hipLaunchKernelGGL(k_zdots, dim3(nblk, m), dim3(kT), 0, ctx->stream, nc, m, Q, stride, v, nblk, part);
    BSP_HIP(hipGetLastError());
    return BSP_OK;
}

int launch_zupdate(const bsp_ctx* ctx, int64_t nc, int m, const double* Q, int64_t stride, const double* h,
                   double sign, double* v) {
    if (nc <= 0 || m <= 0) return BSP_OK;
    hipLaunchKernelGGL(k_zupdate, dim3(zgrid(nc)), dim3(kT), 0, ctx->stream, nc, m, Q, stride, h, sign, v);
    BSP_HIP(hipGetLastError());
    return BSP_OK;
}

int launch_zrotate(const bsp_ctx* ctx, int64_t nc, int m, int kp, const double* Q, int64_t stride, const double* P,
                   double* Qout, int64_t ostride) {
    if (nc <= 0 || m <= 0 || kp <= 0) return BSP_OK;
    hipLaunchKernelGGL(k_zrotate, dim3(zgrid(nc)), dim3(kT), 0, ctx->stream, nc, m, kp, Q, stride, P, Qout, ostride);
    BSP_HIP(hipGetLastError());
    return BSP_OK;
}

}  // namespace bsp
