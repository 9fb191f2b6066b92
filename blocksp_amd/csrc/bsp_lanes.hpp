// bsp_lanes.hpp -- moving a value between the lanes of one block row of a wavefront.
#pragma once

#include <hip/hip_runtime.h>

namespace bsp {

// v from lane (lane - lane % BS + c) of the wavefront, for a block size that divides a quad
// (DPP quad permutation) or any other (bpermute).
template <int BS>
__device__ __forceinline__ double block_lane(double v, int c, int src_lane) {
    if constexpr (BS == 1) {
        return v;
    } else if constexpr (BS == 2 || BS == 4) {
        const int2 w = __builtin_bit_cast(int2, v);
        int lo, hi;
        // quad_perm: lane j of each quad reads lane (j - j % BS + c) of that quad
#define BSP_QP(c0, c1, c2, c3) ((c0) | ((c1) << 2) | ((c2) << 4) | ((c3) << 6))
        switch (BS * 4 + c) {
            case 8: lo = __builtin_amdgcn_mov_dpp(w.x, BSP_QP(0, 0, 2, 2), 0xf, 0xf, false);
                    hi = __builtin_amdgcn_mov_dpp(w.y, BSP_QP(0, 0, 2, 2), 0xf, 0xf, false); break;
            case 9: lo = __builtin_amdgcn_mov_dpp(w.x, BSP_QP(1, 1, 3, 3), 0xf, 0xf, false);
                    hi = __builtin_amdgcn_mov_dpp(w.y, BSP_QP(1, 1, 3, 3), 0xf, 0xf, false); break;
            case 16: lo = __builtin_amdgcn_mov_dpp(w.x, BSP_QP(0, 0, 0, 0), 0xf, 0xf, false);
                     hi = __builtin_amdgcn_mov_dpp(w.y, BSP_QP(0, 0, 0, 0), 0xf, 0xf, false); break;
            case 17: lo = __builtin_amdgcn_mov_dpp(w.x, BSP_QP(1, 1, 1, 1), 0xf, 0xf, false);
                     hi = __builtin_amdgcn_mov_dpp(w.y, BSP_QP(1, 1, 1, 1), 0xf, 0xf, false); break;
            case 18: lo = __builtin_amdgcn_mov_dpp(w.x, BSP_QP(2, 2, 2, 2), 0xf, 0xf, false);
                     hi = __builtin_amdgcn_mov_dpp(w.y, BSP_QP(2, 2, 2, 2), 0xf, 0xf, false); break;
            default: lo = __builtin_amdgcn_mov_dpp(w.x, BSP_QP(3, 3, 3, 3), 0xf, 0xf, false);
                     hi = __builtin_amdgcn_mov_dpp(w.y, BSP_QP(3, 3, 3, 3), 0xf, 0xf, false); break;
        }
#undef BSP_QP
        return __builtin_bit_cast(double, make_int2(lo, hi));
    } else {
        return __shfl(v, src_lane, 64);
    }
}

template <int BS>
__device__ __forceinline__ float block_lane(float v, int c, int src_lane) {
    if constexpr (BS == 1) {
        return v;
    } else if constexpr (BS == 2 || BS == 4) {
        const double d = block_lane<BS>(__builtin_bit_cast(double, make_float2(v, 0.f)), c, src_lane);
        return __builtin_bit_cast(float2, d).x;
    } else {
        return __shfl(v, src_lane, 64);
    }
}

}  // namespace bsp
