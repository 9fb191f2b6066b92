// bsp_factor.hpp -- device block ILU(0) / IChol(0) factorisation (bsp_factor.hip).
#pragma once

#include <type_traits>

#include "bsp_internal.hpp"

namespace bsp {

bool factor_device_supported(int bs);
// Block ILU(0) (chol = false) or IChol(0) (chol = true) of A in the elimination order pos_dev
// (device, pos[i] = step of row i), factored level by level over `lower`, the forward sweep's
// schedule (rows of the strict-lower pattern by elimination position: the same dependencies).
// F <- the factor in A's block layout (ILU0: L strict lower, U upper with diagonal; IChol0: L on
// the blocks with pos[j] <= pos[k], other blocks untouched); dinv <- inverse of each pivot block
// (ILU0: U_kk^{-1}, IChol0: L_kk^{-1}), FP64; diag <- position of each row's diagonal block.
int factor_device(bsp_ctx* ctx, const bsp_bsr* A, const int* pos_dev, const Sweep& lower, bool chol, DevMem& F,
                  DevMem& dinv, DevMem& diag);
// out[k] = inverse of F's block diag[k] (out_t, optional: its transpose), in `precision`
// (BSP_PREC_F64 / F32 / BF16).
int invert_diag_device(bsp_ctx* ctx, int n, int bs, const int* diag, const double* F, void* out, void* out_t,
                       int precision);
// out[q] = F's block map[q] (transposed if asked), converted to `precision`.
int gather_blocks(bsp_ctx* ctx, int64_t cnt, int bs, const int* map, const double* F, void* out, int precision,
                  bool transpose);

}  // namespace bsp
