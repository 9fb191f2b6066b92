// bsp_factor.hpp -- device block ILU(0) / IChol(0) factorisation (bsp_factor.hip).
#pragma once

#include <type_traits>

#include "bsp_internal.hpp"

namespace bsp {

// Symbolic plan of a factorisation in a given elimination order (pos[i] = step of row i),
// built on the host from A's structure.  Row k's tasks are its strict-lower blocks (k, i)
// (pos[i] < pos[k]) in elimination order; task t turns block task_p[t] into L_ki and applies
// the updates upd[task_u[t] .. task_u[t + 1]) to row k: block upd_dst -= L_ki * upd_src
// (ILU0: upd_src = (i, j)) or L_ki * upd_src^T (IChol0: upd_src = (j, i)).  Rows are factored
// level by level; level(k) = 1 + max level(i) over its strict-lower blocks.
struct FactorPlan {
    int n = 0;
    int64_t ntasks = 0, nupd = 0;
    std::vector<int32_t> level_ptr;   // host: rows[level_ptr[l] .. level_ptr[l + 1]) form level l
    DevMem rows, row_task, task_p, task_i, task_u, upd_src, upd_dst, diag;
};

bool factor_device_supported(int bs);
int plan_factor(bsp_ctx* ctx, const bsp_bsr* A, const std::vector<int32_t>& pos, bool chol, FactorPlan& P);
// F <- the factor in A's block layout (ILU0: L strict lower, U upper with diagonal; IChol0: L on
// the blocks with pos[j] <= pos[k], other blocks untouched); dinv <- inverse of each pivot block
// (ILU0: U_kk^{-1}, IChol0: L_kk^{-1}), FP64.
int factor_device(bsp_ctx* ctx, const bsp_bsr* A, const FactorPlan& P, bool chol, DevMem& F, DevMem& dinv);
// out[k] = inverse of F's block diag[k] (out_t, optional: its transpose), in `precision`
// (BSP_PREC_F64 / F32 / BF16).
int invert_diag_device(bsp_ctx* ctx, int n, int bs, const int* diag, const double* F, void* out, void* out_t,
                       int precision);
// out[q] = F's block map[q] (transposed if asked), converted to `precision`.
int gather_blocks(bsp_ctx* ctx, int64_t cnt, int bs, const int* map, const double* F, void* out, int precision,
                  bool transpose);

}  // namespace bsp
