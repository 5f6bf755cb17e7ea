// Standalone Krylov solves on a caller's SparseMatrix (numeric.hip).
#pragma once

#include "mpf_internal.h"

struct mpf_ctx;

namespace mpf {

int krylov_solve(mpf_ctx* ctx, int n, int density, const uint32_t* d_cid, const float* d_val, int R,
                 const float* d_B, int ldb, float* d_X, int ldx, int dc_free, const mpf_solve_opts& o,
                 mpf_solve_stats* stats);

}  // namespace mpf
