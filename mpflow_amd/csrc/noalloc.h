// Allocation-free variant of the normal-equation solve (noalloc.hip).
#pragma once

#include "mpf_internal.h"

struct mpf_ctx;

namespace mpf {

// device address of this GPU's fixed work space; resolved once per device with that
// device current (mpf_ctx_create calls it)
void* noalloc_scratch(int device);

// mpf_normal_solve on the fixed work space: rows <= 1024, cols <= 65536, rows and ldj
// multiples of 4.  Synchronises.
void normal_solve_noalloc(mpf_ctx* ctx, int rows, int cols, const float* d_J, int ldj, double shift,
                          const double* d_r, int gn_refine, double* d_delta);

}  // namespace mpf
