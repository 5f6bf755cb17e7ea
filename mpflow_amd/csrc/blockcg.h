// Block preconditioned conjugate gradients over the right-hand sides of one matrix
// (blockcg.hip).
#pragma once

#include "mpf_internal.h"

namespace mpf {

// One matrix on the product layout (slot-major ELL [wpad][np], values S, R term zero),
// R <= Rp right-hand sides in the FP32 vector `rhs` [np][Rp]; x receives the solution
// from 0 (FP32 [np][Rp]); p and q are FP32 [np][Rp] work vectors (the search directions and
// A times them).  The recurrences for x, r and z run in FP64 buffers of the solver's own.  dc: the operator is A + 1 1^T.  bb_ref: b.b per
// lane the stop rule compares with (device FP64 [Rp]); null: the rhs's own column norms.
struct BcgArgs {
    int n = 0, np = 0, R = 0, Rp = 0, wpad = 0, dc = 0;
    const int32_t* ids = nullptr;
    const float* s = nullptr;
    const float* zero = nullptr;   // [wpad][np] zeros (the R term of the SpMV)
    const float* dinv = nullptr;   // [np] Jacobi inverse
    const float* rhs = nullptr;
    float *x = nullptr, *p = nullptr, *q = nullptr;
    int* bad = nullptr;            // SpMV index check flag
    const double* bb_ref = nullptr;
};

struct BcgResult {
    int steps = 0;        // block iterations
    int rank = 0;         // directions kept in the last iteration
    bool capped = false;  // hit max_iter
    bool broke = false;   // no independent direction left, or P^T A P not positive definite
};

// Runs on `st` and synchronises at every convergence poll (every check_every iterations).
BcgResult block_pcg(const BcgArgs& a, int rule, double tol, int max_iter, int check_every, hipStream_t st);

}  // namespace mpf
