// Standalone Krylov solves on a caller's SparseMatrix (numeric.hip).
#pragma once

#include "mpf_internal.h"

struct mpf_ctx;

namespace mpf {

int krylov_solve(mpf_ctx* ctx, int n, int density, const uint32_t* d_cid, const float* d_val, int R,
                 const float* d_B, int ldb, float* d_X, int ldx, int dc_free, const mpf_solve_opts& o,
                 mpf_solve_stats* stats);

// SparseMatrix::multiply on the reference layout (synchronises)
void sparse_multiply(mpf_ctx* ctx, const float* d_values, const uint32_t* d_cid, const float* d_X,
                     int result_rows, int matrix_rows, int columns, int density, float* d_Y);
// Y = A X on the product layout for one matrix with values S and a zero R term
// (`zero`), Rp lanes; enqueued on st (no synchronisation); an id outside [0, np) sets *bad
void ell_spmv_product(const int32_t* ids, const float* S, const float* zero, int np, int width, int Rp,
                      const float* X, float* Y, int* bad, hipStream_t st);
// HBM streaming rate (GB/s, best of reps launches) of a grid-stride 128-bit kernel over
// n floats: pattern 0 y = x, pattern 1 y = x + 0.5 y; `loads` float4 in flight per thread
double stream_probe(int pattern, int loads, const float* x, float* y, size_t n, int blocks, int reps,
                    hipStream_t st);
// multi-harmonic S/R SpMV on the product layout; alpha is a host array of C values
// (synchronises)
void ell_spmv(mpf_ctx* ctx, int np, int width, const int32_t* d_ids, const float* d_S, const float* d_R, int C,
              const float* alpha, const float* d_X, float* d_Y, int ld);

}  // namespace mpf
