// Coupled complete electrode model: Schur complement on the electrode potentials and the
// pattern fields (cem.hip).
#pragma once

#include "jobs.h"

namespace mpf {

struct CemDev {
    const int32_t* bc_off;    // [L + 1] unit electrode sources Bc, CSR over ports
    const int32_t* bc_row;    // device rows
    const double* bc_val;
    const double* elen;       // [L] electrode lengths
    const double* pi;         // [L][R] Pi = [I Dpat | Mpat]
    const int32_t* prep;      // [pRu] representative pattern column of each unique column
    double coef[kMaxSlices];  // n_c (h beta_c)^2 / z'
    double tsc[kMaxSlices];   // h beta_c / z'
    double zp;
    int L, R, D, M, Dp, Mp, pRu, pRp;
};

// voltages (U32 / U64, nullable), T [frames][C][L][pRp] and the pattern fields
// phi [frames][C][np][pRp] from the unit electrode fields in the jobs' x64 (use64) or x
void launch_cem(const JobDev* d_jobs, int frames, int C, int np, const CemDev& cd, int use64, float* U32,
                double* U64, double* T, double* phi, hipStream_t st);

}  // namespace mpf
