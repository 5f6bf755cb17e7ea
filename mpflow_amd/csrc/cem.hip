// Coupled complete electrode model (cem_coupled): every electrode holds one potential V_l
// for all harmonics.  After the Krylov solves of the L unit electrode sources,
//   K_c Y_c = Bc          (K_c = A_c + Zb / z', one lane per electrode)
// the electrode potentials are eliminated by a Schur complement per frame,
//   S_V = sum_c n_c (h beta_c)^2 / z' (|e| - Bc^T Y_c / z'),   V = S_V^-1 I,
// and the fields of the pattern columns are combinations of the unit fields,
//   Phi_c = Y_c T_c,   T_c = (h beta_c / z') S_V^-1 Pi.
// No counterpart in the reference, whose electrodes use the gap model
// (src/fem/equation.cpp:183-215, src/models/eit.cpp:139-213).
#include "cem.h"

namespace mpf {

constexpr int kCemMaxPorts = 64;

// One block per frame: S_V in FP64 (each thread owns entries (l, l') and sums the
// harmonics in order), its inverse by Gauss-Jordan elimination (S_V is symmetric positive
// definite: no pivoting), the voltages U = Mpat^T S_V^-1 I Dpat and the combination
// matrices T_c.
__global__ void __launch_bounds__(256) k_cem_couple(const JobDev* __restrict__ jobs, CemDev cd, int use64,
                                                    float* __restrict__ U32, double* __restrict__ U64,
                                                    double* __restrict__ T) {
    const int f = blockIdx.x;
    const JobDev& J = jobs[f];
    const int L = cd.L, C = J.C;
    __shared__ double A[kCemMaxPorts][2 * kCemMaxPorts + 1];
    __shared__ double fac[kCemMaxPorts];
    __shared__ double Wpi[kCemMaxPorts][kCemMaxPorts];   // S_V^-1 Pi for the columns in use
    // S_V, and the identity beside it
    for (int i = threadIdx.x; i < L * L; i += blockDim.x) {
        const int l = i / L, l2 = i % L;
        double v = 0.0;
        for (int c = 0; c < C; ++c) {
            double g = 0.0;   // (Bc^T Y_c)[l][l2]
            for (int t = cd.bc_off[l]; t < cd.bc_off[l + 1]; ++t) {
                const size_t vi = vec_index(J, c, cd.bc_row[t], l2);
                g += cd.bc_val[t] * (use64 ? J.x64[vi] : (double)J.x[vi]);
            }
            v += cd.coef[c] * ((l == l2 ? cd.elen[l] : 0.0) - g / cd.zp);
        }
        A[l][l2] = v;
        A[l][L + l2] = l == l2 ? 1.0 : 0.0;
    }
    __syncthreads();
    if (C == 1) {
        // one harmonic: S_V has the constant null vector; currents summing to zero are
        // consistent, and S_V + s 1 1^T gives their solution with 1^T V = 0
        double tr = 0.0;
        for (int l = 0; l < L; ++l) tr += A[l][l];
        __syncthreads();
        for (int i = threadIdx.x; i < L * L; i += blockDim.x) A[i / L][i % L] += tr / L;
        __syncthreads();
    }
    for (int k = 0; k < L; ++k) {
        const double piv = A[k][k];
        __syncthreads();
        for (int j = threadIdx.x; j < 2 * L; j += blockDim.x) A[k][j] /= piv;
        __syncthreads();
        for (int i = threadIdx.x; i < L; i += blockDim.x) fac[i] = A[i][k];
        __syncthreads();
        for (int i = threadIdx.x; i < L * 2 * L; i += blockDim.x) {
            const int r = i / (2 * L), j = i % (2 * L);
            if (r != k) A[r][j] -= fac[r] * A[k][j];
        }
        __syncthreads();
    }
    // V = S_V^-1 (I Dpat): U[d Mp + m] = sum_l Mpat[l][m] V[l][d]
    for (int i = threadIdx.x; i < L * cd.D; i += blockDim.x) {
        const int l = i / cd.D, d = i % cd.D;
        double v = 0.0;
        for (int l2 = 0; l2 < L; ++l2) v += A[l][L + l2] * cd.pi[(size_t)l2 * cd.R + d];
        Wpi[l][d] = v;
    }
    __syncthreads();
    const int mrows = cd.Mp * cd.Dp;
    for (int i = threadIdx.x; i < mrows; i += blockDim.x) {
        const int d = i / cd.Mp, m = i % cd.Mp;
        double v = 0.0;
        if (d < cd.D && m < cd.M)
            for (int l = 0; l < L; ++l) v += cd.pi[(size_t)l * cd.R + cd.D + m] * Wpi[l][d];
        if (U32) U32[(size_t)f * mrows + i] = (float)v;
        if (U64) U64[(size_t)f * mrows + i] = v;
    }
    __syncthreads();
    // T_c[l][u] = (h beta_c / z') (S_V^-1 Pi)[l][rep_u]
    for (int i = threadIdx.x; i < L * cd.pRu; i += blockDim.x) {
        const int l = i / cd.pRu, u = i % cd.pRu;
        double v = 0.0;
        for (int l2 = 0; l2 < L; ++l2) v += A[l][L + l2] * cd.pi[(size_t)l2 * cd.R + cd.prep[u]];
        Wpi[l][u] = v;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < C * L * cd.pRp; i += blockDim.x) {
        const int c = i / (L * cd.pRp), r = i % (L * cd.pRp), l = r / cd.pRp, u = r % cd.pRp;
        T[(((size_t)f * C + c) * L + l) * cd.pRp + u] = u < cd.pRu ? cd.tsc[c] * Wpi[l][u] : 0.0;
    }
}

// Phi_c[row][u] = sum_l Y_c[row][l] T_c[l][u] (FP64) for every (frame, harmonic, tile):
// the tile's rows of Y_c are staged in LDS in chunks of 32 rows (one coalesced read of
// each row's L unit fields), T_c once per block; every output then reads its row of Y
// and its column of T from LDS.
constexpr int kCemRows = 32;   // rows per LDS chunk (dynamic LDS stays below 64 KiB for L, pRp <= 64)

__global__ void __launch_bounds__(256) k_cem_fields(const JobDev* __restrict__ jobs, CemDev cd, int use64,
                                                    const double* __restrict__ T, double* __restrict__ phi) {
    const int tile = blockIdx.x, c = blockIdx.y, f = blockIdx.z;
    const JobDev& J = jobs[f];
    const int L = cd.L, pRp = cd.pRp, ly = L + 1;
    // dynamic LDS: T [L][pRp], then the chunk of Y [kCemRows][L + 1]
    extern __shared__ double sm[];
    double* Ts = sm;
    double* Ys = sm + L * pRp;
    for (int i = threadIdx.x; i < L * pRp; i += blockDim.x) Ts[i] = T[(((size_t)f * J.C + c) * L) * pRp + i];
    double* out = phi + ((size_t)f * J.C + c) * (size_t)J.np * pRp;
    for (int r0 = tile * kTileRows; r0 < (tile + 1) * kTileRows; r0 += kCemRows) {
        __syncthreads();   // Ts written / the previous chunk's readers done
        for (int i = threadIdx.x; i < kCemRows * L; i += blockDim.x) {
            const int rr = i / L, l = i % L, row = r0 + rr;
            double y = 0.0;
            if (row < J.n) {
                const size_t vi = vec_index(J, c, row, l);
                y = use64 ? J.x64[vi] : (double)J.x[vi];
            }
            Ys[rr * ly + l] = y;
        }
        __syncthreads();
        for (int i = threadIdx.x; i < kCemRows * pRp; i += blockDim.x) {
            const int rr = i / pRp, u = i % pRp, row = r0 + rr;
            double v = 0.0;
            if (row < J.n && u < cd.pRu)
                for (int l = 0; l < L; ++l) v += Ys[rr * ly + l] * Ts[l * pRp + u];
            out[(size_t)row * pRp + u] = v;
        }
    }
}

void launch_cem(const JobDev* d_jobs, int frames, int C, int np, const CemDev& cd, int use64, float* U32,
                double* U64, double* T, double* phi, hipStream_t st) {
    MPF_REQUIRE(cd.L <= kCemMaxPorts && cd.pRp <= kCemMaxPorts && cd.D <= kCemMaxPorts,
                "coupled electrode model: at most 64 ports and pattern columns");
    k_cem_couple<<<frames, 256, 0, st>>>(d_jobs, cd, use64, U32, U64, T);
    const size_t lds = sizeof(double) * ((size_t)cd.L * cd.pRp + (size_t)kCemRows * (cd.L + 1));
    k_cem_fields<<<dim3(np / kTileRows, C, frames), 256, lds, st>>>(d_jobs, cd, use64, T, phi);
    MPF_HIP(hipGetLastError());
}

}  // namespace mpf
