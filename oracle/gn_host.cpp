// CPU baseline of a whole reconstruction frame: the device path's algorithm on the
// host -- sigma-weighted assembly of S and R, for every harmonic FP32 two-level (or
// Jacobi) PCG with FP64 scalars, warm-started from the previous Gauss-Newton step's FP64
// fields and followed by FP64 residual refinement, the measured voltages, the exact 2.5D
// Jacobian, and the Tikhonov step in the m x m data space (FP32 Gram J J^T, FP64
// Cholesky, FP64-residual refinement passes).  One frame per OpenMP thread.
// TEST INFRASTRUCTURE ONLY: the timed CPU baseline of bench.py, checked against the FP64
// oracle in tests/.
//
// Reference algorithm: Solver::solveAbsolute (src/solver/solver.cpp:195-225) over
// EIT::solve (src/models/eit.cpp:139-213), ConjugateGradient::solve
// (src/numeric/conjugate_gradient.cpp:61-161) and Inverse::solve
// (src/solver/inverse.cpp:60-238); the Jacobian is the exact 2.5D derivative
// (the device path's jacobian_exact mode).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include <omp.h>

extern "C" {

struct GnhProblem {
    int N, E, k;                      // nodes (aggregates are ranges of consecutive rows), elements
    const int32_t* elements;          // [E][k]
    const double* Ke;                 // [E][k*k] stiffness integrals
    const double* Me;                 // [E][k*k] mass integrals
    int nnz;
    const int32_t* rowptr;            // [N+1] CSR pattern of the system matrix
    const int32_t* col;               // [nnz]
    const int32_t* emap;              // [E][k*k] CSR slot of each element entry
    const double* S0;                 // [nnz] sigma-independent part of S (complete-electrode-
                                      // model blocks), nullable
    int C;
    const double* alpha;              // [C] (2 pi c / h)^2
    const double* beta;               // [C] excitation scale
    const double* jw;                 // [C] Jacobian weight 1 / beta_c
    int Ru;                           // unique right-hand sides
    const double* F;                  // [N][Ru] unscaled excitation columns
    int D, M, Dp, Mp;
    const int32_t* colmap;            // [D+M] pattern column -> unique column
    const double* colscale;           // [D+M]
    const int32_t* w_off;             // [M+1] measurement rows (CSR)
    const int32_t* w_node;
    const double* w_val;
    double ref;
    int logarithmic;
    int na;                           // coarse aggregates (0: Jacobi alone)
    const int32_t* astart;            // [na+1] row ranges of the aggregates
    const float* cinv;                // [C][na][na] inverse coarse matrices at sigma = 1
    const double* d1;                 // [na] stiffness-diagonal sums at sigma = 1
    double tol, refine_tol;
    int refine;                       // FP64 refinement passes per forward solve
    int gn_refine;                    // FP64-residual passes of the normal-equation solve
    double lam2;                      // (reference lambda)^2
    // coupled complete electrode model (cem_coupled): the solves' right-hand sides F are
    // the L unit electrode sources Bc (no scaling), the electrode potentials are
    // eliminated per frame by a Schur complement, and the fields of the pattern's unique
    // columns are combinations of the unit fields (colmap / colscale then index those)
    int cem_coupled;
    int L, pRu;
    const int32_t* bc_off;            // [L+1] Bc as CSR over ports
    const int32_t* bc_row;
    const double* bc_val;
    const double* elen;               // [L] electrode lengths
    const double* pi;                 // [L][D+M] Pi = [I Dpat | Mpat]
    const int32_t* prep;              // [pRu] representative column of each unique pattern column
    const double* coef;               // [C] n_c (h beta_c)^2 / z'
    const double* tsc;                // [C] h beta_c / z'
    double zp;
};

int gnh_threads(void) { return omp_get_max_threads(); }

namespace {

struct Pcg {
    // work vectors [N][Ru]
    std::vector<float> r, p, q, x;
    std::vector<double> g;
    std::vector<float> y;
};

// z = D^-1 r + W Einv W^T r (two-level) for one column block; returns sum r.z per column
void precond(const GnhProblem& P, const std::vector<float>& dinv, const std::vector<float>& einv,
             const float* r, float* z, double* rz, std::vector<double>& g, std::vector<float>& y) {
    const int N = P.N, R = P.Ru, na = P.na;
    for (int i = 0; i < N; ++i)
        for (int c = 0; c < R; ++c) z[(size_t)i * R + c] = dinv[i] * r[(size_t)i * R + c];
    if (na) {
        std::fill(g.begin(), g.end(), 0.0);
        for (int a = 0; a < na; ++a)
            for (int i = P.astart[a]; i < P.astart[a + 1]; ++i)
                for (int c = 0; c < R; ++c) g[(size_t)a * R + c] += r[(size_t)i * R + c];
        for (int a = 0; a < na; ++a) {
            float* ya = &y[(size_t)a * R];
            for (int c = 0; c < R; ++c) ya[c] = 0.0f;
            const float* e = &einv[(size_t)a * na];
            for (int b = 0; b < na; ++b) {
                const float eb = e[b];
                const double* gb = &g[(size_t)b * R];
                for (int c = 0; c < R; ++c) ya[c] += eb * (float)gb[c];
            }
        }
        for (int a = 0; a < na; ++a)
            for (int i = P.astart[a]; i < P.astart[a + 1]; ++i)
                for (int c = 0; c < R; ++c) z[(size_t)i * R + c] += y[(size_t)a * R + c];
    }
    for (int c = 0; c < R; ++c) rz[c] = 0.0;
    for (int i = 0; i < N; ++i)
        for (int c = 0; c < R; ++c) rz[c] += (double)r[(size_t)i * R + c] * z[(size_t)i * R + c];
}

// FP32 PCG on (A + dc 1 1^T) x = b from x = 0, stop when every column has
// ||r|| <= tol * ref_norm[c]; returns steps
int pcg(const GnhProblem& P, const std::vector<float>& A, const std::vector<float>& dinv,
        const std::vector<float>& einv, bool dc, const float* b, const double* ref_bb, double tol, float* x,
        Pcg& w, int max_iter) {
    const int N = P.N, R = P.Ru;
    const size_t V = (size_t)N * R;
    w.r.assign(b, b + V);
    w.p.resize(V);
    w.q.resize(V);
    std::fill(x, x + V, 0.0f);
    std::vector<double> rho(R), nrho(R), pq(R), ps(R), rr(R);
    precond(P, dinv, einv, w.r.data(), w.p.data(), rho.data(), w.g, w.y);
    std::vector<float> z(V);
    std::vector<float> acc(R);
    int step = 0;
    for (; step < max_iter; ++step) {
        for (int c = 0; c < R; ++c) { pq[c] = 0.0; ps[c] = 0.0; rr[c] = 0.0; }
        for (int i = 0; i < N; ++i) {
            for (int c = 0; c < R; ++c) acc[c] = 0.0f;
            for (int s = P.rowptr[i]; s < P.rowptr[i + 1]; ++s) {
                const float v = A[s];
                const float* pr = &w.p[(size_t)P.col[s] * R];
                for (int c = 0; c < R; ++c) acc[c] += v * pr[c];
            }
            for (int c = 0; c < R; ++c) {
                const size_t o = (size_t)i * R + c;
                w.q[o] = acc[c];
                pq[c] += (double)w.p[o] * acc[c];
                ps[c] += w.p[o];
            }
        }
        std::vector<float> al(R), dcs(R);
        for (int c = 0; c < R; ++c) {
            if (dc) pq[c] += ps[c] * ps[c];
            al[c] = (float)(pq[c] != 0.0 ? rho[c] / pq[c] : 0.0);
            dcs[c] = dc ? (float)ps[c] : 0.0f;
        }
        for (size_t o = 0; o < V; ++o) {
            const int c = (int)(o % R);
            x[o] += al[c] * w.p[o];
            w.r[o] -= al[c] * (w.q[o] + dcs[c]);
        }
        precond(P, dinv, einv, w.r.data(), z.data(), nrho.data(), w.g, w.y);
        bool done = true;
        for (int i = 0; i < N; ++i)
            for (int c = 0; c < R; ++c) rr[c] += (double)w.r[(size_t)i * R + c] * w.r[(size_t)i * R + c];
        std::vector<float> be(R);
        for (int c = 0; c < R; ++c) {
            if (ref_bb[c] > 0.0 && rr[c] > tol * tol * ref_bb[c]) done = false;
            be[c] = (float)(rho[c] != 0.0 ? nrho[c] / rho[c] : 0.0);
            rho[c] = nrho[c];
        }
        if (done) return step + 1;
        for (size_t o = 0; o < V; ++o) w.p[o] = z[o] + be[o % R] * w.p[o];
    }
    return step;
}

// FP32 m x m Gram G = sum_e J_e J_e^T of the element-major Jacobian J [E][m] (the
// device's layout), lower triangle in 4 x 16 register tiles over 256-element chunks,
// chunk sums added in FP64, then mirrored
void gram(int m, int E, const float* J, float* G) {
    constexpr int EB = 256, IB = 4, JB = 16;
    std::vector<double> acc((size_t)m * m, 0.0);
    for (int e0 = 0; e0 < E; e0 += EB) {
        const int e1 = std::min(E, e0 + EB);
        for (int i0 = 0; i0 + IB <= m; i0 += IB)
            for (int j0 = 0; j0 < i0 + IB && j0 + JB <= m; j0 += JB) {
                float s[IB][JB] = {};
                for (int e = e0; e < e1; ++e) {
                    const float* Je = J + (size_t)e * m;
                    for (int u = 0; u < IB; ++u) {
                        const float a = Je[i0 + u];
                        for (int v = 0; v < JB; ++v) s[u][v] += a * Je[j0 + v];
                    }
                }
                for (int u = 0; u < IB; ++u)
                    for (int v = 0; v < JB; ++v) acc[(size_t)(i0 + u) * m + j0 + v] += s[u][v];
            }
    }
    for (int i = 0; i < m; ++i)
        for (int j = 0; j <= i; ++j) G[(size_t)i * m + j] = G[(size_t)j * m + i] = (float)acc[(size_t)i * m + j];
}

bool cholesky(int n, std::vector<double>& A) {
    for (int j = 0; j < n; ++j) {
        double d = A[(size_t)j * n + j];
        for (int t = 0; t < j; ++t) d -= A[(size_t)j * n + t] * A[(size_t)j * n + t];
        if (!(d > 0.0)) return false;
        d = std::sqrt(d);
        A[(size_t)j * n + j] = d;
        for (int i = j + 1; i < n; ++i) {
            double s = A[(size_t)i * n + j];
            for (int t = 0; t < j; ++t) s -= A[(size_t)i * n + t] * A[(size_t)j * n + t];
            A[(size_t)i * n + j] = s / d;
        }
    }
    return true;
}

void chol_solve(int n, const std::vector<double>& L, const double* b, double* y) {
    std::vector<double> z(b, b + n);
    for (int i = 0; i < n; ++i) {
        double s = z[i];
        for (int t = 0; t < i; ++t) s -= L[(size_t)i * n + t] * z[t];
        z[i] = s / L[(size_t)i * n + i];
    }
    for (int i = n - 1; i >= 0; --i) {
        double s = z[i];
        for (int t = i + 1; t < n; ++t) s -= L[(size_t)t * n + i] * z[t];
        z[i] = s / L[(size_t)i * n + i];
    }
    std::copy(z.begin(), z.end(), y);
}

// one frame: `steps` Gauss-Newton steps from gamma (in/out)
void frame(const GnhProblem& P, const double* Um, int steps, float* gamma, double* Uout, float* Jout) {
    const int N = P.N, E = P.E, k = P.k, R = P.Ru, C = P.C, kk = k * k;
    const int m = P.Dp * P.Mp;
    const size_t V = (size_t)N * R;
    std::vector<double> x64((size_t)C * V, 0.0), S(P.nnz), Rm(P.nnz), A64(P.nnz), res(V);
    std::vector<float> A32(P.nnz), dinv(N), einv((size_t)std::max(P.na, 1) * std::max(P.na, 1)), b32(V), dx(V);
    std::vector<double> bb(R), sigma(E), dsig(E), U(m), csc(std::max(P.na, 1));
    std::vector<float> J((size_t)m * E);
    std::vector<float> G((size_t)m * m);
    std::vector<double> L((size_t)m * m), r(m), y(m), dy(m), vE(E), res_m(m), delta(E);
    Pcg w;
    w.g.resize((size_t)std::max(P.na, 1) * R);
    w.y.resize((size_t)std::max(P.na, 1) * R);
    for (int step = 0; step < steps; ++step) {
        // sigma (matParam) and the assembly of S, R in FP64
        for (int e = 0; e < E; ++e) {
            const double g = gamma[e];
            sigma[e] = P.logarithmic ? P.ref * std::exp(std::log(10.0) * g / 10.0) : P.ref * g;
            dsig[e] = P.logarithmic ? sigma[e] * std::log(10.0) / 10.0 : P.ref;
        }
        if (P.S0) std::copy(P.S0, P.S0 + P.nnz, S.begin());
        else std::fill(S.begin(), S.end(), 0.0);
        std::fill(Rm.begin(), Rm.end(), 0.0);
        for (int e = 0; e < E; ++e)
            for (int t = 0; t < kk; ++t) {
                S[P.emap[(size_t)e * kk + t]] += sigma[e] * P.Ke[(size_t)e * kk + t];
                Rm[P.emap[(size_t)e * kk + t]] += sigma[e] * P.Me[(size_t)e * kk + t];
            }
        if (P.na)
            for (int a = 0; a < P.na; ++a) {
                double s = 0.0;
                for (int i = P.astart[a]; i < P.astart[a + 1]; ++i)
                    for (int q = P.rowptr[i]; q < P.rowptr[i + 1]; ++q)
                        if (P.col[q] == i) s += S[q];
                csc[a] = 1.0 / std::sqrt(s / P.d1[a]);
            }
        std::fill(U.begin(), U.end(), 0.0);
        std::fill(J.begin(), J.end(), 0.0f);
        for (int c = 0; c < C; ++c) {
            const bool dc = c == 0 && !P.cem_coupled;
            for (int q = 0; q < P.nnz; ++q) {
                A64[q] = S[q] + P.alpha[c] * Rm[q];
                A32[q] = (float)A64[q];
            }
            for (int i = 0; i < N; ++i)
                for (int q = P.rowptr[i]; q < P.rowptr[i + 1]; ++q)
                    if (P.col[q] == i) dinv[i] = A32[q] != 0.0f ? 1.0f / A32[q] : 0.0f;
            if (P.na)
                for (int a = 0; a < P.na; ++a)
                    for (int b = 0; b < P.na; ++b)
                        einv[(size_t)a * P.na + b] =
                            (float)(csc[a] * (double)P.cinv[((size_t)c * P.na + a) * P.na + b] * csc[b]);
            double* x = &x64[(size_t)c * V];
            for (int col = 0; col < R; ++col) {
                bb[col] = 0.0;
                for (int i = 0; i < N; ++i) {
                    const double f = P.beta[c] * P.F[(size_t)i * R + col];
                    bb[col] += f * f;
                }
            }
            // warm start (steps after the first) and refinement passes: correction solves of
            // the FP64 residual; the first solve stops at tol ||beta F||, later ones at
            // refine_tol of their own residual
            for (int pass = 0; pass <= P.refine; ++pass) {
                std::vector<double> cs(R, 0.0);
                if (dc)
                    for (int i = 0; i < N; ++i)
                        for (int col = 0; col < R; ++col) cs[col] += x[(size_t)i * R + col];
                std::vector<double> rb(R, 0.0);
                for (int i = 0; i < N; ++i) {
                    for (int col = 0; col < R; ++col) res[(size_t)i * R + col] = P.beta[c] * P.F[(size_t)i * R + col] - cs[col];
                    for (int q = P.rowptr[i]; q < P.rowptr[i + 1]; ++q) {
                        const double a = A64[q];
                        const double* xr = &x[(size_t)P.col[q] * R];
                        for (int col = 0; col < R; ++col) res[(size_t)i * R + col] -= a * xr[col];
                    }
                    for (int col = 0; col < R; ++col) {
                        b32[(size_t)i * R + col] = (float)res[(size_t)i * R + col];
                        rb[col] += res[(size_t)i * R + col] * res[(size_t)i * R + col];
                    }
                }
                const bool first = pass == 0;
                pcg(P, A32, dinv, einv, dc, b32.data(), first ? bb.data() : rb.data(), first ? P.tol : P.refine_tol,
                    dx.data(), w, 10 * N);
                for (size_t o = 0; o < V; ++o) x[o] += dx[o];
            }
        }
        // the fields the Jacobian contracts, one lane per (unique) pattern column: the
        // solved fields themselves, or (coupled electrodes) their combinations Y_c T_c
        std::vector<double> Tm;
        if (P.cem_coupled) {
            const int L = P.L;
            std::vector<double> A((size_t)L * 2 * L, 0.0);
            for (int l = 0; l < L; ++l)
                for (int l2 = 0; l2 < L; ++l2) {
                    double v = 0.0;
                    for (int c = 0; c < C; ++c) {
                        double g = 0.0;
                        for (int t = P.bc_off[l]; t < P.bc_off[l + 1]; ++t)
                            g += P.bc_val[t] * x64[(size_t)c * V + (size_t)P.bc_row[t] * R + l2];
                        v += P.coef[c] * ((l == l2 ? P.elen[l] : 0.0) - g / P.zp);
                    }
                    A[(size_t)l * 2 * L + l2] = v;
                    A[(size_t)l * 2 * L + L + l2] = l == l2 ? 1.0 : 0.0;
                }
            if (C == 1) {
                double tr = 0.0;
                for (int l = 0; l < L; ++l) tr += A[(size_t)l * 2 * L + l];
                for (int l = 0; l < L; ++l)
                    for (int l2 = 0; l2 < L; ++l2) A[(size_t)l * 2 * L + l2] += tr / L;
            }
            for (int kk2 = 0; kk2 < L; ++kk2) {
                const double piv = A[(size_t)kk2 * 2 * L + kk2];
                for (int j = 0; j < 2 * L; ++j) A[(size_t)kk2 * 2 * L + j] /= piv;
                for (int i = 0; i < L; ++i) {
                    if (i == kk2) continue;
                    const double f = A[(size_t)i * 2 * L + kk2];
                    for (int j = 0; j < 2 * L; ++j) A[(size_t)i * 2 * L + j] -= f * A[(size_t)kk2 * 2 * L + j];
                }
            }
            const int Rt = P.D + P.M;
            auto winv_pi = [&](int l, int col) {
                double v = 0.0;
                for (int l2 = 0; l2 < L; ++l2) v += A[(size_t)l * 2 * L + L + l2] * P.pi[(size_t)l2 * Rt + col];
                return v;
            };
            for (int d = 0; d < P.D; ++d) {
                std::vector<double> Vd(L);
                for (int l = 0; l < L; ++l) Vd[l] = winv_pi(l, d);
                for (int mm = 0; mm < P.M; ++mm) {
                    double v = 0.0;
                    for (int l = 0; l < L; ++l) v += P.pi[(size_t)l * Rt + P.D + mm] * Vd[l];
                    U[(size_t)d * P.Mp + mm] = v;
                }
            }
            Tm.assign((size_t)L * P.pRu, 0.0);
            for (int l = 0; l < L; ++l)
                for (int u = 0; u < P.pRu; ++u) Tm[(size_t)l * P.pRu + u] = winv_pi(l, P.prep[u]);
        }
        const int RJ = P.cem_coupled ? P.pRu : R;
        std::vector<double> phiall;
        for (int c = 0; c < C; ++c) {
            const double* xc = &x64[(size_t)c * V];
            const double* fields = xc;
            if (P.cem_coupled) {
                phiall.assign((size_t)N * RJ, 0.0);
                for (int i = 0; i < N; ++i)
                    for (int u = 0; u < RJ; ++u) {
                        double v = 0.0;
                        for (int l = 0; l < P.L; ++l) v += xc[(size_t)i * R + l] * Tm[(size_t)l * RJ + u];
                        phiall[(size_t)i * RJ + u] = P.tsc[c] * v;
                    }
                fields = phiall.data();
            } else {
                // measured voltages U[d * Mp + mm] += sum_node W[mm][node] Phi[node][d]
                for (int d = 0; d < P.D; ++d) {
                    const int ld = P.colmap[d];
                    const double sd = P.colscale[d];
                    for (int mm = 0; mm < P.M; ++mm) {
                        double s2 = 0.0;
                        for (int t = P.w_off[mm]; t < P.w_off[mm + 1]; ++t) s2 += P.w_val[t] * xc[(size_t)P.w_node[t] * R + ld];
                        U[(size_t)d * P.Mp + mm] += sd * s2;
                    }
                }
            }
            // exact Jacobian: J[(d, m), e] -= sigma'_e w_c Phi_d^T (K_e + alpha_c M_e) Phi_m
            std::vector<double> phi((size_t)k * RJ), T((size_t)k * RJ), Gl((size_t)RJ * RJ);
            for (int e = 0; e < E; ++e) {
                const int32_t* el = &P.elements[(size_t)e * k];
                for (int i = 0; i < k; ++i)
                    for (int col = 0; col < RJ; ++col) phi[(size_t)i * RJ + col] = fields[(size_t)el[i] * RJ + col];
                for (int i = 0; i < k; ++i)
                    for (int col = 0; col < RJ; ++col) {
                        double s2 = 0.0;
                        for (int j = 0; j < k; ++j)
                            s2 += (P.Ke[(size_t)e * kk + i * k + j] + P.alpha[c] * P.Me[(size_t)e * kk + i * k + j]) *
                                  phi[(size_t)j * RJ + col];
                        T[(size_t)i * RJ + col] = s2;
                    }
                for (int a = 0; a < RJ; ++a)
                    for (int b = 0; b < RJ; ++b) {
                        double s2 = 0.0;
                        for (int i = 0; i < k; ++i) s2 += phi[(size_t)i * RJ + a] * T[(size_t)i * RJ + b];
                        Gl[(size_t)a * RJ + b] = s2;
                    }
                const double wgt = -dsig[e] * P.jw[c];
                for (int d = 0; d < P.D; ++d)
                    for (int mm = 0; mm < P.M; ++mm)
                        J[(size_t)e * m + d * P.Mp + mm] +=
                            (float)(wgt * P.colscale[d] * P.colscale[P.D + mm] *
                                    Gl[(size_t)P.colmap[d] * RJ + P.colmap[P.D + mm]]);
            }
        }
        // the voltages and Jacobian of the last step, when asked for
        if (step == steps - 1) {
            if (Uout) std::copy(U.begin(), U.end(), Uout);
            if (Jout) std::copy(J.begin(), J.end(), Jout);
        }
        // data-space Tikhonov step: (J J^T + lam2 I) y = Um - U, delta = J^T y
        for (int i = 0; i < m; ++i) r[i] = Um[i] - U[i];
        gram(m, E, J.data(), G.data());
        for (size_t t = 0; t < (size_t)m * m; ++t) L[t] = G[t];
        for (int i = 0; i < m; ++i) L[(size_t)i * m + i] += P.lam2;
        if (!cholesky(m, L)) return;
        chol_solve(m, L, r.data(), y.data());
        auto jt_y = [&](const std::vector<double>& yv, std::vector<double>& out) {
            for (int e = 0; e < E; ++e) {
                const float* Je = &J[(size_t)e * m];
                double s = 0.0;
                for (int i = 0; i < m; ++i) s += (double)Je[i] * yv[i];
                out[e] = s;
            }
        };
        for (int pass = 0; pass < P.gn_refine; ++pass) {
            jt_y(y, vE);
            for (int i = 0; i < m; ++i) res_m[i] = r[i] - P.lam2 * y[i];
            for (int e = 0; e < E; ++e) {
                const float* Je = &J[(size_t)e * m];
                for (int i = 0; i < m; ++i) res_m[i] -= (double)Je[i] * vE[e];
            }
            chol_solve(m, L, res_m.data(), dy.data());
            for (int i = 0; i < m; ++i) y[i] += dy[i];
        }
        jt_y(y, delta);
        for (int e = 0; e < E; ++e) gamma[e] = (float)((double)gamma[e] + delta[e]);
    }
}

}  // namespace

// `frames` reconstructions of `steps` Gauss-Newton steps each, one frame per thread;
// gamma [frames][E] in/out, Um [frames][Dp*Mp].  Uout [frames][m] and Jout [frames][E][m]
// (nullable) receive the voltages and the Jacobian of each frame's last step.  Returns
// the wall time in seconds.
double gnh_reconstruct(const GnhProblem* P, int frames, const double* Um, int steps, float* gamma, int threads,
                       double* Uout, float* Jout) {
    if (threads > 0) omp_set_num_threads(threads);
    const int m = P->Dp * P->Mp;
    const auto t0 = std::chrono::steady_clock::now();
#pragma omp parallel for schedule(dynamic, 1)
    for (int f = 0; f < frames; ++f)
        frame(*P, Um + (size_t)f * m, steps, gamma + (size_t)f * P->E, Uout ? Uout + (size_t)f * m : nullptr,
              Jout ? Jout + (size_t)f * P->E * m : nullptr);
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // extern "C"
