// Host-side finite-element set-up: element integrals, the ELL pattern, the gather
// plan for sigma-weighted assembly, the gap-model excitation and the measurement map.
//
// Follows FEM::Equation::Equation / initElementalMatrices / initExcitationMatrix /
// initJacobianCalculationMatrix (src/fem/equation.cpp:29-253) and models::EIT::EIT
// (src/models/eit.cpp:29-95), re-laid out for the device kernels.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <thread>

#include "mpf_internal.h"

namespace mpf {
namespace {

// Polynomial in barycentric coordinates: c[p][q][r] * l0^p l1^q l2^r, degree <= 4.
struct Poly {
    double c[5][5][5];
    Poly() { std::memset(c, 0, sizeof(c)); }
};

Poly mul(const Poly& a, const Poly& b) {
    Poly out;
    for (int p = 0; p < 5; ++p) for (int q = 0; q + p < 5; ++q) for (int r = 0; r + p + q < 5; ++r) {
        double v = a.c[p][q][r];
        if (v == 0.0) continue;
        for (int s = 0; p + s < 5; ++s) for (int t = 0; q + t < 5; ++t) for (int u = 0; r + u < 5; ++u) {
            if (p + s + q + t + r + u > 4) continue;
            out.c[p + s][q + t][r + u] += v * b.c[s][t][u];
        }
    }
    return out;
}

Poly deriv(const Poly& a, int var) {
    Poly out;
    for (int p = 0; p < 5; ++p) for (int q = 0; q < 5; ++q) for (int r = 0; r < 5; ++r) {
        double v = a.c[p][q][r];
        if (v == 0.0) continue;
        if (var == 0 && p > 0) out.c[p - 1][q][r] += v * p;
        if (var == 1 && q > 0) out.c[p][q - 1][r] += v * q;
        if (var == 2 && r > 0) out.c[p][q][r - 1] += v * r;
    }
    return out;
}

double fact(int n) { double f = 1; for (int i = 2; i <= n; ++i) f *= i; return f; }

// int_T l0^p l1^q l2^r dA = 2 A p! q! r! / (p+q+r+2)!
double integrate(const Poly& a, double area) {
    double s = 0.0;
    for (int p = 0; p < 5; ++p) for (int q = 0; q < 5; ++q) for (int r = 0; r < 5; ++r) {
        double v = a.c[p][q][r];
        if (v == 0.0) continue;
        s += v * fact(p) * fact(q) * fact(r) / fact(p + q + r + 2);
    }
    return 2.0 * area * s;
}

void nodal_basis(int k, Poly* phi) {
    if (k == 3) {
        phi[0].c[1][0][0] = 1; phi[1].c[0][1][0] = 1; phi[2].c[0][0][1] = 1;
        return;
    }
    // vertices: l (2 l - 1); midpoints: 4 l_a l_b   (order v0 v1 v2 m01 m12 m20)
    phi[0].c[2][0][0] = 2; phi[0].c[1][0][0] = -1;
    phi[1].c[0][2][0] = 2; phi[1].c[0][1][0] = -1;
    phi[2].c[0][0][2] = 2; phi[2].c[0][0][1] = -1;
    phi[3].c[1][1][0] = 4;
    phi[4].c[0][1][1] = 4;
    phi[5].c[1][0][1] = 4;
}

}  // namespace

void element_integrals(int k, const double* xy, double* K, double* Mm) {
    const double x0 = xy[0], y0 = xy[1], x1 = xy[2], y1 = xy[3], x2 = xy[4], y2 = xy[5];
    const double det = (x1 - x0) * (y2 - y0) - (x2 - x0) * (y1 - y0);
    const double area = 0.5 * std::fabs(det);
    // grad l_a = (y_b - y_c, x_c - x_b) / det for (a,b,c) cyclic
    const double gl[3][2] = {{(y1 - y2) / det, (x2 - x1) / det},
                             {(y2 - y0) / det, (x0 - x2) / det},
                             {(y0 - y1) / det, (x1 - x0) / det}};
    // the barycentric integrals are element independent up to the area factor: tabulate
    // them once per element type (unit area)
    struct Tables {
        double m[6][6];
        double kk[6][6][3][3];
    };
    static const auto tabulate = [](int kw) {
        Tables t;
        Poly phi[6];
        nodal_basis(kw, phi);
        Poly dphi[6][3];
        for (int i = 0; i < kw; ++i) for (int a = 0; a < 3; ++a) dphi[i][a] = deriv(phi[i], a);
        for (int i = 0; i < kw; ++i) for (int j = 0; j < kw; ++j) {
            t.m[i][j] = integrate(mul(phi[i], phi[j]), 1.0);
            for (int a = 0; a < 3; ++a) for (int b = 0; b < 3; ++b)
                t.kk[i][j][a][b] = integrate(mul(dphi[i][a], dphi[j][b]), 1.0);
        }
        return t;
    };
    static const Tables tab[2] = {tabulate(3), tabulate(6)};
    const Tables& T = tab[k == 3 ? 0 : 1];
    for (int i = 0; i < k; ++i) for (int j = 0; j < k; ++j) {
        Mm[i * k + j] = area * T.m[i][j];
        double s = 0.0;
        for (int a = 0; a < 3; ++a) for (int b = 0; b < 3; ++b) {
            double g = gl[a][0] * gl[b][0] + gl[a][1] * gl[b][1];
            s += g * (area * T.kk[i][j][a][b]);
        }
        K[i * k + j] = s;
    }
}

// ---------------------------------------------------------------------------
// Ports::circularBoundary (src/fem/ports.cpp:32-77) with math::polar /
// circleParameter / circularPoints (src/mathematics.cpp:26-99)
// ---------------------------------------------------------------------------
static double polar_angle(double x, double y) {
    if (x > 0.0) return std::atan(y / x);
    if (x < 0.0 && y >= 0.0) return std::atan(y / x) + M_PI;
    if (x < 0.0 && y < 0.0) return std::atan(y / x) - M_PI;
    if (x == 0.0 && y > 0.0) return M_PI / 2;
    if (x == 0.0 && y < 0.0) return -M_PI / 2;
    return 0.0;
}

static double circle_parameter(double x, double y, double offset) {
    double r = std::sqrt(x * x + y * y);
    double a = polar_angle(x, y) - offset / r;
    a += (a < -M_PI) ? 2.0 * M_PI : 0.0;
    a -= (a > M_PI) ? 2.0 * M_PI : 0.0;
    return a * r;
}

std::vector<std::vector<int32_t>> circular_ports(int N, const double* nodes, int E, int k,
    const int32_t* elements, int nb, int kb, const int32_t* boundary, int count,
    double width, double offset, bool clockwise) {
    MPF_REQUIRE(count > 0, "ports: count <= 0");
    double radius = 0.0;
    for (int e = 0; e < E; ++e)
        for (int i = 0; i < 3; ++i) {
            int n = elements[e * k + i];
            MPF_REQUIRE(n >= 0 && n < N, "ports: element node out of range");
            radius = std::max(radius, std::sqrt(nodes[2 * n] * nodes[2 * n] + nodes[2 * n + 1] * nodes[2 * n + 1]));
        }
    auto corner = [&](double off, int l, double* px, double* py) {
        double phi = off / radius + l * (2.0 * M_PI * radius / count) / radius;
        *px = radius * std::cos(phi);
        *py = (clockwise ? -1.0 : 1.0) * radius * std::sin(phi);
    };
    std::vector<std::vector<int32_t>> out(count);
    for (int port = 0; port < count; ++port) {
        double sx, sy, ex, ey;
        corner(offset + (clockwise ? width : 0.0), port, &sx, &sy);
        corner(offset + (clockwise ? 0.0 : width), port, &ex, &ey);
        for (int b = 0; b < nb; ++b) {
            const int32_t* ed = boundary + (size_t)b * kb;
            double x0 = nodes[2 * ed[0]], y0 = nodes[2 * ed[0] + 1];
            double xl = nodes[2 * ed[kb - 1]], yl = nodes[2 * ed[kb - 1] + 1];
            double par_last = std::sqrt((xl - x0) * (xl - x0) + (yl - y0) * (yl - y0));
            double off0 = circle_parameter(x0, y0, 0.0);
            double a = circle_parameter(sx, sy, off0);
            double c = circle_parameter(ex, ey, off0);
            if (a < c && 0.0 - a >= -1e-9 && par_last - c <= 1e-9) out[port].push_back(b);
        }
    }
    return out;
}

// ---------------------------------------------------------------------------
// full set-up
// ---------------------------------------------------------------------------
// Lagrange basis of one straight boundary edge with nodes at distances sp[0] = 0 <
// ... < sp[kb-1] from its first node: m[a][b] = int phi_a phi_b ds and v[a] = int phi_a ds,
// by 3-point Gauss-Legendre (exact up to degree 5).
static void edge_integrals(int kb, const double* sp, double (*m)[3], double* v) {
    static const double xg[3] = {-0.7745966692414834, 0.0, 0.7745966692414834};
    static const double wg[3] = {5.0 / 9.0, 8.0 / 9.0, 5.0 / 9.0};
    const double L = sp[kb - 1];
    for (int a = 0; a < kb; ++a) {
        v[a] = 0.0;
        for (int b = 0; b < kb; ++b) m[a][b] = 0.0;
    }
    for (int g = 0; g < 3; ++g) {
        const double t = 0.5 * L * (xg[g] + 1.0), w = 0.5 * L * wg[g];
        double phi[3];
        for (int a = 0; a < kb; ++a) {
            phi[a] = 1.0;
            for (int c = 0; c < kb; ++c)
                if (c != a) phi[a] *= (t - sp[c]) / (sp[a] - sp[c]);
        }
        for (int a = 0; a < kb; ++a) {
            v[a] += w * phi[a];
            for (int b = 0; b < kb; ++b) m[a][b] += w * phi[a] * phi[b];
        }
    }
}

// Device node numbering.  Recursive coordinate bisection: a node set is split at the
// median of its longer coordinate extent, the first part holding a whole number of
// 256-node leaves, until every leaf fits one 256-row tile.  Each tile is then a compact
// patch of the mesh: most gathers of its SpMV rows stay inside it, consecutive tiles are
// neighbouring patches (the XCD-striped work lists keep them in one L2).  Each tile is
// then bisected the same way into leaves of kAggRows nodes (a part holding a whole number
// of them first): the kAggPerTile aggregates of the two-level preconditioner, quarter[node]
// = the leaf's index in its tile.  The tile's rows are then sorted by descending degree
// (the aggregates are sets, not row ranges) so the 16-row groups of the SpMV have
// near-uniform ELL widths.  Returns perm[node] = row.
static std::vector<int32_t> device_order(int N, const double* xy, int E, int k, const int32_t* el,
                                         std::vector<uint8_t>& quarter) {
    std::vector<int32_t> deg(N, 0);
    {
        std::vector<std::vector<int32_t>> nb(N);
        for (int e = 0; e < E; ++e)
            for (int i = 0; i < k; ++i)
                for (int j = 0; j < k; ++j)
                    if (i != j) nb[el[(size_t)e * k + i]].push_back(el[(size_t)e * k + j]);
        for (int n = 0; n < N; ++n) {
            std::sort(nb[n].begin(), nb[n].end());
            deg[n] = (int32_t)(std::unique(nb[n].begin(), nb[n].end()) - nb[n].begin());
        }
    }
    std::vector<int32_t> order(N);
    for (int n = 0; n < N; ++n) order[n] = n;
    // order[b, e) split at `left` along the longer coordinate extent of the set; ties
    // broken by node id, so the split does not depend on the library's algorithm
    auto bisect = [&](int b, int e, int left) {
        double lo[2] = {1e300, 1e300}, hi[2] = {-1e300, -1e300};
        for (int i = b; i < e; ++i)
            for (int a = 0; a < 2; ++a) {
                lo[a] = std::min(lo[a], xy[2 * (size_t)order[i] + a]);
                hi[a] = std::max(hi[a], xy[2 * (size_t)order[i] + a]);
            }
        const int ax = (hi[0] - lo[0]) >= (hi[1] - lo[1]) ? 0 : 1;
        std::nth_element(order.begin() + b, order.begin() + b + left, order.begin() + e, [&](int a, int c) {
            const double va = xy[2 * (size_t)a + ax], vc = xy[2 * (size_t)c + ax];
            return va < vc || (va == vc && a < c);
        });
    };
    // [b, e) split until every part fits one leaf of `leaf` rows
    auto split = [&](int b0, int e0, int leaf) {
        std::vector<std::pair<int, int>> todo{{b0, e0}};
        while (!todo.empty()) {
            const auto [b, e] = todo.back();
            todo.pop_back();
            const int n = e - b;
            if (n <= leaf) continue;
            const int leaves = (n + leaf - 1) / leaf;
            const int left = (leaves / 2) * leaf;
            bisect(b, e, left);
            todo.emplace_back(b + left, e);
            todo.emplace_back(b, b + left);
        }
    };
    split(0, N, kTileRows);
    auto by_degree = [&](int a, int c) { return deg[a] != deg[c] ? deg[a] > deg[c] : a < c; };
    quarter.assign(N, 0);
    for (int w = 0; w < N; w += kTileRows) {
        const int e = std::min(N, w + kTileRows);
        split(w, e, kAggRows);
        for (int i = w; i < e; ++i) quarter[order[i]] = (uint8_t)((i - w) / kAggRows);
        std::sort(order.begin() + w, order.begin() + e, by_degree);
    }
    std::vector<int32_t> perm(N);
    for (int r = 0; r < N; ++r) perm[order[r]] = r;
    return perm;
}

// In-place inverse of an SPD matrix (n x n, row-major, FP64) by Cholesky: A = L L^T,
// A^-1 = L^-T L^-1.  Returns false if A is not numerically positive definite.
static bool spd_inverse(int n, double* A) {
    // Cholesky A = L L^T, dot-product form over contiguous row prefixes
    std::vector<double> L((size_t)n * n, 0.0);
    for (int j = 0; j < n; ++j) {
        const double* lj = &L[(size_t)j * n];
        double d = A[(size_t)j * n + j];
        for (int q = 0; q < j; ++q) d -= lj[q] * lj[q];
        if (!(d > 0.0)) return false;
        const double ljj = std::sqrt(d);
        L[(size_t)j * n + j] = ljj;
        for (int i = j + 1; i < n; ++i) {
            const double* li = &L[(size_t)i * n];
            double v = A[(size_t)i * n + j];
            for (int q = 0; q < j; ++q) v -= li[q] * lj[q];
            L[(size_t)i * n + j] = v / ljj;
        }
    }
    // M = L^-1, one row at a time: M_i = (e_i - sum_{q < i} L_iq M_q) / L_ii, the sum as
    // row updates over M_q's nonzero prefix [0, q]
    std::vector<double> M((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) {
        double* mi = &M[(size_t)i * n];
        const double* li = &L[(size_t)i * n];
        for (int q = 0; q < i; ++q) {
            const double c = li[q];
            const double* mq = &M[(size_t)q * n];
            for (int u = 0; u <= q; ++u) mi[u] -= c * mq[u];
        }
        mi[i] += 1.0;
        const double inv = 1.0 / li[i];
        for (int u = 0; u <= i; ++u) mi[u] *= inv;
    }
    // A^-1 = M^T M, accumulated as rank-one row updates: row i (j <= i) += M_qi M_q[0, i]
    std::fill(A, A + (size_t)n * n, 0.0);
    for (int q = 0; q < n; ++q) {
        const double* mq = &M[(size_t)q * n];
        for (int i = 0; i <= q; ++i) {
            const double c = mq[i];
            double* ai = &A[(size_t)i * n];
            for (int j = 0; j <= i; ++j) ai[j] += c * mq[j];
        }
    }
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < i; ++j) A[(size_t)j * n + i] = A[(size_t)i * n + j];
    return true;
}

// Coarse level of the two-level preconditioner (aggregates = the kAggRows-node leaves of
// each 256-row tile, aggregate kAggPerTile t + q holding the rows with agg_q = q; W = the
// aggregate indicator): per harmonic, the inverse
// of W^T A_c(1) W at sigma = 1 (plus
// n n^T, n = real rows per aggregate, on the dcFree harmonic 0, matching the A + 1 1^T
// operator), and per aggregate the sum of the sigma = 1 stiffness diagonal that scales
// it to a frame's conductivity.  Skipped (coarse_nt = 0) above kCoarseMaxTiles tiles.
static void coarse_setup(MeshSetup& s) {
    const int np = s.np, nt = np / kTileRows, na = kAggPerTile * nt;
    s.coarse_nt = 0;
    if (nt > kCoarseMaxTiles) return;
    const size_t nent = (size_t)s.wpad * np;
    std::vector<double> sk(nent, 0.0), sm(nent, 0.0);
    for (size_t ent = 0; ent < nent; ++ent)
        for (int i = s.ctb_off[ent]; i < s.ctb_off[ent + 1]; ++i) { sk[ent] += s.ctb_k[i]; sm[ent] += s.ctb_m[i]; }
    // aggregate of each real row
    std::vector<int> agg(s.N);
    for (int row = 0; row < s.N; ++row) agg[row] = kAggPerTile * (row / kTileRows) + s.agg_q[row];
    std::vector<double> cnt(na, 0.0);
    for (int row = 0; row < s.N; ++row) cnt[agg[row]] += 1.0;
    s.coarse_d1.assign(na, 0.0);
    for (int row = 0; row < s.N; ++row) s.coarse_d1[agg[row]] += sk[(size_t)s.diag_slot[row] * np + row];
    const int ld = (na + 3) / 4 * 4;
    s.coarse_inv.assign((size_t)s.C * na * ld, 0.0f);
    // one host thread per harmonic (each forms and inverts its own coarse matrix)
    std::vector<char> ok(s.C, 0);
    auto harmonic = [&](int c) {
        std::vector<double> Ec((size_t)na * na, 0.0);
        for (int q = 0; q < s.wpad; ++q)
            for (int row = 0; row < s.N; ++row) {
                const size_t ent = (size_t)q * np + row;
                if (s.ctb_off[ent] == s.ctb_off[ent + 1] || s.ids[ent] >= s.N) continue;   // padding slot
                Ec[(size_t)agg[row] * na + agg[s.ids[ent]]] += sk[ent] + s.alpha[c] * sm[ent];
            }
        if (c == 0 && !s.fixed && !s.cem_coupled)
            for (int t = 0; t < na; ++t)
                for (int u = 0; u < na; ++u) Ec[(size_t)t * na + u] += cnt[t] * cnt[u];
        // aggregates without real rows (past the last real row of the last tile) keep an
        // identity block
        for (int t = 0; t < na; ++t)
            if (cnt[t] == 0.0) Ec[(size_t)t * na + t] = 1.0;
        if (!spd_inverse(na, Ec.data())) return;
        for (int t = 0; t < na; ++t)
            for (int u = 0; u < na; ++u)
                s.coarse_inv[((size_t)c * na + t) * ld + u] = (float)Ec[(size_t)t * na + u];
        ok[c] = 1;
    };
    std::vector<std::thread> pool;
    for (int c = 1; c < s.C; ++c) pool.emplace_back(harmonic, c);
    harmonic(0);
    for (auto& th : pool) th.join();
    for (int c = 0; c < s.C; ++c)
        if (!ok[c]) return;
    s.coarse_nt = nt;
    s.coarse_ld = ld;
}

// Laplacian regularisation R = L^T L (Inverse::calcRegularizationMatrix,
// src/solver/inverse.cpp:96-117): L[edge][element] = +|edge| for the first element that
// uses an edge and -|edge| for the second, over the elements' vertex edges.  Hence
// R_ee = sum of |edge|^2 over e's edges and R_ab = -|edge|^2 for elements a, b sharing
// an edge; kept as the diagonal and up to three (neighbour, value) pairs per element.
static void laplacian_setup(const mpf_model_desc& d, MeshSetup& s) {
    const int E = d.n_elements, k = d.nodes_per_element;
    s.lap_diag.assign(E, 0.0);
    s.lap_nb.assign((size_t)E * 3, -1);
    s.lap_val.assign((size_t)E * 3, 0.0);
    std::map<std::pair<int, int>, std::pair<int, int>> first;   // edge -> (element, local edge)
    for (int e = 0; e < E; ++e) {
        const int32_t* el = d.elements + (size_t)e * k;
        for (int i = 0; i < 3; ++i) {
            const int a = el[i], b = el[(i + 1) % 3];
            const double dx = d.nodes[2 * a] - d.nodes[2 * b], dy = d.nodes[2 * a + 1] - d.nodes[2 * b + 1];
            const double l2 = dx * dx + dy * dy;
            s.lap_diag[e] += l2;
            const auto key = std::make_pair(std::min(a, b), std::max(a, b));
            const auto it = first.find(key);
            if (it == first.end()) {
                first.emplace(key, std::make_pair(e, i));
                continue;
            }
            const int f = it->second.first, j = it->second.second;
            MPF_REQUIRE(s.lap_nb[(size_t)f * 3 + j] < 0, "laplacian: an edge is shared by more than two elements");
            s.lap_nb[(size_t)f * 3 + j] = e;
            s.lap_val[(size_t)f * 3 + j] = -l2;
            s.lap_nb[(size_t)e * 3 + i] = f;
            s.lap_val[(size_t)e * 3 + i] = -l2;
        }
    }
}

// Fixed sources: rows D (a nonzero clipped excitation) become identity rows and the
// columns of D leave the other rows, so the device solves the symmetric system
// [I 0; 0 A_II] x = [b_D; -A_ID b_D] — the solution of the reference's row-replaced
// system A_mod x = b.  The dropped contributions are kept per row: couplings of rows
// outside D to columns in D (for the right-hand side) and the full rows of D (for the
// electrode currents A_c x on D).
static void fixed_setup(MeshSetup& s, const std::vector<std::map<int, double>>& Bn, int Nu) {
    const int np = s.np, wp = s.wpad;
    std::vector<char> dir(np, 0);
    for (int n = 0; n < Nu; ++n)
        for (auto& kv : Bn[n]) if (kv.second != 0.0) dir[n] = 1;
    struct Ctb { int32_t node, elem; double k, m; };
    std::vector<std::vector<Ctb>> cpl(np), drow(np);
    std::vector<int32_t> off(1, 0), elem;
    std::vector<double> ck, cm;
    for (int q = 0; q < wp; ++q)
        for (int row = 0; row < np; ++row) {
            const size_t ent = (size_t)q * np + row;
            const int col = s.ids[ent];
            for (int i = s.ctb_off[ent]; i < s.ctb_off[ent + 1]; ++i) {
                const Ctb c{col, s.ctb_elem[i], s.ctb_k[i], s.ctb_m[i]};
                if (dir[row]) drow[row].push_back(c);
                else if (dir[col]) cpl[row].push_back(c);
                else { elem.push_back(c.elem); ck.push_back(c.k); cm.push_back(c.m); }
            }
            if (dir[row] && q == s.diag_slot[row]) { elem.push_back(-1); ck.push_back(1.0); cm.push_back(0.0); }
            off.push_back((int32_t)elem.size());
        }
    s.ctb_off.swap(off); s.ctb_elem.swap(elem); s.ctb_k.swap(ck); s.ctb_m.swap(cm);
    auto flatten = [&](const std::vector<std::vector<Ctb>>& v, std::vector<int32_t>& o, std::vector<int32_t>& nd,
                       std::vector<int32_t>& el, std::vector<double>& kk, std::vector<double>& mm) {
        o.assign(1, 0);
        for (int row = 0; row < np; ++row) {
            for (const Ctb& c : v[row]) { nd.push_back(c.node); el.push_back(c.elem); kk.push_back(c.k); mm.push_back(c.m); }
            o.push_back((int32_t)nd.size());
        }
    };
    flatten(cpl, s.cpl_off, s.cpl_node, s.cpl_elem, s.cpl_k, s.cpl_m);
    flatten(drow, s.drow_off, s.drow_node, s.drow_elem, s.drow_k, s.drow_m);
}

void build_setup(const mpf_model_desc& caller, MeshSetup& s) {
    const mpf_model_desc& c0 = caller;
    MPF_REQUIRE(c0.nodes && c0.elements && c0.boundary, "model: null mesh array");
    MPF_REQUIRE(c0.n_nodes > 0 && c0.n_elements > 0, "model: empty mesh");
    MPF_REQUIRE(c0.nodes_per_element == 3 || c0.nodes_per_element == 6, "model: nodes_per_element must be 3 or 6");
    MPF_REQUIRE(c0.nodes_per_edge == (c0.nodes_per_element == 3 ? 2 : 3), "model: nodes_per_edge does not match the element type");
    for (size_t i = 0; i < (size_t)c0.n_elements * c0.nodes_per_element; ++i)
        MPF_REQUIRE(c0.elements[i] >= 0 && c0.elements[i] < c0.n_nodes, "model: element node index out of range");
    for (size_t i = 0; i < (size_t)c0.n_boundary * c0.nodes_per_edge; ++i)
        MPF_REQUIRE(c0.boundary[i] >= 0 && c0.boundary[i] < c0.n_nodes, "model: boundary node index out of range");
    // renumber the mesh into the device order; everything below works in that numbering
    mpf_model_desc d = caller;
    std::vector<double> nodes_r;
    std::vector<int32_t> el_r, bd_r;
    // aggregate of each caller node within its tile (kept order: consecutive kAggRows runs)
    std::vector<uint8_t> quarter(caller.n_nodes);
    if (caller.keep_node_order) {
        s.perm.resize(caller.n_nodes);
        for (int n = 0; n < caller.n_nodes; ++n) {
            s.perm[n] = n;
            quarter[n] = (uint8_t)((n % kTileRows) / kAggRows);
        }
    } else {
        s.perm = device_order(caller.n_nodes, caller.nodes, caller.n_elements, caller.nodes_per_element,
                              caller.elements, quarter);
        nodes_r.resize((size_t)2 * caller.n_nodes);
        for (int n = 0; n < caller.n_nodes; ++n) {
            nodes_r[2 * (size_t)s.perm[n]] = caller.nodes[2 * (size_t)n];
            nodes_r[2 * (size_t)s.perm[n] + 1] = caller.nodes[2 * (size_t)n + 1];
        }
        el_r.resize((size_t)caller.n_elements * caller.nodes_per_element);
        for (size_t i = 0; i < el_r.size(); ++i) el_r[i] = s.perm[caller.elements[i]];
        bd_r.resize((size_t)caller.n_boundary * caller.nodes_per_edge);
        for (size_t i = 0; i < bd_r.size(); ++i) bd_r[i] = s.perm[caller.boundary[i]];
        d.nodes = nodes_r.data();
        d.elements = el_r.data();
        d.boundary = bd_r.data();
    }
    MPF_REQUIRE(d.n_ports > 0 && d.port_edges, "model: no ports");
    MPF_REQUIRE(d.n_drive > 0 && d.n_meas > 0 && d.drive_pattern && d.meas_pattern && d.drive_current, "model: no sources");
    MPF_REQUIRE(d.components >= 1 && d.components <= kMaxSlices, "model: components must be in [1, 8]");
    MPF_REQUIRE(d.height > 0.0, "model: height <= 0.0");
    MPF_REQUIRE(d.port_height > 0.0, "model: portHeight <= 0.0");
    const int N = d.n_nodes, E = d.n_elements, k = d.nodes_per_element;
    MPF_REQUIRE(d.contact_impedance >= 0.0, "model: contact_impedance < 0");
    const bool cem = d.contact_impedance > 0.0;
    // the coupled electrode model eliminates the electrode potentials (Schur complement on
    // V after the solves): its unknowns are the mesh nodes alone
    const bool coupled = cem && d.cem_coupled != 0;
    MPF_REQUIRE(d.cem_coupled == 0 || cem, "model: cem_coupled needs contact_impedance > 0");
    MPF_REQUIRE(!coupled || d.jacobian_exact, "model: the coupled electrode model needs jacobian_exact = 1");
    s.cem_coupled = coupled ? 1 : 0;
    // unknowns: the mesh nodes, then (harmonic-wise complete electrode model) one potential
    // per port
    const int Nu = N + (cem && !coupled ? d.n_ports : 0);
    s.N = Nu; s.E = E; s.k = k;
    s.np = round_up(Nu, kTileRows);
    // aggregate within the tile of every row: the mesh nodes' from the ordering, the
    // electrode rows (harmonic-wise electrode model) and padding rows by position
    s.agg_q.resize(s.np);
    for (int r = 0; r < s.np; ++r) s.agg_q[r] = (uint8_t)((r % kTileRows) / kAggRows);
    for (int n = 0; n < N; ++n) s.agg_q[s.perm[n]] = quarter[n];
    s.P = d.n_ports; s.D = d.n_drive; s.M = d.n_meas; s.R = s.D + s.M;
    MPF_REQUIRE(s.R <= 64, "model: D + M > 64 is not supported");
    // Pi = [I*Dpat | Mpat] (sources.cpp:57-71); columns that are scalar multiples of an
    // earlier column (adjacent measurement = adjacent drive / I) share one solve
    {
        auto Pi = [&](int p, int j) {
            return j < s.D ? d.drive_current[j] * d.drive_pattern[(size_t)p * s.D + j]
                           : d.meas_pattern[(size_t)p * s.M + (j - s.D)];
        };
        std::vector<std::vector<double>> uniq;   // normalised representatives
        std::vector<double> uamp;
        s.colmap.assign(s.R, 0);
        s.colscale.assign(s.R, 0.0);
        for (int j = 0; j < s.R; ++j) {
            double amp = 0.0;
            for (int p = 0; p < s.P; ++p)
                if (std::fabs(Pi(p, j)) > std::fabs(amp)) amp = Pi(p, j);
            std::vector<double> nrm(s.P, 0.0);
            if (amp != 0.0)
                for (int p = 0; p < s.P; ++p) nrm[p] = Pi(p, j) / amp;
            int u = -1;
            for (size_t k = 0; k < uniq.size() && u < 0; ++k) {
                bool same = true;
                for (int p = 0; p < s.P && same; ++p) same = std::fabs(uniq[k][p] - nrm[p]) <= 1e-14;
                if (same) u = (int)k;
            }
            if (u < 0) {
                u = (int)uniq.size();
                uniq.push_back(nrm);
                uamp.push_back(amp);
            }
            s.colmap[j] = u;
            s.colscale[j] = uamp[u] != 0.0 ? amp / uamp[u] : 0.0;
        }
        s.Ru = (int)uniq.size();
    }
    s.Rp = s.Ru <= 16 ? 16 : (s.Ru <= 32 ? 32 : 64);
    if (coupled) {
        // the pattern columns are combinations of the L unit electrode fields: the solves
        // run one lane per port; the pattern's unique columns index the combined fields
        MPF_REQUIRE(d.n_ports <= 64, "model: the coupled electrode model supports at most 64 ports");
        s.pRu = s.Ru;
        s.pRp = s.Rp;
        s.Ru = d.n_ports;
        s.Rp = s.Ru <= 16 ? 16 : (s.Ru <= 32 ? 32 : 64);
    }
    s.Mp = ceil16(s.M); s.Dp = ceil16(s.D); s.mrows = s.Mp * s.Dp;
    s.C = d.components;
    s.ref = d.reference_value; s.height = d.height; s.port_height = d.port_height;
    s.logarithmic = d.logarithmic; s.jacobian_exact = d.jacobian_exact;
    for (int c = 0; c < s.C; ++c) {
        s.alpha[c] = std::pow(2.0 * c * M_PI / d.height, 2);
        s.beta[c] = c == 0 ? 1.0 / d.height
                           : 2.0 * std::sin(c * M_PI * d.port_height / d.height) / (c * M_PI * d.port_height);
    }
    const int cap = d.ell_max_width > 0 ? d.ell_max_width : 32;
    MPF_REQUIRE(cap <= 64, "model: ell_max_width > 64");

    s.elements.assign(d.elements, d.elements + (size_t)E * k);

    // element integrals
    std::vector<double> Ke((size_t)E * k * k), Me((size_t)E * k * k);
    for (int e = 0; e < E; ++e) {
        double xy[12];
        const int32_t* el = d.elements + (size_t)e * k;
        for (int i = 0; i < 3; ++i) { xy[2 * i] = d.nodes[2 * el[i]]; xy[2 * i + 1] = d.nodes[2 * el[i] + 1]; }
        if (k == 6) {
            static const int pr[3][2] = {{0, 1}, {1, 2}, {2, 0}};
            double hx = std::fabs(xy[2] - xy[0]) + std::fabs(xy[3] - xy[1]) + 1e-300;
            for (int t = 0; t < 3; ++t) {
                int a = el[pr[t][0]], b = el[pr[t][1]], m = el[3 + t];
                double mx = 0.5 * (d.nodes[2 * a] + d.nodes[2 * b]), my = 0.5 * (d.nodes[2 * a + 1] + d.nodes[2 * b + 1]);
                MPF_REQUIRE(std::fabs(mx - d.nodes[2 * m]) + std::fabs(my - d.nodes[2 * m + 1]) < 1e-9 * (1.0 + hx) * 1e3,
                            "model: P2 element nodes must be ordered [v0 v1 v2 m01 m12 m20] with straight midpoints");
            }
        }
        element_integrals(k, xy, &Ke[(size_t)e * k * k], &Me[(size_t)e * k * k]);
    }
    s.elem_k.resize(Ke.size());
    s.elem_m.resize(Me.size());
    for (size_t i = 0; i < Ke.size(); ++i) { s.elem_k[i] = (float)Ke[i]; s.elem_m[i] = (float)Me[i]; }

    // ELL pattern: first-insertion order of the element loop (equation.cpp:131-144)
    std::vector<int32_t> cols((size_t)Nu * cap, -1);
    std::vector<int32_t> cnt(Nu, 0);
    std::vector<int32_t> slot_of((size_t)E * k * k);
    int width = 0;
    auto slot = [&](int row, int col) {
        int32_t* rc = &cols[(size_t)row * cap];
        for (int q = 0; q < cnt[row]; ++q) if (rc[q] == col) return q;
        if (cnt[row] >= cap)
            throw Error(MPF_EFULL, "mpFlow::numeric::SparseMatrix::setValue: sparse format full");
        const int sl = cnt[row]++;
        rc[sl] = col;
        width = std::max(width, cnt[row]);
        return sl;
    };
    for (int e = 0; e < E; ++e) {
        const int32_t* el = d.elements + (size_t)e * k;
        for (int i = 0; i < k; ++i) for (int j = 0; j < k; ++j)
            slot_of[((size_t)e * k + i) * k + j] = slot(el[i], el[j]);
    }
    // complete electrode model: sigma-independent blocks (1/z) int phi_i phi_j, -(1/z)
    // int phi_i and |e|/z on the port edges, kept as (row, slot, value) contributions
    struct Const { int row, sl; double v; };
    std::vector<Const> consts;
    if (coupled) {
        // Zb / z' on the port edges; the sources Bc and the electrode block live in the
        // Schur complement (bc_*, elen)
        s.zp = d.contact_impedance * d.height / d.port_height;
        const double iz = 1.0 / s.zp;
        const int kb = d.nodes_per_edge;
        s.elen.assign(s.P, 0.0);
        std::vector<std::map<int, double>> bc(s.P);
        for (int p = 0; p < s.P; ++p)
            for (int q = 0; q < d.port_max_edges; ++q) {
                const int eidx = d.port_edges[(size_t)p * d.port_max_edges + q];
                if (eidx < 0) continue;
                MPF_REQUIRE(eidx < d.n_boundary, "model: port edge index out of range");
                const int32_t* ed = d.boundary + (size_t)eidx * kb;
                double sp[3], m[3][3], v[3];
                for (int a = 0; a < kb; ++a) {
                    const double dx = d.nodes[2 * ed[a]] - d.nodes[2 * ed[0]];
                    const double dy = d.nodes[2 * ed[a] + 1] - d.nodes[2 * ed[0] + 1];
                    sp[a] = std::sqrt(dx * dx + dy * dy);
                }
                edge_integrals(kb, sp, m, v);
                for (int a = 0; a < kb; ++a) {
                    for (int b = 0; b < kb; ++b) consts.push_back({ed[a], slot(ed[a], ed[b]), iz * m[a][b]});
                    bc[p][ed[a]] += v[a];
                }
                s.elen[p] += sp[kb - 1];
            }
        s.bc_off.assign(s.P + 1, 0);
        s.bc_row.clear();
        s.bc_val.clear();
        for (int p = 0; p < s.P; ++p) {
            for (auto& kv : bc[p]) { s.bc_row.push_back(kv.first); s.bc_val.push_back(kv.second); }
            s.bc_off[p + 1] = (int32_t)s.bc_row.size();
        }
        for (int c = 0; c < s.C; ++c) s.n_c[c] = c == 0 ? d.height : 0.5 * d.height;
    } else if (cem) {
        const double iz = 1.0 / d.contact_impedance;
        const int kb = d.nodes_per_edge;
        for (int p = 0; p < s.P; ++p) {
            const int ep = N + p;
            for (int q = 0; q < d.port_max_edges; ++q) {
                const int eidx = d.port_edges[(size_t)p * d.port_max_edges + q];
                if (eidx < 0) continue;
                MPF_REQUIRE(eidx < d.n_boundary, "model: port edge index out of range");
                const int32_t* ed = d.boundary + (size_t)eidx * kb;
                double sp[3], m[3][3], v[3];
                for (int a = 0; a < kb; ++a) {
                    const double dx = d.nodes[2 * ed[a]] - d.nodes[2 * ed[0]];
                    const double dy = d.nodes[2 * ed[a] + 1] - d.nodes[2 * ed[0] + 1];
                    sp[a] = std::sqrt(dx * dx + dy * dy);
                }
                edge_integrals(kb, sp, m, v);
                for (int a = 0; a < kb; ++a) {
                    for (int b = 0; b < kb; ++b) consts.push_back({ed[a], slot(ed[a], ed[b]), iz * m[a][b]});
                    consts.push_back({ed[a], slot(ed[a], ep), -iz * v[a]});
                    consts.push_back({ep, slot(ep, ed[a]), -iz * v[a]});
                }
                consts.push_back({ep, slot(ep, ep), iz * sp[kb - 1]});
            }
        }
    }
    s.width = width;
    s.wpad = round_up(width, kSlotPad);
    const int np = s.np;
    const int wp = s.wpad;
    s.ids.assign((size_t)wp * np, 0);
    s.diag_slot.assign(np, 0);
    s.nnz = 0;
    for (int row = 0; row < np; ++row) {
        for (int q = 0; q < wp; ++q) {
            int col = (row < Nu && q < cnt[row]) ? cols[(size_t)row * cap + q] : row;
            s.ids[(size_t)q * np + row] = col;
            if (row < Nu && q < cnt[row] && col == row) s.diag_slot[row] = q;
        }
        if (row < Nu) s.nnz += cnt[row];
    }
    s.group_width.assign(np / 16, 0);
    for (int row = 0; row < Nu; ++row)
        s.group_width[row / 16] = (uint8_t)std::max<int>(s.group_width[row / 16], cnt[row]);
    // gather plan, contributions in element order per ELL entry
    const size_t nent = (size_t)wp * np;
    s.ctb_off.assign(nent + 1, 0);
    for (int e = 0; e < E; ++e) {
        const int32_t* el = d.elements + (size_t)e * k;
        for (int i = 0; i < k; ++i) for (int j = 0; j < k; ++j)
            s.ctb_off[(size_t)slot_of[((size_t)e * k + i) * k + j] * np + el[i] + 1]++;
    }
    for (const auto& c : consts) s.ctb_off[(size_t)c.sl * np + c.row + 1]++;
    for (size_t i = 0; i < nent; ++i) s.ctb_off[i + 1] += s.ctb_off[i];
    const size_t total = (size_t)E * k * k + consts.size();
    s.ctb_elem.resize(total); s.ctb_k.resize(total); s.ctb_m.resize(total);
    std::vector<int32_t> fill(s.ctb_off.begin(), s.ctb_off.end() - 1);
    for (int e = 0; e < E; ++e) {
        const int32_t* el = d.elements + (size_t)e * k;
        for (int i = 0; i < k; ++i) for (int j = 0; j < k; ++j) {
            size_t ent = (size_t)slot_of[((size_t)e * k + i) * k + j] * np + el[i];
            int32_t pos = fill[ent]++;
            s.ctb_elem[pos] = e;
            s.ctb_k[pos] = Ke[((size_t)e * k + i) * k + j];
            s.ctb_m[pos] = Me[((size_t)e * k + i) * k + j];
        }
    }
    for (const auto& c : consts) {
        const int32_t pos = fill[(size_t)c.sl * np + c.row]++;
        s.ctb_elem[pos] = -1;   // sigma-independent
        s.ctb_k[pos] = c.v;
        s.ctb_m[pos] = 0.0;
    }

    // gap-model excitation B (equation.cpp:183-215); B stored sparse per node.  With the
    // complete electrode model the current enters at the electrode rows instead: B = [0; I].
    const int kb = d.nodes_per_edge;
    std::vector<std::map<int, double>> Bn(Nu);
    if (cem && !coupled)
        for (int p = 0; p < s.P; ++p) Bn[N + p][p] = 1.0;
    for (int p = 0; p < (cem ? 0 : s.P); ++p) {
        std::vector<int> edges;
        for (int q = 0; q < d.port_max_edges; ++q) {
            int eidx = d.port_edges[(size_t)p * d.port_max_edges + q];
            if (eidx < 0) continue;
            MPF_REQUIRE(eidx < d.n_boundary, "model: port edge index out of range");
            edges.push_back(eidx);
        }
        double length = 0.0;
        for (int eidx : edges) {
            const int32_t* ed = d.boundary + (size_t)eidx * kb;
            double dx = d.nodes[2 * ed[kb - 1]] - d.nodes[2 * ed[0]], dy = d.nodes[2 * ed[kb - 1] + 1] - d.nodes[2 * ed[0] + 1];
            length += std::sqrt(dx * dx + dy * dy);
        }
        for (int eidx : edges) {
            const int32_t* ed = d.boundary + (size_t)eidx * kb;
            double sp[3];
            for (int q = 0; q < kb; ++q) {
                double dx = d.nodes[2 * ed[q]] - d.nodes[2 * ed[0]], dy = d.nodes[2 * ed[q] + 1] - d.nodes[2 * ed[0] + 1];
                sp[q] = std::sqrt(dx * dx + dy * dy);
            }
            double w[3];
            if (kb == 2) {
                w[0] = w[1] = 0.5 * sp[1];
            } else {
                // 1D quadratic Lagrange basis on nodes (0, s1, s2) integrated over [0, s2]
                double s1 = sp[1], s2 = sp[2];
                double i0 = s2 - s2 * s2 * (s1 + s2) / (2.0 * s1 * s2) + s2 * s2 * s2 / (3.0 * s1 * s2);
                double i1 = (s2 * s2 * s2 / 3.0 - s2 * s2 * s2 / 2.0) / (s1 * (s1 - s2));
                double i2 = (s2 * s2 * s2 / 3.0 - s1 * s2 * s2 / 2.0) / (s2 * (s2 - s1));
                w[0] = i0; w[1] = i1;
                // the generated code evaluates the last node with node 1's coefficients
                // (src/fem/codegen/templates/basis.cpp.mako:72-84)
                w[2] = d.quirks ? i1 : i2;
            }
            for (int q = 0; q < kb; ++q) Bn[ed[q]][p] += w[q] / length;
        }
    }
    // Fixed sources (applyMixedBoundaryConditionKernel, eit_kernel.cu:34-54): the
    // excitation is clipped to B_clip = (|B| >= 1e-9); every row with a nonzero clipped
    // entry becomes an identity row
    MPF_REQUIRE(d.source_type == MPF_SOURCE_CURRENT || d.source_type == MPF_SOURCE_VOLTAGE,
                "model: unknown source type");
    s.fixed = d.source_type == MPF_SOURCE_VOLTAGE ? 1 : 0;
    if (s.fixed) {
        MPF_REQUIRE(!cem, "model: voltage sources use the gap model (contact_impedance 0)");
        MPF_REQUIRE(!d.jacobian_exact, "model: voltage sources use the reference Jacobian (jacobian_exact 0)");
        for (int n = 0; n < Nu; ++n)
            for (auto& kv : Bn[n]) kv.second = std::fabs(kv.second) >= 1e-9 ? 1.0 : 0.0;
        fixed_setup(s, Bn, Nu);
    }
    // F = B * Pi for the unique columns, Pi = [I*Dpat | Mpat]  (sources.cpp:57-71)
    s.F.assign((size_t)np * s.Rp, 0.0);
    auto Pi = [&](int p, int j) {
        return j < s.D ? d.drive_current[j] * d.drive_pattern[(size_t)p * s.D + j]
                       : d.meas_pattern[(size_t)p * s.M + (j - s.D)];
    };
    if (coupled) {
        // the solves' right-hand sides: the L unit electrode sources Bc[:, l]; Pi and the
        // representative of each unique pattern column for the Schur complement
        for (int p = 0; p < s.P; ++p)
            for (int t = s.bc_off[p]; t < s.bc_off[p + 1]; ++t) s.F[(size_t)s.bc_row[t] * s.Rp + p] = s.bc_val[t];
        s.pi.assign((size_t)s.P * s.R, 0.0);
        for (int p = 0; p < s.P; ++p)
            for (int j = 0; j < s.R; ++j) s.pi[(size_t)p * s.R + j] = Pi(p, j);
        s.prep.assign(s.pRu, -1);
        for (int j = 0; j < s.R; ++j)
            if (s.prep[s.colmap[j]] < 0) s.prep[s.colmap[j]] = j;
        for (int u = 0; u < s.pRu; ++u) s.colscale[s.prep[u]] = 1.0;
        s.w_off.assign(s.M + 1, 0);
        s.w_node.clear(); s.w_val.clear();
        coarse_setup(s);
        laplacian_setup(d, s);
        return;
    }
    std::vector<int> rep(s.Ru, -1);   // first original column of each unique column
    for (int j = 0; j < s.R; ++j)
        if (rep[s.colmap[j]] < 0) rep[s.colmap[j]] = j;
    for (int n = 0; n < Nu; ++n)
        for (auto& kv : Bn[n])
            for (int u = 0; u < s.Ru; ++u) s.F[(size_t)n * s.Rp + u] += kv.second * Pi(kv.first, rep[u]);
    // a representative column has scale 1 by construction
    for (int u = 0; u < s.Ru; ++u) s.colscale[rep[u]] = 1.0;
    // W = Mpat^T B^T  (eit.cpp:93-94) as CSR (with the electrode model: Mpat^T on the
    // electrode potentials)
    s.w_off.assign(s.M + 1, 0);
    s.w_node.clear(); s.w_val.clear();
    for (int m = 0; m < s.M; ++m) {
        for (int n = 0; n < Nu; ++n) {
            double v = 0.0;
            bool any = false;
            for (auto& kv : Bn[n]) { v += d.meas_pattern[(size_t)kv.first * s.M + m] * kv.second; any = true; }
            // voltage sources measure height * (A_c Phi) through the clipped attachment
            if (s.fixed) v *= d.height;
            if (any && v != 0.0) { s.w_node.push_back(n); s.w_val.push_back(v); }
        }
        s.w_off[m + 1] = (int32_t)s.w_node.size();
    }
    coarse_setup(s);
    laplacian_setup(d, s);
}

}  // namespace mpf
