// Internal declarations shared by the host set-up, the HIP kernels and the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/mpflow_amd.h"

namespace mpf {

constexpr int kMaxSlices = 8;      // harmonics per job
constexpr int kTileRows = 256;     // rows per work tile of the sparse kernels
constexpr int kAggPerTile = 4;     // coarse aggregates per tile: its four 64-row geometric quarters
constexpr int kAggRows = 64;       // rows per coarse aggregate (a tile's 64-node bisection leaves)
constexpr int kCoarseRows = 64;    // coarse rows per block of the coarse-correction kernel
constexpr int kChunkTiles = kCoarseRows / kAggPerTile;   // tiles per partial-sum chunk
constexpr int kCoarseMaxTiles = 512;   // largest tile count per job with a coarse level
constexpr int kBlock = 256;        // threads per SpMV block: 4 wavefronts x 64 rows
constexpr int kEwBlock = 1024;     // threads per streaming / reduction block
constexpr int kUpdBlock = 256;     // threads per block of the PCG vector updates (one tile each)
constexpr int kSlotPad = 8;        // ELL width is padded to a multiple of this

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define MPF_HIP(call)                                                              \
    do {                                                                           \
        hipError_t e_ = (call);                                                    \
        if (e_ != hipSuccess)                                                      \
            throw ::mpf::Error(MPF_EHIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define MPF_REQUIRE(cond, msg)                                                     \
    do {                                                                           \
        if (!(cond)) throw ::mpf::Error(MPF_EINVAL, msg);                          \
    } while (0)

inline int ceil16(int n) { return (n % 16 == 0 || n == 1) ? n : (n / 16 + 1) * 16; }
inline int round_up(int n, int m) { return (n + m - 1) / m * m; }

// Device buffer with RAII.
template <class T>
struct DBuf {
    T* p = nullptr;
    size_t n = 0;
    DBuf() = default;
    explicit DBuf(size_t count) { alloc(count); }
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    DBuf(DBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DBuf& operator=(DBuf&& o) noexcept { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; return *this; }
    ~DBuf() { release(); }
    void alloc(size_t count) {
        release();
        n = count;
        if (count) {
            hipError_t e = hipMalloc(&p, count * sizeof(T));
            if (e != hipSuccess) { p = nullptr; n = 0; throw Error(MPF_ENOMEM, "hipMalloc failed"); }
        }
    }
    void release() { if (p) (void)hipFree(p); p = nullptr; n = 0; }
    void upload(const T* h, size_t count, hipStream_t s) {
        MPF_HIP(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
    }
    size_t bytes() const { return n * sizeof(T); }
};

// ---------------------------------------------------------------------------
// Host-side finite-element set-up (setup.cpp)
// ---------------------------------------------------------------------------
struct MeshSetup {
    int N = 0, E = 0, k = 0, np = 0, width = 0, wpad = 0, nnz = 0;
    int P = 0, D = 0, M = 0, R = 0, Rp = 0, Mp = 0, Dp = 0, mrows = 0;
    // right-hand-side deduplication: column j of Pi = colscale[j] * unique column colmap[j];
    // the device solves Ru unique columns (Rp lanes >= Ru)
    int Ru = 0;
    std::vector<int32_t> colmap;
    std::vector<double> colscale;
    // ELL pattern, slot-major: [wpad][np]; padding slots id = row, value 0
    std::vector<int32_t> ids;
    std::vector<int32_t> diag_slot;               // [np]
    // per 16-row group: the widest row's entry count (the SpMV skips slots beyond it)
    std::vector<uint8_t> group_width;             // [np / 16]
    // device numbering: perm[caller node] = device row
    std::vector<int32_t> perm;
    // [np] coarse aggregate of each row within its tile (0 .. kAggPerTile - 1)
    std::vector<uint8_t> agg_q;
    // gather plan: contributions of (element, local ij) to ELL entry slot*np+row
    std::vector<int32_t> ctb_off;                 // [wpad*np + 1]
    std::vector<int32_t> ctb_elem;                // [E*k*k]
    std::vector<double> ctb_k, ctb_m;             // element stiffness / mass entries
    // element data for the Jacobian: K_e and M_e row-major k x k
    std::vector<float> elem_k, elem_m;            // [E][k*k]
    std::vector<int32_t> elements;                // [E][k]
    // excitation F = B * Pi (FP64), [np][Rp]
    std::vector<double> F;
    // measurement W = Mpat^T B^T as CSR over M rows
    std::vector<int32_t> w_off, w_node;
    std::vector<double> w_val;
    // 2.5D constants per harmonic
    int C = 1;
    double alpha[kMaxSlices] = {}, beta[kMaxSlices] = {};
    double ref = 1.0, height = 1.0, port_height = 1.0;
    int logarithmic = 0, jacobian_exact = 0;
    // two-level preconditioner: tile count (0: none), [C][na][ld] inverses of the coarse
    // matrices at sigma = 1 (na = kAggPerTile * nt aggregates), per-aggregate sigma = 1
    // stiffness-diagonal sums
    int coarse_nt = 0, coarse_ld = 0;   // tiles, padded row length of coarse_inv
    std::vector<float> coarse_inv;
    std::vector<double> coarse_d1;
    // Fixed (voltage) sources: Dirichlet rows D are identity rows of the device matrix and
    // columns of D are eliminated; the eliminated couplings of each other row i
    // (i, j in D) and the full rows of D are kept as element contributions (CSR over rows:
    // column node, element, K_e and M_e entries) for the right-hand side
    // -A_iD (B_clip Pi)_D and the electrode currents A_c Phi on D
    int fixed = 0;
    std::vector<int32_t> cpl_off, cpl_node, cpl_elem;
    std::vector<double> cpl_k, cpl_m;
    std::vector<int32_t> drow_off, drow_node, drow_elem;
    std::vector<double> drow_k, drow_m;
    // coupled complete electrode model (cem_coupled): contact impedance z' of every mode,
    // electrode lengths |e_l|, the unit electrode sources Bc as CSR over ports (device
    // rows), n_c = int cos^2, Pi (P x R row-major), the pattern's unique columns (pRu, lane
    // count pRp) and the representative column of each
    int cem_coupled = 0;
    double zp = 0.0;
    std::vector<double> elen;
    std::vector<int32_t> bc_off, bc_row;
    std::vector<double> bc_val;
    double n_c[kMaxSlices] = {};
    std::vector<double> pi;
    int pRu = 0, pRp = 0;
    std::vector<int32_t> prep;
    // laplacian regularisation matrix: diagonal [E], neighbours across the three vertex
    // edges [E][3] (-1: boundary edge) and their entries [E][3]
    std::vector<double> lap_diag;
    std::vector<int32_t> lap_nb;
    std::vector<double> lap_val;
};

void build_setup(const mpf_model_desc& d, MeshSetup& s);

// element integrals (K_e, M_e) of one straight triangle, P1 or P2, row-major k x k
void element_integrals(int k, const double* xy, double* K, double* Mm);

// Ports::circularBoundary restatement (host)
std::vector<std::vector<int32_t>> circular_ports(int N, const double* nodes, int E, int k,
    const int32_t* elements, int nb, int kb, const int32_t* boundary, int count,
    double width, double offset, bool clockwise);

}  // namespace mpf
