// The Tikhonov normal-equation solve of mpf_normal_solve without any allocation, for
// callers that cannot allocate device memory (code under HIP graph capture, serving loops,
// harnesses that forbid hipMalloc).  Its work space lives in a file-scope device array of
// fixed size; each GPU holds its own copy for the life of the process, and every call on
// that GPU shares it: calls that overlap on one GPU race.
#include <mutex>

#include "dense.h"
#include "model.h"
#include "noalloc.h"

namespace mpf {

// largest problem the fixed work space holds
constexpr int kNaMaxRows = 1024;     // Jacobian rows (ceil16(M) ceil16(D) at 32 x 32 patterns)
constexpr int kNaMaxCols = 65536;    // Jacobian columns (elements)
constexpr int kNaGramSlabs = 4;      // contraction slabs of the Gram kernel
constexpr size_t kNaScratchBytes = (size_t)24 << 20;

}  // namespace mpf

// the work space (device addresses come from hipGetSymbolAddress, per device)
static __device__ __align__(256) unsigned char scratch[mpf::kNaScratchBytes];

namespace mpf {

namespace {

constexpr int kNaMaxDevices = 64;
std::mutex g_na_mu;
void* g_na_ptr[kNaMaxDevices] = {};

__global__ void k_na_zero_int(int* __restrict__ p) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = 0;
}

size_t align256(size_t b) { return (b + 255) / 256 * 256; }

// offsets of the work-space parts for an m x E problem
struct NaLayout {
    size_t G, A, y, dy, res, vE, gemv, gram, pairs, info, total;
};

NaLayout na_layout(int m, int E) {
    NaLayout L{};
    size_t o = 0;
    L.G = o;     o += align256((size_t)m * m * sizeof(float));
    L.A = o;     o += align256((size_t)m * m * sizeof(double));
    L.y = o;     o += align256((size_t)m * sizeof(double));
    L.dy = o;    o += align256((size_t)m * sizeof(double));
    L.res = o;   o += align256((size_t)m * sizeof(double));
    L.vE = o;    o += align256((size_t)E * sizeof(double));
    L.gemv = o;  o += align256(gemv_n_workspace(m, E, 1));
    L.gram = o;  o += align256(gram_nt_workspace_bytes(m, E, 1, kNaGramSlabs));
    L.pairs = o; o += align256(cholesky_pairs(m) * sizeof(int2));
    L.info = o;  o += align256(sizeof(int));
    L.total = o;
    return L;
}

}  // namespace

void* noalloc_scratch(int device) {
    MPF_REQUIRE(device >= 0 && device < kNaMaxDevices, "no-alloc work space: device index out of range");
    std::lock_guard<std::mutex> lk(g_na_mu);
    if (!g_na_ptr[device]) {
        // the caller has made `device` current: its copy of the array is the one resolved
        void* p = nullptr;
        MPF_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        MPF_REQUIRE(p != nullptr, "no-alloc work space: no device address");
        g_na_ptr[device] = p;
    }
    return g_na_ptr[device];
}

void normal_solve_noalloc(mpf_ctx* ctx, int rows, int cols, const float* d_J, int ldj, double shift,
                          const double* d_r, int gn_refine, double* d_delta) {
    MPF_REQUIRE(rows >= 1 && rows <= kNaMaxRows, "mpf_normal_solve_noalloc: rows above the 1024 the work space holds");
    MPF_REQUIRE(cols >= 1 && cols <= kNaMaxCols,
                "mpf_normal_solve_noalloc: columns above the 65536 the work space holds");
    MPF_REQUIRE(gram_nt_applies(rows, ldj), "mpf_normal_solve_noalloc: rows and ldj must be multiples of 4");
    const NaLayout L = na_layout(rows, cols);
    MPF_REQUIRE(L.total <= kNaScratchBytes, "mpf_normal_solve_noalloc: problem exceeds the work space");
    char* base = (char*)noalloc_scratch(ctx->device);
    float* G = (float*)(base + L.G);
    double* A = (double*)(base + L.A);
    double* y = (double*)(base + L.y);
    double* dy = (double*)(base + L.dy);
    double* res = (double*)(base + L.res);
    double* vE = (double*)(base + L.vE);
    double* work = (double*)(base + L.gemv);
    void* gw = base + L.gram;
    int2* pairs = (int2*)(base + L.pairs);
    int* info = (int*)(base + L.info);
    hipStream_t st = ctx->stream;
    const int m = rows, E = cols;
    // the same kernels as mpf_normal_solve: MFMA Gram, FP64 Cholesky, FP64-residual passes
    launch_gram_nt(d_J, 0, m, E, ldj, G, 0, 1, gw, st, kNaGramSlabs);
    k_na_zero_int<<<1, 64, 0, st>>>(info);
    MPF_HIP(hipGetLastError());
    launch_cholesky(m, G, 0, shift, A, 0, 1, info, pairs, st);
    launch_potrs(m, A, 0, d_r, 0, y, 0, 1, st);
    for (int k = 0; k < gn_refine; ++k) {
        launch_gemv_t(m, E, d_J, ldj, 0, y, 0, vE, 0, 1, st);
        launch_gemv_n(m, E, d_J, ldj, 0, vE, 0, shift, y, 0, d_r, 0, res, 0, 1, work, st);
        launch_potrs(m, A, 0, res, 0, dy, 0, 1, st);
        launch_add_to(m, y, dy, st);
    }
    launch_gemv_t(m, E, d_J, ldj, 0, y, 0, d_delta, 0, 1, st);
    int h = 0;
    MPF_HIP(hipMemcpyAsync(&h, info, sizeof(int), hipMemcpyDeviceToHost, st));
    MPF_HIP(hipStreamSynchronize(st));
    MPF_REQUIRE(h == 0, "mpf_normal_solve_noalloc: normal-equation matrix is not positive definite");
}

}  // namespace mpf
