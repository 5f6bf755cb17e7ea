// C ABI of mpflow_amd: every entry point catches, records the message and returns a
// status code (the reference throws std::invalid_argument / runtime_error instead).
#include <chrono>
#include <cstring>
#include <string>

#include "mlp.h"
#include "noalloc.h"
#include "fem.h"
#include "numeric.h"
#include "model.h"
#include "complex.h"

using mpf::Error;

namespace {
thread_local std::string g_last_error;

template <class F>
int guarded(F&& f) {
    try {
        g_last_error.clear();
        return f();
    } catch (const Error& e) {
        g_last_error = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        g_last_error = "host allocation failed";
        return MPF_ENOMEM;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return MPF_EINVAL;
    }
}

void set_device(mpf_ctx* ctx) { MPF_HIP(hipSetDevice(ctx->device)); }
}  // namespace

struct mpf_model {
    mpf::Model* impl = nullptr;             // real conductivity
    mpf::ComplexModel* cimpl = nullptr;     // complex conductivity (desc.complex_gamma)
    mpf_ctx* ctx = nullptr;
};

struct mpf_batch {
    mpf::Batch* impl;
};

struct mpf_mlp {
    mpf::Mlp* impl;
    mpf_ctx* ctx;
};

extern "C" {

const char* mpf_version(void) { return "mpflow_amd 0.1.0 (gfx950)"; }

const char* mpf_last_error(void) { return g_last_error.c_str(); }

int mpf_ctx_create(int device, void* hip_stream, mpf_ctx** out) {
    return guarded([&] {
        MPF_REQUIRE(out != nullptr, "mpf_ctx_create: out == nullptr");
        int n = 0;
        MPF_HIP(hipGetDeviceCount(&n));
        MPF_REQUIRE(device >= 0 && device < n, "mpf_ctx_create: no such device");
        auto* c = new mpf_ctx;
        c->device = device;
        MPF_HIP(hipSetDevice(device));
        // the device address of this GPU's fixed work space (mpf_normal_solve_noalloc)
        mpf::noalloc_scratch(device);
        if (hip_stream) {
            c->stream = (hipStream_t)hip_stream;
            c->own_stream = false;
        } else {
            MPF_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
            c->own_stream = true;
        }
        *out = c;
        return MPF_OK;
    });
}

int mpf_ctx_destroy(mpf_ctx* ctx) {
    return guarded([&] {
        if (!ctx) return MPF_OK;
        if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
        for (auto& e : ctx->ev) {
            (void)hipEventDestroy(e.first);
            (void)hipEventDestroy(e.second);
        }
        if (ctx->own_stream && ctx->stream) (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return MPF_OK;
    });
}

int mpf_ctx_synchronize(mpf_ctx* ctx) {
    return guarded([&] {
        MPF_REQUIRE(ctx, "mpf_ctx_synchronize: ctx == nullptr");
        MPF_HIP(hipStreamSynchronize(ctx->stream));
        return MPF_OK;
    });
}

void* mpf_ctx_stream(mpf_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int mpf_ctx_profile(mpf_ctx* ctx, int32_t enable) {
    return guarded([&] {
        MPF_REQUIRE(ctx, "mpf_ctx_profile: ctx == nullptr");
        ctx->prof = enable != 0;
        ctx->prof_ms = 0.0;
        ctx->prof_launches = 0;
        ctx->prof_bytes = 0.0;
        ctx->ev_used = 0;
        return MPF_OK;
    });
}

int mpf_ctx_profile_read(mpf_ctx* ctx, double out[3]) {
    return guarded([&] {
        MPF_REQUIRE(ctx && out, "mpf_ctx_profile_read: null argument");
        out[0] = ctx->prof_ms;
        out[1] = (double)ctx->prof_launches;
        out[2] = ctx->prof_bytes;
        return MPF_OK;
    });
}

int mpf_ctx_stream_probe(mpf_ctx* ctx, int32_t pattern, int32_t loads, const float* d_x, float* d_y, int64_t n,
                         int32_t blocks, int32_t reps, double* gbps) {
    return guarded([&] {
        MPF_REQUIRE(ctx && d_x && d_y && gbps, "mpf_ctx_stream_probe: null argument");
        MPF_REQUIRE(n > 0, "mpf_ctx_stream_probe: n <= 0");
        set_device(ctx);
        *gbps = mpf::stream_probe(pattern, loads, d_x, d_y, (size_t)n, blocks, reps, ctx->stream);
        return MPF_OK;
    });
}

void mpf_solve_opts_default(mpf_solve_opts* o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->stop_rule = MPF_STOP_REFERENCE;
    o->tol = 1e-6;
    o->max_iter = 0;
    o->precision = MPF_PREC_FP32;
    o->refine_steps = 2;
    o->refine_tol = 1e-3;
    o->krylov = MPF_KRYLOV_CG;
    o->check_every = 16;
}

int mpf_ctx_autotune_read(mpf_ctx* ctx, int32_t which, int64_t* keys, int32_t* values, int32_t cap,
                          int32_t* count) {
    return guarded([&] {
        MPF_REQUIRE(ctx && count, "mpf_ctx_autotune_read: null argument");
        MPF_REQUIRE(which == 0 || which == 1, "mpf_ctx_autotune_read: which must be 0 or 1");
        std::lock_guard<std::mutex> lk(ctx->tune.mu);
        const auto& mp = which == 0 ? ctx->tune.spmv : ctx->tune.gram;
        int n = 0;
        for (const auto& kv : mp) {
            if (n < cap && keys && values) { keys[n] = kv.first; values[n] = kv.second; }
            ++n;
        }
        *count = n;
        return MPF_OK;
    });
}

int mpf_ports_circular(int32_t n_nodes, const double* nodes, int32_t n_elements,
                       int32_t nodes_per_element, const int32_t* elements, int32_t n_boundary,
                       int32_t nodes_per_edge, const int32_t* boundary, int32_t count, double width,
                       double offset, int32_t clockwise, int32_t* out, int32_t out_capacity,
                       int32_t* max_edges_out) {
    return guarded([&] {
        MPF_REQUIRE(nodes && elements && boundary && max_edges_out, "mpf_ports_circular: null argument");
        auto ports = mpf::circular_ports(n_nodes, nodes, n_elements, nodes_per_element, elements, n_boundary,
                                         nodes_per_edge, boundary, count, width, offset, clockwise != 0);
        int w = 0;
        for (auto& p : ports) w = std::max<int>(w, (int)p.size());
        *max_edges_out = w;
        if (out) {
            MPF_REQUIRE(out_capacity >= count * w, "mpf_ports_circular: out too small");
            for (int p = 0; p < count; ++p)
                for (int q = 0; q < w; ++q) out[p * w + q] = q < (int)ports[p].size() ? ports[p][q] : -1;
        }
        return MPF_OK;
    });
}

int mpf_element_integrals(int32_t nodes_per_element, const double* xy, double* K, double* M) {
    return guarded([&] {
        MPF_REQUIRE(xy && K && M, "mpf_element_integrals: null argument");
        MPF_REQUIRE(nodes_per_element == 3 || nodes_per_element == 6, "mpf_element_integrals: k must be 3 or 6");
        mpf::element_integrals(nodes_per_element, xy, K, M);
        return MPF_OK;
    });
}

int mpf_model_create(mpf_ctx* ctx, const mpf_model_desc* desc, mpf_model** out) {
    return guarded([&] {
        MPF_REQUIRE(ctx && desc && out, "mpf_model_create: null argument");
        set_device(ctx);
        auto* m = new mpf_model;
        m->ctx = ctx;
        try {
            if (desc->complex_gamma) m->cimpl = new mpf::ComplexModel(ctx, *desc);
            else m->impl = new mpf::Model(ctx, *desc);
        } catch (...) {
            delete m;
            throw;
        }
        *out = m;
        return MPF_OK;
    });
}

int mpf_model_destroy(mpf_model* m) {
    return guarded([&] {
        if (!m) return MPF_OK;
        if (m->impl || m->cimpl) (void)hipStreamSynchronize(m->ctx->stream);
        delete m->impl;
        delete m->cimpl;
        delete m;
        return MPF_OK;
    });
}

int mpf_model_dims(const mpf_model* m, int64_t dims[16]) {
    return guarded([&] {
        MPF_REQUIRE(m && dims, "mpf_model_dims: null argument");
        const auto& S = m->impl ? m->impl->S : m->cimpl->setup();
        dims[0] = S.N; dims[1] = S.E; dims[2] = S.np; dims[3] = S.width; dims[4] = S.mrows;
        dims[5] = S.Mp; dims[6] = S.Dp; dims[7] = S.R; dims[8] = S.Rp;
        dims[9] = m->impl ? m->impl->frames : m->cimpl->frames();
        dims[10] = S.C; dims[11] = S.nnz; dims[12] = S.Ru; dims[13] = S.wpad;
        dims[14] = 0; dims[15] = 0;
        return MPF_OK;
    });
}

int mpf_eit_solve(mpf_model* m, const float* d_gamma, float* d_U, double* d_U64, float* d_J,
                  const mpf_solve_opts* opts, mpf_solve_stats* stats) {
    return guarded([&] {
        MPF_REQUIRE(m && d_gamma, "mpf_eit_solve: null argument");
        auto t0 = std::chrono::steady_clock::now();
        set_device(m->ctx);
        mpf_solve_opts o;
        if (opts) o = *opts; else mpf_solve_opts_default(&o);
        if (stats) std::memset(stats, 0, sizeof(*stats));
        if (m->cimpl) {
            // complex conductivity: gamma, U, U64 and J hold (re, im) pairs
            int rc = m->cimpl->forward(o, reinterpret_cast<const float2*>(d_gamma), reinterpret_cast<float2*>(d_U),
                                       reinterpret_cast<mpf::zd*>(d_U64), reinterpret_cast<float2*>(d_J), stats);
            if (stats) stats->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            if (rc == MPF_ENOCONV) g_last_error = "mpf_eit_solve: a Krylov solve hit its iteration cap";
            return rc;
        }
        auto& M = *m->impl;
        MPF_HIP(hipMemcpyAsync(M.gamma.p, d_gamma, sizeof(float) * (size_t)M.S.E * M.frames,
                               hipMemcpyDeviceToDevice, M.ctx->stream));
        int rc = M.forward(o, d_U, d_U64, d_J, stats);
        if (stats) stats->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (rc == MPF_ENOCONV) g_last_error = "mpf_eit_solve: a Krylov solve hit its iteration cap";
        return rc;
    });
}

int mpf_gn_set_regularization(mpf_model* m, int32_t type, const float* d_gamma_ref, const mpf_solve_opts* opts,
                              mpf_solve_stats* stats) {
    return guarded([&] {
        MPF_REQUIRE(m, "mpf_gn_set_regularization: null model");
        MPF_REQUIRE(type == MPF_REG_IDENTITY || type == MPF_REG_NOSER || type == MPF_REG_LAPLACIAN,
                    "mpf_gn_set_regularization: unknown regularisation type");
        MPF_REQUIRE(m->impl, "mpf_gn_set_regularization: not available for a complex-conductivity model");
        set_device(m->impl->ctx);
        mpf_solve_opts o;
        if (opts) o = *opts; else mpf_solve_opts_default(&o);
        if (stats) std::memset(stats, 0, sizeof(*stats));
        m->impl->set_regularization(type, d_gamma_ref, o, stats);
        return MPF_OK;
    });
}

int mpf_diff_prepare(mpf_model* m, const float* d_gamma_ref, double lambda, const mpf_solve_opts* opts,
                     mpf_solve_stats* stats) {
    return guarded([&] {
        MPF_REQUIRE(m && d_gamma_ref, "mpf_diff_prepare: null argument");
        MPF_REQUIRE(m->impl, "mpf_diff_prepare: not available for a complex-conductivity model");
        set_device(m->impl->ctx);
        mpf_solve_opts o;
        if (opts) o = *opts; else mpf_solve_opts_default(&o);
        if (stats) std::memset(stats, 0, sizeof(*stats));
        m->impl->diff_prepare(d_gamma_ref, lambda, o, stats);
        return MPF_OK;
    });
}

int mpf_set_sensitivity_mask(mpf_model* m, const float* d_mask) {
    return guarded([&] {
        MPF_REQUIRE(m, "mpf_set_sensitivity_mask: null model");
        MPF_REQUIRE(m->impl, "mpf_set_sensitivity_mask: not available for a complex-conductivity model");
        set_device(m->impl->ctx);
        m->impl->set_sensitivity_mask(d_mask);
        return MPF_OK;
    });
}

int mpf_diff_solve(mpf_model* m, const double* d_Umeas, int32_t gn_refine, float* d_gamma) {
    return guarded([&] {
        MPF_REQUIRE(m && d_Umeas && d_gamma, "mpf_diff_solve: null argument");
        MPF_REQUIRE(gn_refine >= 0, "mpf_diff_solve: gn_refine < 0");
        MPF_REQUIRE(m->impl, "mpf_diff_solve: not available for a complex-conductivity model");
        set_device(m->impl->ctx);
        m->impl->diff_solve(d_Umeas, gn_refine, d_gamma);
        return MPF_OK;
    });
}

int mpf_fem_assemble(mpf_model* m, const float* d_gamma, double alpha, uint32_t* d_column_ids, float* d_values) {
    return guarded([&] {
        MPF_REQUIRE(m && d_gamma && d_values, "mpf_fem_assemble: null argument");
        MPF_REQUIRE(m->impl, "mpf_fem_assemble: not available for a complex-conductivity model");
        set_device(m->impl->ctx);
        mpf::fem_assemble(*m->impl, d_gamma, alpha, d_column_ids, d_values);
        MPF_HIP(hipStreamSynchronize(m->impl->ctx->stream));
        return MPF_OK;
    });
}

int mpf_eit_measure(mpf_model* m, const float* d_phi, int32_t ldphi, int32_t additive, float* d_U) {
    return guarded([&] {
        MPF_REQUIRE(m && d_phi && d_U, "mpf_eit_measure: null argument");
        MPF_REQUIRE(m->impl, "mpf_eit_measure: not available for a complex-conductivity model");
        set_device(m->impl->ctx);
        mpf::fem_measure(*m->impl, d_phi, ldphi, additive, d_U);
        MPF_HIP(hipStreamSynchronize(m->impl->ctx->stream));
        return MPF_OK;
    });
}

int mpf_jacobian(mpf_model* m, const float* d_phi, int32_t ldphi, const float* d_gamma, int32_t additive,
                 float* d_J) {
    return guarded([&] {
        MPF_REQUIRE(m && d_phi && d_gamma && d_J, "mpf_jacobian: null argument");
        MPF_REQUIRE(m->impl, "mpf_jacobian: not available for a complex-conductivity model");
        set_device(m->impl->ctx);
        mpf::fem_jacobian(*m->impl, d_phi, ldphi, d_gamma, additive, d_J);
        MPF_HIP(hipStreamSynchronize(m->impl->ctx->stream));
        return MPF_OK;
    });
}

int mpf_gn_step(mpf_model* m, float* d_gamma, const double* d_Umeas, double lambda, int32_t gn_refine,
                const mpf_solve_opts* opts, double* d_Ucalc, mpf_solve_stats* stats) {
    return guarded([&] {
        MPF_REQUIRE(m && d_gamma && d_Umeas, "mpf_gn_step: null argument");
        auto t0 = std::chrono::steady_clock::now();
        set_device(m->ctx);
        mpf_solve_opts o;
        if (opts) o = *opts; else mpf_solve_opts_default(&o);
        if (stats) std::memset(stats, 0, sizeof(*stats));
        int rc;
        if (m->cimpl)
            // complex conductivity: gamma, U_meas and U_calc hold (re, im) pairs; the
            // normal equations are formed in FP64, so gn_refine has nothing to refine
            rc = m->cimpl->gn_step(reinterpret_cast<float2*>(d_gamma), reinterpret_cast<const mpf::zd*>(d_Umeas),
                                   lambda, o, reinterpret_cast<mpf::zd*>(d_Ucalc), stats);
        else
            rc = m->impl->gn_step(d_gamma, d_Umeas, lambda, gn_refine, o, d_Ucalc, stats);
        if (stats) stats->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        return rc;
    });
}

int mpf_batch_create(mpf_ctx* ctx, mpf_model* const* models, int32_t n_models, mpf_batch** out) {
    return guarded([&] {
        MPF_REQUIRE(ctx && models && out && n_models > 0, "mpf_batch_create: null argument or no models");
        set_device(ctx);
        std::vector<mpf::Model*> ms;
        for (int i = 0; i < n_models; ++i) {
            MPF_REQUIRE(models[i] && models[i]->impl, "mpf_batch_create: null or complex-conductivity model");
            ms.push_back(models[i]->impl);
        }
        auto* b = new mpf_batch;
        try {
            b->impl = new mpf::Batch(ctx, ms);
        } catch (...) {
            delete b;
            throw;
        }
        *out = b;
        return MPF_OK;
    });
}

int mpf_batch_destroy(mpf_batch* b) {
    return guarded([&] {
        if (!b) return MPF_OK;
        if (b->impl) {
            (void)hipStreamSynchronize(b->impl->ctx->stream);
            delete b->impl;
        }
        delete b;
        return MPF_OK;
    });
}

int mpf_batch_solve(mpf_batch* b, const float* const* d_gamma, float* const* d_U, double* const* d_U64,
                    float* const* d_J, const mpf_solve_opts* opts, mpf_solve_stats* stats) {
    return guarded([&] {
        MPF_REQUIRE(b && d_gamma, "mpf_batch_solve: null argument");
        auto t0 = std::chrono::steady_clock::now();
        set_device(b->impl->ctx);
        mpf_solve_opts o;
        if (opts) o = *opts; else mpf_solve_opts_default(&o);
        if (stats) std::memset(stats, 0, sizeof(*stats));
        int rc = b->impl->solve(d_gamma, d_U, d_U64, d_J, o, stats);
        if (stats) stats->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (rc == MPF_ENOCONV) g_last_error = "mpf_batch_solve: a Krylov solve hit its iteration cap";
        return rc;
    });
}

int mpf_batch_gn_step(mpf_batch* b, float* const* d_gamma, const double* const* d_Umeas, double lambda,
                      int32_t gn_refine, const mpf_solve_opts* opts, double* const* d_Ucalc,
                      mpf_solve_stats* stats) {
    return guarded([&] {
        MPF_REQUIRE(b && d_gamma && d_Umeas, "mpf_batch_gn_step: null argument");
        auto t0 = std::chrono::steady_clock::now();
        set_device(b->impl->ctx);
        mpf_solve_opts o;
        if (opts) o = *opts; else mpf_solve_opts_default(&o);
        if (stats) std::memset(stats, 0, sizeof(*stats));
        int rc = b->impl->gn_step(d_gamma, d_Umeas, lambda, gn_refine, o, d_Ucalc, stats);
        if (stats) stats->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        return rc;
    });
}

int mpf_ell_spmv(mpf_ctx* ctx, int32_t np, int32_t width, const int32_t* d_ids, const float* d_S,
                 const float* d_R, int32_t C, const float* alpha, const float* d_X, float* d_Y, int32_t R,
                 int32_t ld) {
    return guarded([&] {
        MPF_REQUIRE(ctx && d_ids && d_S && d_R && alpha && d_X && d_Y, "mpf_ell_spmv: null argument");
        MPF_REQUIRE(np > 0 && np % mpf::kTileRows == 0, "mpf_ell_spmv: np must be a positive multiple of 256");
        MPF_REQUIRE(ld == 16 || ld == 32 || ld == 64, "mpf_ell_spmv: ld must be 16, 32 or 64");
        MPF_REQUIRE(R >= 1 && R <= ld, "mpf_ell_spmv: R out of range");
        MPF_REQUIRE(C >= 1 && C <= mpf::kMaxSlices, "mpf_ell_spmv: C out of range");
        MPF_REQUIRE(width >= 8 && width <= 64 && width % 8 == 0, "mpf_ell_spmv: width must be 8, 16, ..., 64");
        set_device(ctx);
        mpf::ell_spmv(ctx, np, width, d_ids, d_S, d_R, C, alpha, d_X, d_Y, ld);
        return MPF_OK;
    });
}

int mpf_sparse_multiply(mpf_ctx* ctx, const float* d_values, const uint32_t* d_column_ids, const float* d_matrix,
                        uint32_t result_rows, uint32_t matrix_rows, uint32_t columns, uint32_t density,
                        float* d_result) {
    return guarded([&] {
        MPF_REQUIRE(ctx && d_values && d_column_ids && d_matrix && d_result, "mpf_sparse_multiply: null argument");
        MPF_REQUIRE(result_rows >= 1 && matrix_rows >= 1 && columns >= 1 && result_rows < (1u << 24) &&
                        matrix_rows < (1u << 24) && columns < (1u << 16),
                    "mpf_sparse_multiply: bad shape");
        MPF_REQUIRE(density <= 32, "mpf_sparse_multiply: density above the 32 slots of a row");
        set_device(ctx);
        mpf::sparse_multiply(ctx, d_values, d_column_ids, d_matrix, (int)result_rows, (int)matrix_rows, (int)columns,
                             (int)density, d_result);
        return MPF_OK;
    });
}

int mpf_gram_bf16(mpf_ctx* ctx, int32_t trans, int32_t rows, int32_t k, const float* d_A, int32_t lda,
                  float* d_G) {
    return guarded([&] {
        MPF_REQUIRE(ctx && d_A && d_G, "mpf_gram_bf16: null argument");
        MPF_REQUIRE(rows > 0 && k > 0 && lda >= rows, "mpf_gram_bf16: bad shape");
        set_device(ctx);
        const int n = trans ? k : rows, nk = trans ? rows : k;
        if (trans == 0 && mpf::gram_bf16_applies(n, lda)) {
            // one rounding pass to BF16, then the 256-tile kernel
            mpf::DBuf<uint16_t> ab((size_t)lda * nk);
            mpf::launch_to_bf16((size_t)lda * nk, d_A, ab.p, ctx->stream);
            mpf::DBuf<char> work(mpf::gram_bf16_workspace_bytes(n, nk, 1));
            mpf::launch_gram_bf16(ab.p, 0, n, nk, lda, d_G, 0, 1, work.p, ctx->stream);
            MPF_HIP(hipStreamSynchronize(ctx->stream));
        } else if (trans == 0 && mpf::gram_nt_applies(n, lda)) {
            mpf::DBuf<char> work(mpf::gram_nt_workspace_bytes(n, nk, 1));
            mpf::launch_gram_nt(d_A, 0, n, nk, lda, d_G, 0, 1, work.p, ctx->stream);
            MPF_HIP(hipStreamSynchronize(ctx->stream));
        } else {
            mpf::DBuf<char> work(mpf::gram_workspace_bytes(n, nk));
            mpf::launch_gram(d_A, trans, n, nk, lda, d_G, work.p, ctx->stream);
        }
        MPF_HIP(hipStreamSynchronize(ctx->stream));
        return MPF_OK;
    });
}

int mpf_gemv_t(mpf_ctx* ctx, int32_t rows, int32_t cols, const float* d_A, int32_t lda, const double* d_x,
               double* d_y) {
    return guarded([&] {
        MPF_REQUIRE(ctx && d_A && d_x && d_y, "mpf_gemv_t: null argument");
        MPF_REQUIRE(rows > 0 && cols > 0 && lda >= rows, "mpf_gemv_t: bad shape");
        set_device(ctx);
        mpf::launch_gemv_t(rows, cols, d_A, lda, 0, d_x, 0, d_y, 0, 1, ctx->stream);
        MPF_HIP(hipStreamSynchronize(ctx->stream));
        return MPF_OK;
    });
}

int mpf_spd_solve(mpf_ctx* ctx, int32_t n, const float* d_G, double shift, const double* d_b, double* d_y) {
    return guarded([&] {
        MPF_REQUIRE(ctx && d_G && d_b && d_y, "mpf_spd_solve: null argument");
        MPF_REQUIRE(n > 0, "mpf_spd_solve: n <= 0");
        set_device(ctx);
        mpf::DBuf<double> A((size_t)n * n);
        mpf::DBuf<int2> pairs(mpf::cholesky_pairs(n));
        mpf::DBuf<int> info(1);
        MPF_HIP(hipMemsetAsync(info.p, 0, sizeof(int), ctx->stream));
        mpf::launch_cholesky(n, d_G, 0, shift, A.p, 0, 1, info.p, pairs.p, ctx->stream);
        mpf::launch_potrs(n, A.p, 0, d_b, 0, d_y, 0, 1, ctx->stream);
        int h = 0;
        MPF_HIP(hipMemcpyAsync(&h, info.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        MPF_HIP(hipStreamSynchronize(ctx->stream));
        MPF_REQUIRE(h == 0, "mpf_spd_solve: matrix is not positive definite");
        return MPF_OK;
    });
}

int mpf_normal_solve(mpf_ctx* ctx, int32_t rows, int32_t cols, const float* d_J, int32_t ldj, double shift,
                     const double* d_r, int32_t gn_refine, double* d_delta) {
    return guarded([&] {
        MPF_REQUIRE(ctx && d_J && d_r && d_delta, "mpf_normal_solve: null argument");
        MPF_REQUIRE(rows > 0 && cols > 0 && ldj >= rows && gn_refine >= 0 && shift > 0.0,
                    "mpf_normal_solve: bad shape, shift or refinement count");
        set_device(ctx);
        hipStream_t st = ctx->stream;
        const int m = rows, E = cols;
        mpf::DBuf<float> G((size_t)m * m);
        mpf::DBuf<double> A((size_t)m * m), y(m), dy(m), res(m), vE(E);
        mpf::DBuf<double> work(mpf::gemv_n_workspace(m, E, 1) / sizeof(double) + 1);
        mpf::DBuf<int2> pairs(mpf::cholesky_pairs(m));
        mpf::DBuf<int> info(1);
        if (mpf::gram_nt_applies(m, ldj)) {
            mpf::DBuf<char> gw(mpf::gram_nt_workspace_bytes(m, E, 1));
            mpf::launch_gram_nt(d_J, 0, m, E, ldj, G.p, 0, 1, gw.p, st);
            MPF_HIP(hipStreamSynchronize(st));
        } else {
            mpf::DBuf<char> gw(mpf::gram_workspace_bytes(m, E));
            mpf::launch_gram(d_J, 0, m, E, ldj, G.p, gw.p, st);
            MPF_HIP(hipStreamSynchronize(st));
        }
        MPF_HIP(hipMemsetAsync(info.p, 0, sizeof(int), st));
        mpf::launch_cholesky(m, G.p, 0, shift, A.p, 0, 1, info.p, pairs.p, st);
        mpf::launch_potrs(m, A.p, 0, d_r, 0, y.p, 0, 1, st);
        for (int k = 0; k < gn_refine; ++k) {
            mpf::launch_gemv_t(m, E, d_J, ldj, 0, y.p, 0, vE.p, 0, 1, st);
            mpf::launch_gemv_n(m, E, d_J, ldj, 0, vE.p, 0, shift, y.p, 0, d_r, 0, res.p, 0, 1, work.p, st);
            mpf::launch_potrs(m, A.p, 0, res.p, 0, dy.p, 0, 1, st);
            mpf::launch_add_to(m, y.p, dy.p, st);
        }
        mpf::launch_gemv_t(m, E, d_J, ldj, 0, y.p, 0, d_delta, 0, 1, st);
        int h = 0;
        MPF_HIP(hipMemcpyAsync(&h, info.p, sizeof(int), hipMemcpyDeviceToHost, st));
        MPF_HIP(hipStreamSynchronize(st));
        MPF_REQUIRE(h == 0, "mpf_normal_solve: normal-equation matrix is not positive definite");
        return MPF_OK;
    });
}

int mpf_normal_solve_noalloc(mpf_ctx* ctx, int32_t rows, int32_t cols, const float* d_J, int32_t ldj,
                             double shift, const double* d_r, int32_t gn_refine, double* d_delta) {
    return guarded([&] {
        MPF_REQUIRE(ctx && d_J && d_r && d_delta, "mpf_normal_solve_noalloc: null argument");
        MPF_REQUIRE(rows > 0 && cols > 0 && ldj >= rows && gn_refine >= 0 && shift > 0.0,
                    "mpf_normal_solve_noalloc: bad shape, shift or refinement count");
        set_device(ctx);
        mpf::normal_solve_noalloc(ctx, rows, cols, d_J, ldj, shift, d_r, gn_refine, d_delta);
        return MPF_OK;
    });
}

}  // extern "C"

int mpf_mlp_create(mpf_ctx* ctx, int32_t n_layers, const int32_t* dims, uint64_t seed, mpf_mlp** out) {
    return guarded([&] {
        MPF_REQUIRE(ctx && dims && out && n_layers >= 1, "mpf_mlp_create: null argument or no layers");
        set_device(ctx);
        std::vector<int> d(dims, dims + n_layers + 1);
        auto* m = new mpf_mlp;
        m->ctx = ctx;
        try {
            m->impl = new mpf::Mlp(ctx->stream, d, seed);
        } catch (...) {
            delete m;
            throw;
        }
        *out = m;
        return MPF_OK;
    });
}

int mpf_mlp_destroy(mpf_mlp* m) {
    return guarded([&] {
        if (!m) return MPF_OK;
        if (m->impl) {
            (void)hipStreamSynchronize(m->ctx->stream);
            delete m->impl;
        }
        delete m;
        return MPF_OK;
    });
}

int mpf_mlp_fwd(mpf_mlp* m, int32_t batch, const float* d_X, float* d_Y) {
    return guarded([&] {
        MPF_REQUIRE(m && d_X, "mpf_mlp_fwd: null argument");
        set_device(m->ctx);
        m->impl->forward(batch, d_X, d_Y);
        MPF_HIP(hipStreamSynchronize(m->ctx->stream));
        return MPF_OK;
    });
}

int mpf_mlp_bwd(mpf_mlp* m, const float* d_T, double* loss) {
    return guarded([&] {
        MPF_REQUIRE(m && d_T, "mpf_mlp_bwd: null argument");
        set_device(m->ctx);
        const double l = m->impl->backward(d_T);
        if (loss) *loss = l;
        return MPF_OK;
    });
}

int mpf_mlp_adam(mpf_mlp* m, float lr, float beta1, float beta2, float eps) {
    return guarded([&] {
        MPF_REQUIRE(m, "mpf_mlp_adam: null argument");
        MPF_REQUIRE(lr > 0.0f && beta1 >= 0.0f && beta1 < 1.0f && beta2 >= 0.0f && beta2 < 1.0f && eps > 0.0f,
                    "mpf_mlp_adam: bad hyper-parameters");
        set_device(m->ctx);
        m->impl->adam(lr, beta1, beta2, eps);
        return MPF_OK;
    });
}

int64_t mpf_mlp_param_count(const mpf_mlp* m) { return m && m->impl ? m->impl->count() : -1; }

int mpf_mlp_params(mpf_mlp* m, int32_t which, int32_t set, float* host) {
    return guarded([&] {
        MPF_REQUIRE(m && host, "mpf_mlp_params: null argument");
        set_device(m->ctx);
        m->impl->params(which, set != 0, host);
        return MPF_OK;
    });
}

static int krylov_entry(const char* name, int krylov, mpf_ctx* ctx, int32_t n, int32_t density,
                        const uint32_t* d_cid, const float* d_val, int32_t R, const float* d_B, int32_t ldb,
                        float* d_X, int32_t ldx, int32_t dc_free, const mpf_solve_opts* opts,
                        mpf_solve_stats* stats) {
    return guarded([&] {
        MPF_REQUIRE(ctx && d_cid && d_val && d_B && d_X, std::string(name) + ": null argument");
        auto t0 = std::chrono::steady_clock::now();
        set_device(ctx);
        mpf_solve_opts o;
        if (opts) o = *opts; else mpf_solve_opts_default(&o);
        o.krylov = krylov;
        if (stats) std::memset(stats, 0, sizeof(*stats));
        const int rc = mpf::krylov_solve(ctx, n, density, d_cid, d_val, R, d_B, ldb, d_X, ldx, dc_free, o, stats);
        if (stats) stats->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (rc == MPF_ENOCONV) g_last_error = std::string(name) + ": the iteration cap was hit or the solve broke down";
        return rc;
    });
}

int mpf_pcg_solve(mpf_ctx* ctx, int32_t n, int32_t density, const uint32_t* d_column_ids, const float* d_values,
                  int32_t R, const float* d_B, int32_t ldb, float* d_X, int32_t ldx, int32_t dc_free,
                  const mpf_solve_opts* opts, mpf_solve_stats* stats) {
    return krylov_entry("mpf_pcg_solve", MPF_KRYLOV_CG, ctx, n, density, d_column_ids, d_values, R, d_B, ldb,
                        d_X, ldx, dc_free, opts, stats);
}

int mpf_bicgstab_solve(mpf_ctx* ctx, int32_t n, int32_t density, const uint32_t* d_column_ids,
                       const float* d_values, int32_t R, const float* d_B, int32_t ldb, float* d_X, int32_t ldx,
                       int32_t dc_free, const mpf_solve_opts* opts, mpf_solve_stats* stats) {
    return krylov_entry("mpf_bicgstab_solve", MPF_KRYLOV_BICGSTAB, ctx, n, density, d_column_ids, d_values, R,
                        d_B, ldb, d_X, ldx, dc_free, opts, stats);
}

int mpf_block_pcg_solve(mpf_ctx* ctx, int32_t n, int32_t density, const uint32_t* d_column_ids,
                        const float* d_values, int32_t R, const float* d_B, int32_t ldb, float* d_X, int32_t ldx,
                        int32_t dc_free, const mpf_solve_opts* opts, mpf_solve_stats* stats) {
    return krylov_entry("mpf_block_pcg_solve", MPF_KRYLOV_BLOCK_CG, ctx, n, density, d_column_ids, d_values, R,
                        d_B, ldb, d_X, ldx, dc_free, opts, stats);
}
