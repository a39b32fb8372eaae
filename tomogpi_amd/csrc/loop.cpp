// loop.cpp - device-resident simple-gradient reconstruction
// (Iter3D::doSimpleGradient, C/Iter3D_Algo_Gradient.cu:25-223) with the
// generalised-Gaussian beta=2 (quadratic) regulariser
// (C/GeneralizedGaussianRegularizer_CPU.cu:158-226).  Per iteration k:
//   est = P x;  r = y - est;  dJ = B r;  dJ *= -2
//   if lambda != 0:  dJ += (2 lambda) L(L x);
//                    if k < Kopt: normdJProjReg = 2 sum (L dJ)^2
//   if k < Kopt:  est = P dJ;  normdJ = sum dJ^2;
//                 normdJProj = 2 sum est^2 + lambda normdJProjReg
//                 alpha = -normdJ / normdJProj;  alpha_min = max(alpha, alpha_min)
//   else          alpha = alpha_min
//   x += (float)alpha dJ, clamped at 0 if positivity
// With a 16-bit storage type every operator reads a 16-bit copy of its FP32
// input (the mixed-precision form); the rest of the loop is unchanged.
// Volumes and sinograms stay on the device; only the FP64 norms return to the
// host for the step rule.  The SIRT loop (tomo_sirt_run) follows.
#include <algorithm>
#include <cmath>

#include "tomo_internal.hpp"

namespace tomo {
namespace {

struct DevBuf {
    void* p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
};

// A loop's work buffer from the stream-ordered pool of the loop's stream,
// released on that stream (after the work queued on it), so allocating and
// releasing it neither waits for the device nor stalls other streams.
struct PoolBuf {
    void* p = nullptr;
    hipStream_t s = nullptr;
    hipError_t alloc(size_t bytes, hipStream_t st) {
        s = st;
        return hipMallocAsync(&p, bytes, st);
    }
    ~PoolBuf() {
        if (p) (void)hipFreeAsync(p, s);
    }
};

// The projector's two copies (z-pair, or A-pair for RS on 16-bit volumes:
// RS loops size them for the larger), allocated once for a loop's run and
// passed to each of its projections (which otherwise take them from the
// stream-ordered pool per call: two volume-sized allocations per projection,
// whose pool allocations stalled the host for up to seconds now and then
// between the loop's synchronisations).  The compat Siddon reads the volume
// itself and takes none; a Siddon volume beyond its 32-bit traversal index
// takes none either, so that the projector reports it as unsupported.
struct LoopCopies {
    PoolBuf qx, qy;
    ProjExtra extra;
    int alloc(const tomo_geom* g, const tomo_gd_params& pr, int storage, hipStream_t s) {
        const int nx = g->info.nx, ny = g->info.ny, nz = g->info.nz;
        if (pr.pair == TOMO_PAIR_SIDDONVI &&
            ((pr.flags & TOMO_FLAG_SIDDON_COMPAT) || zpair_elems_x(nx, ny, nz) >= (1ull << 32) ||
             zpair_elems_y(nx, ny, nz) >= (1ull << 32)))
            return TOMO_OK;
        const size_t es = storage == TOMO_F32 ? 4 : 2;
        const bool rs = pr.pair != TOMO_PAIR_SIDDONVI;
        TOMO_HIP(qx.alloc(rs ? proj_copy_bytes_x(nx, ny, nz, es) : zpair_elems_x(nx, ny, nz) * es, s));
        TOMO_HIP(qy.alloc(rs ? proj_copy_bytes_y(nx, ny, nz, es) : zpair_elems_y(nx, ny, nz) * es, s));
        extra.qx = qx.p;
        extra.qy = qy.p;
        return TOMO_OK;
    }
};

int project(const tomo_geom* g, const tomo_gd_params& pr, const float* vol, float* sino, hipStream_t s,
            const ProjExtra& x) {
    Buf in{vol, TOMO_F32, 1.0f};
    MutBuf out{sino, TOMO_F32, 1.0f};
    if (pr.pair == TOMO_PAIR_SIDDONVI)
        return launch_siddon(g, (pr.flags & TOMO_FLAG_SIDDON_COMPAT) ? 1 : 0, in, out, s, x);
    return launch_rs(g, (pr.flags & TOMO_FLAG_RS_FOV_CUBE) ? 1 : 0, in, out, s, x);
}

// Operator inputs of the mixed-precision loops (tomo_gd_params.storage
// TOMO_F16 / TOMO_BF16): each operator reads a 16-bit copy of its FP32
// input, made in `tmp` (room for the larger of a volume and a sinogram);
// FP16 copies are scaled by the power of two that puts max|x| in
// [2^13, 2^14), BF16 copies are unscaled.  Outputs stay FP32.  The copy is
// reused by the next operator of the same stream, after this one read it.
struct Mixed {
    int dtype = TOMO_F32;
    PoolBuf tmp;
    int init(int storage, size_t n, hipStream_t s) {
        dtype = storage;
        if (dtype != TOMO_F32) TOMO_HIP(tmp.alloc(n * 2, s));
        return TOMO_OK;
    }
    int stage(const float* x, size_t n, Buf* in, hipStream_t s) {
        if (dtype == TOMO_F32) {
            *in = Buf{x, TOMO_F32, 1.0f};
            return TOMO_OK;
        }
        float scale = 1.0f;
        if (dtype == TOMO_F16) {
            float m = 0.0f;
            const int rc = reduce_max_abs(x, n, &m, s);
            if (rc) return rc;
            if (m > 0.0f && std::isfinite(m)) {
                int e = 0;
                (void)std::frexp(m, &e);  // m = f 2^e, f in [0.5, 1)
                scale = std::ldexp(1.0f, std::min(100, std::max(-100, 14 - e)));
            }
        }
        *in = Buf{tmp.p, dtype, scale};
        return launch_convert(Buf{x, TOMO_F32, 1.0f}, MutBuf{tmp.p, dtype, scale}, n, s);
    }
};

int project(const tomo_geom* g, const tomo_gd_params& pr, Mixed& mx, const float* vol, float* sino, hipStream_t s,
            const ProjExtra& x) {
    Buf in{};
    int rc = mx.stage(vol, g->nvox(), &in, s);
    if (rc) return rc;
    const MutBuf out{sino, TOMO_F32, 1.0f};
    if (pr.pair == TOMO_PAIR_SIDDONVI)
        return launch_siddon(g, (pr.flags & TOMO_FLAG_SIDDON_COMPAT) ? 1 : 0, in, out, s, x);
    return launch_rs(g, (pr.flags & TOMO_FLAG_RS_FOV_CUBE) ? 1 : 0, in, out, s, x);
}

int backproject(const tomo_geom* g, Mixed& mx, const float* sino, float* vol, hipStream_t s) {
    Buf in{};
    int rc = mx.stage(sino, g->nsino(), &in, s);
    if (rc) return rc;
    return launch_vi(g, in, MutBuf{vol, TOMO_F32, 1.0f}, s);
}

}  // namespace

int gd_run(tomo_ctx* ctx, const tomo_geom* g, const tomo_gd_params& pr, const float* sino, const float* weight,
           float* vol, tomo_gd_iter_stat* stats, hipStream_t s) {
    const size_t nv = g->nvox(), ns = g->nsino();
    PoolBuf est, dj, t1, t2;
    TOMO_HIP(est.alloc(ns * sizeof(float), s));
    TOMO_HIP(dj.alloc(nv * sizeof(float), s));
    if (pr.lambda != 0.0) {
        TOMO_HIP(t1.alloc(nv * sizeof(float), s));
        TOMO_HIP(t2.alloc(nv * sizeof(float), s));
    }
    Mixed mx;
    int rc = mx.init(pr.storage, std::max(nv, ns), s);
    if (rc) return rc;
    LoopCopies cp;
    if ((rc = cp.alloc(g, pr, pr.storage, s))) return rc;
    float* e = (float*)est.p;
    float* d = (float*)dj.p;
    double alpha_min = pr.alpha_min_init;
    for (int k = 1; k <= pr.gradient_iteration_nb; ++k) {
        tomo_gd_iter_stat st{0, 0, 0, 0, 0, 0};
        if ((rc = project(g, pr, mx, vol, e, s, cp.extra))) return rc;
        if ((rc = launch_diff(sino, e, e, ns, s))) return rc;
        if ((rc = reduce_sq_norm(ctx, Buf{e, TOMO_F32, 1.0f}, ns, &st.residual_sq, s))) return rc;
        if (weight && (rc = launch_mul(e, weight, ns, s))) return rc;  // weighted form: r = W (y - P x)
        if ((rc = backproject(g, mx, e, d, s))) return rc;
        if ((rc = launch_scale(d, -2.0f, nv, s))) return rc;
        double norm_reg = 0.0;
        if (pr.lambda != 0.0) {
            float* a = (float*)t1.p;
            float* b = (float*)t2.p;
            if ((rc = launch_laplacian(vol, a, g->info.nx, g->info.ny, g->info.nz, s))) return rc;
            if ((rc = reduce_sq_norm(ctx, Buf{a, TOMO_F32, 1.0f}, nv, &st.j_reg, s))) return rc;
            if ((rc = launch_laplacian(a, b, g->info.nx, g->info.ny, g->info.nz, s))) return rc;
            const float w = (float)(2.0 * (double)(float)pr.lambda);
            if ((rc = launch_axpy(d, b, w, 0, nv, s))) return rc;
            if (k < pr.optimal_step_iteration_nb) {
                if ((rc = launch_laplacian(d, a, g->info.nx, g->info.ny, g->info.nz, s))) return rc;
                double sq = 0.0;
                if ((rc = reduce_sq_norm(ctx, Buf{a, TOMO_F32, 1.0f}, nv, &sq, s))) return rc;
                norm_reg = 2.0 * sq;
            }
        }
        double alpha;
        if (k < pr.optimal_step_iteration_nb) {
            if ((rc = project(g, pr, mx, d, e, s, cp.extra))) return rc;
            if ((rc = reduce_sq_norm(ctx, Buf{d, TOMO_F32, 1.0f}, nv, &st.norm_dj, s))) return rc;
            double pe = 0.0;
            if ((rc = reduce_sq_norm(ctx, Buf{e, TOMO_F32, 1.0f}, ns, &pe, s))) return rc;
            st.norm_dj_proj = 2.0 * pe + pr.lambda * norm_reg;
            alpha = -st.norm_dj / st.norm_dj_proj;
            alpha_min = std::fmax(alpha, alpha_min);
        } else {
            alpha = alpha_min;
        }
        st.alpha = alpha;
        if ((rc = launch_axpy(vol, d, (float)alpha, pr.positivity, nv, s))) return rc;
        if (stats) stats[k - 1] = st;
    }
    TOMO_HIP(hipStreamSynchronize(s));
    return TOMO_OK;
}

// Conjugate gradient on the criterion of gd_run (Fletcher-Reeves):
//   g = -2 B (y - P x) + 2 lambda L(L x)
//   d = g (k = 1),  d = beta d + g,  beta = |g|^2 / |g_prev|^2
//   alpha = -(g . d) / (2 |P d|^2 + lambda 2 |L d|^2);  x += alpha d
// The reference's doConjugateGradient (C/Iter3D_Algo_Gradient.cu:375-559)
// starts its direction at 0 and divides by that zero norm at iteration 2,
// so this loop defines the method instead of restating it.
int cg_run(tomo_ctx* ctx, const tomo_geom* g, const tomo_gd_params& pr, const float* sino, float* vol,
           tomo_gd_iter_stat* stats, hipStream_t s) {
    const size_t nv = g->nvox(), ns = g->nsino();
    PoolBuf est, gb, db, t1, t2;
    TOMO_HIP(est.alloc(ns * sizeof(float), s));
    TOMO_HIP(gb.alloc(nv * sizeof(float), s));
    TOMO_HIP(db.alloc(nv * sizeof(float), s));
    TOMO_HIP(t1.alloc(nv * sizeof(float), s));
    if (pr.lambda != 0.0) TOMO_HIP(t2.alloc(nv * sizeof(float), s));
    float* e = (float*)est.p;
    float* gr = (float*)gb.p;
    float* d = (float*)db.p;
    float* a = (float*)t1.p;
    float* b = (float*)t2.p;
    const int nx = g->info.nx, ny = g->info.ny, nz = g->info.nz;
    Mixed mx;
    int rc = mx.init(pr.storage, std::max(nv, ns), s);
    if (rc) return rc;
    LoopCopies cp;
    if ((rc = cp.alloc(g, pr, pr.storage, s))) return rc;
    double gg_prev = 0.0;
    for (int k = 1; k <= pr.gradient_iteration_nb; ++k) {
        tomo_gd_iter_stat st{0, 0, 0, 0, 0, 0};
        if ((rc = project(g, pr, mx, vol, e, s, cp.extra))) return rc;
        if ((rc = launch_diff(sino, e, e, ns, s))) return rc;
        if ((rc = reduce_sq_norm(ctx, Buf{e, TOMO_F32, 1.0f}, ns, &st.residual_sq, s))) return rc;
        if ((rc = backproject(g, mx, e, gr, s))) return rc;
        if ((rc = launch_scale(gr, -2.0f, nv, s))) return rc;
        if (pr.lambda != 0.0) {
            if ((rc = launch_laplacian(vol, a, nx, ny, nz, s))) return rc;
            if ((rc = reduce_sq_norm(ctx, Buf{a, TOMO_F32, 1.0f}, nv, &st.j_reg, s))) return rc;
            if ((rc = launch_laplacian(a, b, nx, ny, nz, s))) return rc;
            if ((rc = launch_axpy(gr, b, (float)(2.0 * (double)(float)pr.lambda), 0, nv, s))) return rc;
        }
        double gg = 0.0;
        if ((rc = reduce_sq_norm(ctx, Buf{gr, TOMO_F32, 1.0f}, nv, &gg, s))) return rc;
        if (k == 1) {
            TOMO_HIP(hipMemcpyAsync(d, gr, nv * sizeof(float), hipMemcpyDeviceToDevice, s));
        } else {
            st.beta = gg / gg_prev;
            if ((rc = launch_scale(d, (float)st.beta, nv, s))) return rc;
            if ((rc = launch_axpy(d, gr, 1.0f, 0, nv, s))) return rc;
        }
        gg_prev = gg;
        double gd = 0.0;
        if ((rc = reduce_dot(ctx, gr, d, nv, &gd, s))) return rc;
        if ((rc = project(g, pr, mx, d, e, s, cp.extra))) return rc;
        double pe = 0.0, reg = 0.0;
        if ((rc = reduce_sq_norm(ctx, Buf{e, TOMO_F32, 1.0f}, ns, &pe, s))) return rc;
        if (pr.lambda != 0.0) {
            if ((rc = launch_laplacian(d, a, nx, ny, nz, s))) return rc;
            double sq = 0.0;
            if ((rc = reduce_sq_norm(ctx, Buf{a, TOMO_F32, 1.0f}, nv, &sq, s))) return rc;
            reg = 2.0 * sq;
        }
        st.norm_dj = gg;
        st.norm_dj_proj = 2.0 * pe + pr.lambda * reg;
        st.alpha = -gd / st.norm_dj_proj;
        if ((rc = launch_axpy(vol, d, (float)st.alpha, pr.positivity, nv, s))) return rc;
        if (stats) stats[k - 1] = st;
    }
    TOMO_HIP(hipStreamSynchronize(s));
    return TOMO_OK;
}

// SIRT (defined by this library; the reference has none):
//   R = 1/(P 1), C = relaxation/(B 1), x0 = 0,
//   r = (y - P x) * R;  x = x + C * (B r)  (+ clamp at 0).
int sirt_run(tomo_ctx* ctx, const tomo_geom* g, const tomo_sirt_params& pr, const float* sino, float* vol,
             double* residual_sq, hipStream_t s) {
    const size_t nv = g->nvox(), ns = g->nsino();
    PoolBuf rw, cw, est, b;
    TOMO_HIP(rw.alloc(ns * sizeof(float), s));
    TOMO_HIP(est.alloc(ns * sizeof(float), s));
    TOMO_HIP(cw.alloc(nv * sizeof(float), s));
    TOMO_HIP(b.alloc(nv * sizeof(float), s));
    tomo_gd_params gp{};
    gp.pair = pr.pair;
    gp.flags = pr.flags;
    LoopCopies cp;
    if (const int arc = cp.alloc(g, gp, TOMO_F32, s)) return arc;
    float* R = (float*)rw.p;
    float* Cv = (float*)cw.p;
    float* e = (float*)est.p;
    float* bv = (float*)b.p;
    int rc;
    // weights from the operators applied to ones
    if ((rc = launch_fill(bv, 1.0f, nv, s))) return rc;
    if ((rc = project(g, gp, bv, R, s, cp.extra))) return rc;
    if ((rc = launch_recip_weight(R, 1.0f, R, ns, s))) return rc;
    if ((rc = launch_fill(e, 1.0f, ns, s))) return rc;
    if ((rc = launch_vi(g, Buf{e, TOMO_F32, 1.0f}, MutBuf{Cv, TOMO_F32, 1.0f}, s))) return rc;
    if ((rc = launch_recip_weight(Cv, (float)pr.relaxation, Cv, nv, s))) return rc;
    if ((rc = launch_fill(vol, 0.0f, nv, s))) return rc;
    for (int k = 0; k < pr.iterations; ++k) {
        if ((rc = project(g, gp, vol, e, s, cp.extra))) return rc;
        if ((rc = launch_diff(sino, e, e, ns, s))) return rc;
        if (residual_sq) {
            if ((rc = reduce_sq_norm(ctx, Buf{e, TOMO_F32, 1.0f}, ns, &residual_sq[k], s))) return rc;
        }
        if ((rc = launch_mul(e, R, ns, s))) return rc;
        if ((rc = launch_vi(g, Buf{e, TOMO_F32, 1.0f}, MutBuf{bv, TOMO_F32, 1.0f}, s))) return rc;
        if ((rc = launch_waxpy(vol, Cv, bv, pr.positivity, nv, s))) return rc;
    }
    TOMO_HIP(hipStreamSynchronize(s));
    return TOMO_OK;
}

}  // namespace tomo

extern "C" int tomo_sirt_run(tomo_ctx* ctx, const tomo_geom* g, const tomo_sirt_params* params,
                             const float* sinogram, float* volume, double* residual_sq, void* stream) {
    using namespace tomo;
    if (!ctx || !g || !params || !sinogram || !volume) return fail(TOMO_ERR_INVALID_ARG, "null argument");
    TOMO_HIP(hipSetDevice(ctx->device));
    if (params->pair != TOMO_PAIR_RSVI && params->pair != TOMO_PAIR_SIDDONVI)
        return fail(TOMO_ERR_INVALID_ARG, "unknown pair");
    if (params->iterations < 0) return fail(TOMO_ERR_INVALID_ARG, "negative iteration count");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (!(params->flags & TOMO_FLAG_HOST_PTR)) return sirt_run(ctx, g, *params, sinogram, volume, residual_sq, s);
    const size_t nv = g->nvox(), ns = g->nsino();
    DevBuf ds, dv;
    TOMO_HIP(hipMalloc(&ds.p, ns * sizeof(float)));
    TOMO_HIP(hipMalloc(&dv.p, nv * sizeof(float)));
    TOMO_HIP(hipMemcpyAsync(ds.p, sinogram, ns * sizeof(float), hipMemcpyHostToDevice, s));
    int rc = sirt_run(ctx, g, *params, (const float*)ds.p, (float*)dv.p, residual_sq, s);
    if (rc) return rc;
    TOMO_HIP(hipMemcpyAsync(volume, dv.p, nv * sizeof(float), hipMemcpyDeviceToHost, s));
    TOMO_HIP(hipStreamSynchronize(s));
    return TOMO_OK;
}

namespace {
// Shared entry of tomo_gd_run / tomo_cg_run: argument checks and host staging.
template <typename Run>
int run_loop(tomo_ctx* ctx, const tomo_geom* g, const tomo_gd_params* params, const float* sinogram, float* volume,
             tomo_gd_iter_stat* stats, void* stream, Run run) {
    using namespace tomo;
    if (!ctx || !g || !params || !sinogram || !volume) return fail(TOMO_ERR_INVALID_ARG, "null argument");
    TOMO_HIP(hipSetDevice(ctx->device));
    if (params->pair != TOMO_PAIR_RSVI && params->pair != TOMO_PAIR_SIDDONVI)
        return fail(TOMO_ERR_INVALID_ARG, "unknown pair");
    if (params->gradient_iteration_nb < 0) return fail(TOMO_ERR_INVALID_ARG, "negative iteration count");
    if (params->storage != TOMO_F32 && params->storage != TOMO_F16 && params->storage != TOMO_BF16)
        return fail(TOMO_ERR_INVALID_ARG, "storage must be TOMO_F32, TOMO_F16 or TOMO_BF16");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (!(params->flags & TOMO_FLAG_HOST_PTR)) return run(ctx, g, *params, sinogram, volume, stats, s);
    const size_t nv = g->nvox(), ns = g->nsino();
    DevBuf ds, dv;
    TOMO_HIP(hipMalloc(&ds.p, ns * sizeof(float)));
    TOMO_HIP(hipMalloc(&dv.p, nv * sizeof(float)));
    TOMO_HIP(hipMemcpyAsync(ds.p, sinogram, ns * sizeof(float), hipMemcpyHostToDevice, s));
    TOMO_HIP(hipMemcpyAsync(dv.p, volume, nv * sizeof(float), hipMemcpyHostToDevice, s));
    int rc = run(ctx, g, *params, (const float*)ds.p, (float*)dv.p, stats, s);
    if (rc) return rc;
    TOMO_HIP(hipMemcpyAsync(volume, dv.p, nv * sizeof(float), hipMemcpyDeviceToHost, s));
    TOMO_HIP(hipStreamSynchronize(s));
    return TOMO_OK;
}
}  // namespace

extern "C" int tomo_cg_run(tomo_ctx* ctx, const tomo_geom* g, const tomo_gd_params* params, const float* sinogram,
                           float* volume, tomo_gd_iter_stat* stats, void* stream) {
    return run_loop(ctx, g, params, sinogram, volume, stats, stream, tomo::cg_run);
}

extern "C" int tomo_gd_run(tomo_ctx* ctx, const tomo_geom* g, const tomo_gd_params* params, const float* sinogram,
                           float* volume, tomo_gd_iter_stat* stats, void* stream) {
    return run_loop(ctx, g, params, sinogram, volume, stats, stream,
                    [](tomo_ctx* c, const tomo_geom* gg, const tomo_gd_params& pr, const float* y, float* x,
                       tomo_gd_iter_stat* st, hipStream_t s) { return tomo::gd_run(c, gg, pr, y, nullptr, x, st, s); });
}

extern "C" int tomo_wgd_run(tomo_ctx* ctx, const tomo_geom* g, const tomo_gd_params* params, const float* sinogram,
                            const float* weight, float* volume, tomo_gd_iter_stat* stats, void* stream) {
    using namespace tomo;
    if (!weight) return fail(TOMO_ERR_INVALID_ARG, "null weight sinogram");
    DevBuf dw;
    const float* w = weight;
    if (ctx && g && params && (params->flags & TOMO_FLAG_HOST_PTR)) {
        TOMO_HIP(hipSetDevice(ctx->device));
        const size_t ns = g->nsino();
        TOMO_HIP(hipMalloc(&dw.p, ns * sizeof(float)));
        TOMO_HIP(hipMemcpy(dw.p, weight, ns * sizeof(float), hipMemcpyHostToDevice));
        w = (const float*)dw.p;
    }
    return run_loop(ctx, g, params, sinogram, volume, stats, stream,
                    [w](tomo_ctx* c, const tomo_geom* gg, const tomo_gd_params& pr, const float* y, float* x,
                        tomo_gd_iter_stat* st, hipStream_t s) { return gd_run(c, gg, pr, y, w, x, st, s); });
}
