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
// Volumes and sinograms stay on the device; only the FP64 norms return to the
// host for the step rule.  The SIRT loop (tomo_sirt_run) follows.
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

int project(const tomo_geom* g, const tomo_gd_params& pr, const float* vol, float* sino, hipStream_t s) {
    Buf in{vol, TOMO_F32, 1.0f};
    MutBuf out{sino, TOMO_F32, 1.0f};
    if (pr.pair == TOMO_PAIR_SIDDONVI) return launch_siddon(g, (pr.flags & TOMO_FLAG_SIDDON_COMPAT) ? 1 : 0, in, out, s);
    return launch_rs(g, (pr.flags & TOMO_FLAG_RS_FOV_CUBE) ? 1 : 0, in, out, s);
}

}  // namespace

int gd_run(tomo_ctx* ctx, const tomo_geom* g, const tomo_gd_params& pr, const float* sino, float* vol,
           tomo_gd_iter_stat* stats, hipStream_t s) {
    const size_t nv = g->nvox(), ns = g->nsino();
    DevBuf est, dj, t1, t2;
    TOMO_HIP(hipMalloc(&est.p, ns * sizeof(float)));
    TOMO_HIP(hipMalloc(&dj.p, nv * sizeof(float)));
    if (pr.lambda != 0.0) {
        TOMO_HIP(hipMalloc(&t1.p, nv * sizeof(float)));
        TOMO_HIP(hipMalloc(&t2.p, nv * sizeof(float)));
    }
    float* e = (float*)est.p;
    float* d = (float*)dj.p;
    double alpha_min = pr.alpha_min_init;
    int rc;
    for (int k = 1; k <= pr.gradient_iteration_nb; ++k) {
        tomo_gd_iter_stat st{0, 0, 0, 0};
        if ((rc = project(g, pr, vol, e, s))) return rc;
        if ((rc = launch_diff(sino, e, e, ns, s))) return rc;
        if ((rc = reduce_sq_norm(ctx, Buf{e, TOMO_F32, 1.0f}, ns, &st.residual_sq, s))) return rc;
        if ((rc = launch_vi(g, Buf{e, TOMO_F32, 1.0f}, MutBuf{d, TOMO_F32, 1.0f}, s))) return rc;
        if ((rc = launch_scale(d, -2.0f, nv, s))) return rc;
        double norm_reg = 0.0;
        if (pr.lambda != 0.0) {
            float* a = (float*)t1.p;
            float* b = (float*)t2.p;
            if ((rc = launch_laplacian(vol, a, g->info.nx, g->info.ny, g->info.nz, s))) return rc;
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
            if ((rc = project(g, pr, d, e, s))) return rc;
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

// SIRT (defined by this library; the reference has none):
//   R = 1/(P 1), C = relaxation/(B 1), x0 = 0,
//   r = (y - P x) * R;  x = x + C * (B r)  (+ clamp at 0).
int sirt_run(tomo_ctx* ctx, const tomo_geom* g, const tomo_sirt_params& pr, const float* sino, float* vol,
             double* residual_sq, hipStream_t s) {
    const size_t nv = g->nvox(), ns = g->nsino();
    DevBuf rw, cw, est, b;
    TOMO_HIP(hipMalloc(&rw.p, ns * sizeof(float)));
    TOMO_HIP(hipMalloc(&est.p, ns * sizeof(float)));
    TOMO_HIP(hipMalloc(&cw.p, nv * sizeof(float)));
    TOMO_HIP(hipMalloc(&b.p, nv * sizeof(float)));
    tomo_gd_params gp{};
    gp.pair = pr.pair;
    gp.flags = pr.flags;
    float* R = (float*)rw.p;
    float* Cv = (float*)cw.p;
    float* e = (float*)est.p;
    float* bv = (float*)b.p;
    int rc;
    // weights from the operators applied to ones
    if ((rc = launch_fill(bv, 1.0f, nv, s))) return rc;
    if ((rc = project(g, gp, bv, R, s))) return rc;
    if ((rc = launch_recip_weight(R, 1.0f, R, ns, s))) return rc;
    if ((rc = launch_fill(e, 1.0f, ns, s))) return rc;
    if ((rc = launch_vi(g, Buf{e, TOMO_F32, 1.0f}, MutBuf{Cv, TOMO_F32, 1.0f}, s))) return rc;
    if ((rc = launch_recip_weight(Cv, (float)pr.relaxation, Cv, nv, s))) return rc;
    if ((rc = launch_fill(vol, 0.0f, nv, s))) return rc;
    for (int k = 0; k < pr.iterations; ++k) {
        if ((rc = project(g, gp, vol, e, s))) return rc;
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

extern "C" int tomo_gd_run(tomo_ctx* ctx, const tomo_geom* g, const tomo_gd_params* params, const float* sinogram,
                           float* volume, tomo_gd_iter_stat* stats, void* stream) {
    using namespace tomo;
    if (!ctx || !g || !params || !sinogram || !volume) return fail(TOMO_ERR_INVALID_ARG, "null argument");
    TOMO_HIP(hipSetDevice(ctx->device));
    if (params->pair != TOMO_PAIR_RSVI && params->pair != TOMO_PAIR_SIDDONVI)
        return fail(TOMO_ERR_INVALID_ARG, "unknown pair");
    if (params->gradient_iteration_nb < 0) return fail(TOMO_ERR_INVALID_ARG, "negative iteration count");
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    if (!(params->flags & TOMO_FLAG_HOST_PTR)) return gd_run(ctx, g, *params, sinogram, volume, stats, s);
    // Host-resident data (the reference's mem_CPU storage): stage in, run,
    // copy the volume back.
    const size_t nv = g->nvox(), ns = g->nsino();
    DevBuf ds, dv;
    TOMO_HIP(hipMalloc(&ds.p, ns * sizeof(float)));
    TOMO_HIP(hipMalloc(&dv.p, nv * sizeof(float)));
    TOMO_HIP(hipMemcpyAsync(ds.p, sinogram, ns * sizeof(float), hipMemcpyHostToDevice, s));
    TOMO_HIP(hipMemcpyAsync(dv.p, volume, nv * sizeof(float), hipMemcpyHostToDevice, s));
    int rc = gd_run(ctx, g, *params, (const float*)ds.p, (float*)dv.p, stats, s);
    if (rc) return rc;
    TOMO_HIP(hipMemcpyAsync(volume, dv.p, nv * sizeof(float), hipMemcpyDeviceToHost, s));
    TOMO_HIP(hipStreamSynchronize(s));
    return TOMO_OK;
}
