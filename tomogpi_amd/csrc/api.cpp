// api.cpp - the extern "C" entry points of libtomogpi_mi355x.so (contract
// in include/tomogpi.h).  Validates arguments, resolves streams and host
// staging, and forwards to the launchers of the kernel files.
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "config.hpp"
#include "batch.hpp"
#include "tomo_internal.hpp"

namespace tomo {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }
int fail(int status, const std::string& msg) {
    set_error(msg);
    return status;
}
int hip_fail(hipError_t e, const char* what) {
    set_error(std::string(what) + ": " + hipGetErrorString(e));
    return e == hipErrorOutOfMemory ? TOMO_ERR_OUT_OF_MEMORY : TOMO_ERR_HIP;
}

namespace {

int bind(tomo_ctx* ctx) {
    if (!ctx) return fail(TOMO_ERR_INVALID_ARG, "null context");
    TOMO_HIP(hipSetDevice(ctx->device));
    return TOMO_OK;
}

hipStream_t stream_of(tomo_ctx* ctx, void* stream) {
    (void)ctx;
    return reinterpret_cast<hipStream_t>(stream);
}

bool storage_ok(int dt) { return dt == TOMO_F32 || dt == TOMO_F16 || dt == TOMO_BF16; }

int check_scale(int dt, float scale) {
    if (!storage_ok(dt)) return fail(TOMO_ERR_INVALID_ARG, "element type must be TOMO_F32, TOMO_F16 or TOMO_BF16");
    if (!(scale > 0.0f) || !std::isfinite(scale)) return fail(TOMO_ERR_INVALID_ARG, "scale must be finite and > 0");
    if (dt == TOMO_F32 && scale != 1.0f) return fail(TOMO_ERR_INVALID_ARG, "F32 storage takes scale 1");
    return TOMO_OK;
}

// Host staging for TOMO_FLAG_HOST_PTR: device copies of the input and the
// output, the input copied in, the output copied back after the call.
struct Staging {
    void* d_in = nullptr;
    void* d_out = nullptr;
    ~Staging() {
        if (d_in) (void)hipFree(d_in);
        if (d_out) (void)hipFree(d_out);
    }
};

using Launch = int (*)(const tomo_geom*, int, Buf, MutBuf, hipStream_t);

int run_operator(tomo_ctx* ctx, const tomo_geom* g, int flags, const void* in, int in_dt, float in_scale,
                 size_t in_count, void* out, int out_dt, float out_scale, size_t out_count, void* stream,
                 int (*fn)(const tomo_geom*, int, Buf, MutBuf, hipStream_t), int mode) {
    int rc = bind(ctx);
    if (rc) return rc;
    if (!g || g->ctx != ctx) return fail(TOMO_ERR_INVALID_ARG, "geometry does not belong to this context");
    if (!in || !out) return fail(TOMO_ERR_INVALID_ARG, "null data pointer");
    if ((rc = check_scale(in_dt, in_scale))) return rc;
    if ((rc = check_scale(out_dt, out_scale))) return rc;
    hipStream_t s = stream_of(ctx, stream);
    if (!(flags & TOMO_FLAG_HOST_PTR)) return fn(g, mode, Buf{in, in_dt, in_scale}, MutBuf{out, out_dt, out_scale}, s);
    Staging st;
    const size_t ib = in_count * dtype_size(in_dt), ob = out_count * dtype_size(out_dt);
    TOMO_HIP(hipMalloc(&st.d_in, ib));
    TOMO_HIP(hipMalloc(&st.d_out, ob));
    TOMO_HIP(hipMemcpyAsync(st.d_in, in, ib, hipMemcpyHostToDevice, s));
    rc = fn(g, mode, Buf{st.d_in, in_dt, in_scale}, MutBuf{st.d_out, out_dt, out_scale}, s);
    if (rc) return rc;
    TOMO_HIP(hipMemcpyAsync(out, st.d_out, ob, hipMemcpyDeviceToHost, s));
    TOMO_HIP(hipStreamSynchronize(s));
    return TOMO_OK;
}

// Projector modes: bit 0 = cube FOV (RS) / compat (Siddon), bit 1 = index
// reads, bit 2 = LDS-staged taps.
constexpr int kModeIndexReads = 2, kModeStaged = 4;
int proj_mode(int flags, int fov_or_compat) {
    return ((flags & fov_or_compat) ? 1 : 0) | ((flags & TOMO_FLAG_INDEX_READS) ? kModeIndexReads : 0) |
           ((flags & TOMO_FLAG_STAGED) ? kModeStaged : 0);
}
ProjExtra extra_of(int mode) {
    ProjExtra x;
    x.index_reads = (mode & kModeIndexReads) != 0;
    x.staged = (mode & kModeStaged) != 0;
    return x;
}
int do_rs(const tomo_geom* g, int mode, Buf in, MutBuf out, hipStream_t s) {
    return launch_rs(g, mode & 1, in, out, s, extra_of(mode));
}
int do_siddon(const tomo_geom* g, int mode, Buf in, MutBuf out, hipStream_t s) {
    return launch_siddon(g, mode & 1, in, out, s, extra_of(mode));
}
int do_vi(const tomo_geom* g, int fdk, Buf in, MutBuf out, hipStream_t s) { return launch_vi(g, in, out, s, fdk); }
int do_rs_adj(const tomo_geom* g, int cube, Buf in, MutBuf out, hipStream_t s) {
    return launch_rs_adjoint(g, cube, in, out, s);
}
int do_siddon_adj(const tomo_geom* g, int compat, Buf in, MutBuf out, hipStream_t s) {
    if (compat) return fail(TOMO_ERR_UNSUPPORTED, "no transpose of the compat Siddon (its quirks are not linear-exact)");
    return launch_siddon_adjoint(g, in, out, s);
}
int do_vi_adj(const tomo_geom* g, int, Buf in, MutBuf out, hipStream_t s) { return launch_vi_adjoint(g, in, out, s); }

int upload_geometry(tomo_geom* g) {
    TOMO_HIP(hipMalloc((void**)&g->d_view, sizeof(double) * g->view.size()));
    TOMO_HIP(hipMemcpy(g->d_view, g->view.data(), sizeof(double) * g->view.size(), hipMemcpyHostToDevice));
    TOMO_HIP(hipMalloc((void**)&g->d_view_f, sizeof(float) * g->view_f.size()));
    TOMO_HIP(hipMemcpy(g->d_view_f, g->view_f.data(), sizeof(float) * g->view_f.size(), hipMemcpyHostToDevice));
    g->dev.view = g->d_view;
    g->dev.view_f = g->d_view_f;
    // the VI box table depends on the geometry only: made once, here
    TOMO_HIP(hipMalloc(&g->d_vi_boxes, vi_box_count(g->dev) * sizeof(ViBox)));
    const int rc = launch_vi_boxes(g->dev, g->d_vi_boxes, nullptr);
    if (rc) return rc;
    TOMO_HIP(hipStreamSynchronize(nullptr));
    return TOMO_OK;
}

}  // namespace
}  // namespace tomo

using namespace tomo;

extern "C" {

int tomo_abi_version(void) { return TOMOGPI_ABI_VERSION; }
const char* tomo_last_error(void) { return g_last_error.c_str(); }

const char* tomo_status_string(int status) {
    switch (status) {
        case TOMO_OK: return "ok";
        case TOMO_ERR_INVALID_ARG: return "invalid argument";
        case TOMO_ERR_HIP: return "HIP runtime error";
        case TOMO_ERR_UNSUPPORTED: return "unsupported";
        case TOMO_ERR_OUT_OF_MEMORY: return "out of device memory";
        case TOMO_ERR_IO: return "I/O error";
        case TOMO_ERR_CONFIG: return "configuration error";
        case TOMO_ERR_BUSY: return "busy";
        default: return "unknown status";
    }
}

int tomo_ctx_create(int device, tomo_ctx** out) {
    if (!out) return fail(TOMO_ERR_INVALID_ARG, "null output");
    int n = 0;
    TOMO_HIP(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return fail(TOMO_ERR_INVALID_ARG, "no such device " + std::to_string(device));
    TOMO_HIP(hipSetDevice(device));
    int rc = noalloc_scratch_init(device);
    if (rc) return rc;
    if ((rc = batch_grids_init(device))) return rc;
    // Keep up to a quarter of the device's memory in its stream-ordered pool
    // between operator calls.  The projectors' per-call copies and the loops'
    // work buffers then come back from the pool; with the pool's default
    // threshold (0) every synchronisation handed them back to the driver and
    // the next allocation mapped them again, which left the device idle for
    // up to seconds per reconstruction, by an amount that varied between runs
    // of the same one.  tomo_ctx_destroy trims what is unused.
    hipMemPool_t pool = nullptr;
    size_t free_b = 0, total_b = 0;
    TOMO_HIP(hipDeviceGetDefaultMemPool(&pool, device));
    TOMO_HIP(hipMemGetInfo(&free_b, &total_b));
    uint64_t keep = (uint64_t)(total_b / 4);
    TOMO_HIP(hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep));
    auto* c = new tomo_ctx();
    c->device = device;
    *out = c;
    return TOMO_OK;
}

int tomo_ctx_destroy(tomo_ctx* ctx) {
    if (!ctx) return TOMO_OK;
    hipMemPool_t pool = nullptr;
    if (hipSetDevice(ctx->device) == hipSuccess && hipDeviceGetDefaultMemPool(&pool, ctx->device) == hipSuccess)
        (void)hipMemPoolTrimTo(pool, 0);
    delete ctx;
    return TOMO_OK;
}

int tomo_ctx_mem_info(tomo_ctx* ctx, size_t* free_bytes, size_t* total_bytes) {
    int rc = bind(ctx);
    if (rc) return rc;
    size_t f = 0, t = 0;
    TOMO_HIP(hipMemGetInfo(&f, &t));
    if (free_bytes) *free_bytes = f;
    if (total_bytes) *total_bytes = t;
    return TOMO_OK;
}

int tomo_alloc(tomo_ctx* ctx, size_t bytes, void** dptr) {
    int rc = bind(ctx);
    if (rc) return rc;
    if (!dptr) return fail(TOMO_ERR_INVALID_ARG, "null output");
    TOMO_HIP(hipMalloc(dptr, bytes));
    return TOMO_OK;
}
int tomo_free(tomo_ctx* ctx, void* dptr) {
    int rc = bind(ctx);
    if (rc) return rc;
    TOMO_HIP(hipFree(dptr));
    return TOMO_OK;
}
int tomo_copy_to_device(tomo_ctx* ctx, void* dst, const void* src, size_t bytes) {
    int rc = bind(ctx);
    if (rc) return rc;
    TOMO_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return TOMO_OK;
}
int tomo_copy_to_host(tomo_ctx* ctx, void* dst, const void* src, size_t bytes) {
    int rc = bind(ctx);
    if (rc) return rc;
    TOMO_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return TOMO_OK;
}
int tomo_synchronize(tomo_ctx* ctx, void* stream) {
    int rc = bind(ctx);
    if (rc) return rc;
    TOMO_HIP(hipStreamSynchronize(reinterpret_cast<hipStream_t>(stream)));
    return TOMO_OK;
}

int tomo_geom_create(tomo_ctx* ctx, const tomo_geom_desc* desc, tomo_geom** out) {
    int rc = bind(ctx);
    if (rc) return rc;
    if (!desc || !out) return fail(TOMO_ERR_INVALID_ARG, "null argument");
    std::unique_ptr<tomo_geom> g(new tomo_geom());
    g->ctx = ctx;
    if ((rc = derive_geometry(*desc, g.get()))) return rc;
    if ((rc = upload_geometry(g.get()))) {
        if (g->d_view) (void)hipFree(g->d_view);
        if (g->d_view_f) (void)hipFree(g->d_view_f);
        if (g->d_vi_boxes) (void)hipFree(g->d_vi_boxes);
        return rc;
    }
    *out = g.release();
    return TOMO_OK;
}

int tomo_geom_destroy(tomo_geom* g) {
    if (!g) return TOMO_OK;
    if (g->ctx) (void)hipSetDevice(g->ctx->device);
    if (g->d_view) (void)hipFree(g->d_view);
    if (g->d_view_f) (void)hipFree(g->d_view_f);
    if (g->d_vi_boxes) (void)hipFree(g->d_vi_boxes);
    delete g;
    return TOMO_OK;
}

int tomo_geom_get_info(const tomo_geom* g, tomo_geom_info* out) {
    if (!g || !out) return fail(TOMO_ERR_INVALID_ARG, "null argument");
    *out = g->info;
    return TOMO_OK;
}

int tomo_geom_get_phi(const tomo_geom* g, double* phi_out) {
    if (!g || !phi_out) return fail(TOMO_ERR_INVALID_ARG, "null argument");
    std::memcpy(phi_out, g->phi.data(), sizeof(double) * g->phi.size());
    return TOMO_OK;
}

int tomo_project(tomo_ctx* ctx, const tomo_geom* g, int projector, int flags, const void* volume, int volume_dtype,
                 float volume_scale, void* sinogram, int sinogram_dtype, float sinogram_scale, void* stream) {
    if (!g) return fail(TOMO_ERR_INVALID_ARG, "null geometry");
    if (projector == TOMO_PROJ_RS)
        return run_operator(ctx, g, flags, volume, volume_dtype, volume_scale, g->nvox(), sinogram, sinogram_dtype,
                            sinogram_scale, g->nsino(), stream, do_rs, proj_mode(flags, TOMO_FLAG_RS_FOV_CUBE));
    if (projector == TOMO_PROJ_SIDDON)
        return run_operator(ctx, g, flags, volume, volume_dtype, volume_scale, g->nvox(), sinogram, sinogram_dtype,
                            sinogram_scale, g->nsino(), stream, do_siddon, proj_mode(flags, TOMO_FLAG_SIDDON_COMPAT));
    return fail(TOMO_ERR_INVALID_ARG, "unknown projector");
}

int tomo_backproject(tomo_ctx* ctx, const tomo_geom* g, int backprojector, int flags, const void* sinogram,
                     int sinogram_dtype, float sinogram_scale, void* volume, int volume_dtype, float volume_scale,
                     void* stream) {
    if (!g) return fail(TOMO_ERR_INVALID_ARG, "null geometry");
    if (backprojector != TOMO_BP_VI) return fail(TOMO_ERR_INVALID_ARG, "unknown backprojector");
    return run_operator(ctx, g, flags, sinogram, sinogram_dtype, sinogram_scale, g->nsino(), volume, volume_dtype,
                        volume_scale, g->nvox(), stream, do_vi, (flags & TOMO_FLAG_FDK) ? 1 : 0);
}

int tomo_project_adjoint(tomo_ctx* ctx, const tomo_geom* g, int projector, int flags, const void* sinogram,
                         int sinogram_dtype, float sinogram_scale, void* volume, int volume_dtype, float volume_scale,
                         void* stream) {
    if (!g) return fail(TOMO_ERR_INVALID_ARG, "null geometry");
    if (projector == TOMO_PROJ_RS)
        return run_operator(ctx, g, flags, sinogram, sinogram_dtype, sinogram_scale, g->nsino(), volume, volume_dtype,
                            volume_scale, g->nvox(), stream, do_rs_adj, (flags & TOMO_FLAG_RS_FOV_CUBE) ? 1 : 0);
    if (projector == TOMO_PROJ_SIDDON)
        return run_operator(ctx, g, flags, sinogram, sinogram_dtype, sinogram_scale, g->nsino(), volume, volume_dtype,
                            volume_scale, g->nvox(), stream, do_siddon_adj,
                            (flags & TOMO_FLAG_SIDDON_COMPAT) ? 1 : 0);
    return fail(TOMO_ERR_INVALID_ARG, "unknown projector");
}

int tomo_backproject_adjoint(tomo_ctx* ctx, const tomo_geom* g, int backprojector, int flags, const void* volume,
                             int volume_dtype, float volume_scale, void* sinogram, int sinogram_dtype,
                             float sinogram_scale, void* stream) {
    if (!g) return fail(TOMO_ERR_INVALID_ARG, "null geometry");
    if (backprojector != TOMO_BP_VI) return fail(TOMO_ERR_INVALID_ARG, "unknown backprojector");
    return run_operator(ctx, g, flags, volume, volume_dtype, volume_scale, g->nvox(), sinogram, sinogram_dtype,
                        sinogram_scale, g->nsino(), stream, do_vi_adj, 0);
}

int tomo_project_work(tomo_ctx* ctx, const tomo_geom* g, int projector, int flags, uint64_t* units, void* stream) {
    int rc = bind(ctx);
    if (rc) return rc;
    if (!g || !units) return fail(TOMO_ERR_INVALID_ARG, "null argument");
    hipStream_t s = stream_of(ctx, stream);
    if (projector == TOMO_PROJ_RS) return count_rs_samples(ctx, g, (flags & TOMO_FLAG_RS_FOV_CUBE) ? 1 : 0, units, s);
    if (projector == TOMO_PROJ_SIDDON) {
        if (flags & TOMO_FLAG_SIDDON_COMPAT) return fail(TOMO_ERR_UNSUPPORTED, "no work counter for compat Siddon");
        return count_siddon_segments(ctx, g, units, s);
    }
    return fail(TOMO_ERR_INVALID_ARG, "unknown projector");
}

int tomo_diff(tomo_ctx* ctx, const float* a, const float* b, float* out, size_t n, void* stream) {
    int rc = bind(ctx);
    if (rc) return rc;
    if (!a || !b || !out) return fail(TOMO_ERR_INVALID_ARG, "null pointer");
    return launch_diff(a, b, out, n, stream_of(ctx, stream));
}
int tomo_scale(tomo_ctx* ctx, float* x, float s, size_t n, void* stream) {
    int rc = bind(ctx);
    if (rc) return rc;
    if (!x) return fail(TOMO_ERR_INVALID_ARG, "null pointer");
    return launch_scale(x, s, n, stream_of(ctx, stream));
}
int tomo_axpy(tomo_ctx* ctx, float* y, const float* x, double lambda, int positive, size_t n, void* stream) {
    int rc = bind(ctx);
    if (rc) return rc;
    if (!x || !y) return fail(TOMO_ERR_INVALID_ARG, "null pointer");
    return launch_axpy(y, x, (float)lambda, positive, n, stream_of(ctx, stream));
}
int tomo_sq_norm(tomo_ctx* ctx, const void* x, int dtype, float scale, size_t n, double* out, void* stream) {
    int rc = bind(ctx);
    if (rc) return rc;
    if (!x || !out) return fail(TOMO_ERR_INVALID_ARG, "null pointer");
    if ((rc = check_scale(dtype, scale))) return rc;
    return reduce_sq_norm(ctx, Buf{x, dtype, scale}, n, out, stream_of(ctx, stream));
}
int tomo_sq_norm_noalloc(tomo_ctx* ctx, const void* x, int dtype, float scale, size_t n, double* out,
                         void* stream) {
    int rc = bind(ctx);
    if (rc) return rc;
    if (!x || !out) return fail(TOMO_ERR_INVALID_ARG, "null pointer");
    if ((rc = check_scale(dtype, scale))) return rc;
    return reduce_sq_norm_noalloc(ctx->device, Buf{x, dtype, scale}, n, out, stream_of(ctx, stream));
}
int tomo_dot(tomo_ctx* ctx, const float* a, const float* b, size_t n, double* out, void* stream) {
    int rc = bind(ctx);
    if (rc) return rc;
    if (!a || !b || !out) return fail(TOMO_ERR_INVALID_ARG, "null pointer");
    return reduce_dot(ctx, a, b, n, out, stream_of(ctx, stream));
}
int tomo_laplacian(tomo_ctx* ctx, const tomo_geom* g, const float* x, float* y, void* stream) {
    int rc = bind(ctx);
    if (rc) return rc;
    if (!g || !x || !y) return fail(TOMO_ERR_INVALID_ARG, "null pointer");
    if (x == y) return fail(TOMO_ERR_INVALID_ARG, "laplacian cannot run in place");
    return launch_laplacian(x, y, g->info.nx, g->info.ny, g->info.nz, stream_of(ctx, stream));
}
int tomo_convert(tomo_ctx* ctx, const void* in, int in_dtype, float in_scale, void* out, int out_dtype,
                 float out_scale, size_t n, void* stream) {
    int rc = bind(ctx);
    if (rc) return rc;
    if (!in || !out) return fail(TOMO_ERR_INVALID_ARG, "null pointer");
    if ((rc = check_scale(in_dtype, in_scale)) || (rc = check_scale(out_dtype, out_scale))) return rc;
    return launch_convert(Buf{in, in_dtype, in_scale}, MutBuf{out, out_dtype, out_scale}, n, stream_of(ctx, stream));
}

// ---- phantoms
namespace {
// Shepp-Logan geometry (Data3D_0256/phantom3d.m:186-200) with densities
// 2,-.98,-.02,-.02,.01 x6.  Columns: A a b c x0 y0 z0 phi theta psi.
const double kSheppLogan[10][10] = {
    {2.00, .6900, .920, .810, 0.00, 0.0000, 0.00, 0, 0, 0},
    {-0.98, .6624, .874, .780, 0.00, -0.0184, 0.00, 0, 0, 0},
    {-0.02, .1100, .310, .220, 0.22, 0.0000, 0.00, -18, 0, 10},
    {-0.02, .1600, .410, .280, -0.22, 0.0000, 0.00, 18, 0, 10},
    {0.01, .2100, .250, .410, 0.00, 0.3500, -0.15, 0, 0, 0},
    {0.01, .0460, .046, .050, 0.00, 0.1000, 0.25, 0, 0, 0},
    {0.01, .0460, .046, .050, 0.00, -0.1000, 0.25, 0, 0, 0},
    {0.01, .0460, .023, .050, -0.08, -0.6050, 0.00, 0, 0, 0},
    {0.01, .0230, .023, .020, 0.00, -0.6060, 0.00, 0, 0, 0},
    {0.01, .0230, .046, .020, 0.06, -0.6050, 0.00, 0, 0, 0},
};

// z-x-z Euler rotation of phantom3d.m:96-103.
void euler(double phd, double thd, double psd, double* m) {
    const double ph = phd * M_PI / 180.0, th = thd * M_PI / 180.0, ps = psd * M_PI / 180.0;
    const double cph = std::cos(ph), sph = std::sin(ph), cth = std::cos(th), sth = std::sin(th);
    const double cps = std::cos(ps), sps = std::sin(ps);
    m[0] = cps * cph - cth * sph * sps;  m[1] = cps * sph + cth * cph * sps;  m[2] = sps * sth;
    m[3] = -sps * cph - cth * sph * cps; m[4] = -sps * sph + cth * cph * cps; m[5] = cps * sth;
    m[6] = sth * sph;                    m[7] = -sth * cph;                   m[8] = cth;
}

uint64_t splitmix64(uint64_t* s) {
    uint64_t z = (*s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
double u01(uint64_t* s) { return (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0); }

int run_ellipsoids(tomo_ctx* ctx, const tomo_geom* g, int mode, const std::vector<double>& tab, int count, float* vol,
                   hipStream_t s) {
    double* d_tab = nullptr;
    TOMO_HIP(hipMalloc((void**)&d_tab, sizeof(double) * (tab.empty() ? 1 : tab.size())));
    if (!tab.empty()) {
        hipError_t e = hipMemcpy(d_tab, tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            (void)hipFree(d_tab);
            return hip_fail(e, "hipMemcpy(ellipsoid table)");
        }
    }
    int rc = launch_ellipsoids(g, mode, d_tab, count, vol, s);
    hipError_t e = hipStreamSynchronize(s);
    (void)hipFree(d_tab);
    if (rc) return rc;
    if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize(ellipsoids)");
    (void)ctx;
    return TOMO_OK;
}
}  // namespace

int tomo_phantom_shepp_logan(tomo_ctx* ctx, const tomo_geom* g, float* volume, void* stream) {
    int rc = bind(ctx);
    if (rc) return rc;
    if (!g || !volume) return fail(TOMO_ERR_INVALID_ARG, "null pointer");
    std::vector<double> tab(16 * 10);
    for (int e = 0; e < 10; ++e) {
        double* t = &tab[16 * e];
        euler(kSheppLogan[e][7], kSheppLogan[e][8], kSheppLogan[e][9], t);
        t[9] = kSheppLogan[e][4];
        t[10] = kSheppLogan[e][5];
        t[11] = kSheppLogan[e][6];
        t[12] = kSheppLogan[e][1] * kSheppLogan[e][1];
        t[13] = kSheppLogan[e][2] * kSheppLogan[e][2];
        t[14] = kSheppLogan[e][3] * kSheppLogan[e][3];
        t[15] = kSheppLogan[e][0];
    }
    return run_ellipsoids(ctx, g, 0, tab, 10, volume, stream_of(ctx, stream));
}

// Per ellipsoid, 10 draws of splitmix64 uniforms: centre (inner 80% of each
// axis, voxels from the centre), semi-axes in [amin,amax], three uniforms of
// a Shoemake random unit quaternion, density.
int tomo_phantom_ellipsoids(tomo_ctx* ctx, const tomo_geom* g, int count, uint64_t seed, double amin, double amax,
                            float* volume, void* stream) {
    int rc = bind(ctx);
    if (rc) return rc;
    if (!g || !volume || count < 0 || !(amin > 0.0) || !(amax >= amin))
        return fail(TOMO_ERR_INVALID_ARG, "bad phantom arguments");
    std::vector<double> tab(16 * (size_t)count);
    uint64_t s = seed;
    const int n[3] = {g->info.nx, g->info.ny, g->info.nz};
    for (int e = 0; e < count; ++e) {
        double* t = &tab[16 * (size_t)e];
        for (int a = 0; a < 3; ++a) t[9 + a] = (u01(&s) - 0.5) * 0.8 * (double)n[a];
        for (int a = 0; a < 3; ++a) {
            const double semi = amin + (amax - amin) * u01(&s);
            t[12 + a] = 1.0 / (semi * semi);
        }
        const double u1 = u01(&s), u2 = u01(&s), u3 = u01(&s);
        const double r1 = std::sqrt(1.0 - u1), r2 = std::sqrt(u1);
        const double qx = r1 * std::sin(2.0 * M_PI * u2), qy = r1 * std::cos(2.0 * M_PI * u2);
        const double qz = r2 * std::sin(2.0 * M_PI * u3), qw = r2 * std::cos(2.0 * M_PI * u3);
        t[0] = 1.0 - 2.0 * (qy * qy + qz * qz); t[1] = 2.0 * (qx * qy - qz * qw);       t[2] = 2.0 * (qx * qz + qy * qw);
        t[3] = 2.0 * (qx * qy + qz * qw);       t[4] = 1.0 - 2.0 * (qx * qx + qz * qz); t[5] = 2.0 * (qy * qz - qx * qw);
        t[6] = 2.0 * (qx * qz - qy * qw);       t[7] = 2.0 * (qy * qz + qx * qw);       t[8] = 1.0 - 2.0 * (qx * qx + qy * qy);
        t[15] = u01(&s);
    }
    return run_ellipsoids(ctx, g, 1, tab, count, volume, stream_of(ctx, stream));
}

}  // extern "C"
