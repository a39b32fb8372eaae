// tomo_internal.hpp - internal types of libtomogpi_mi355x.so shared by the
// host API (api.cpp, loop.cpp, config.cpp) and the kernel files (*.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstddef>
#include <string>
#include <vector>

#include "tomogpi.h"

namespace tomo {

// ---------------------------------------------------------------- errors
// Thread-local message of the last failure (tomo_last_error()).
void set_error(const std::string& msg);
int fail(int status, const std::string& msg);
int hip_fail(hipError_t e, const char* what);

#define TOMO_HIP(call)                                         \
    do {                                                       \
        hipError_t e_ = (call);                                \
        if (e_ != hipSuccess) return ::tomo::hip_fail(e_, #call); \
    } while (0)

// -------------------------------------------------------------- geometry
// Scalars every kernel needs, passed by value.  Types follow the reference
// (float detector / voxel sizes and centres, double distances).
struct DevGeom {
    int32_t nx, ny, nz, nu, nv, nphi;
    double fod, fdd;
    float du, dv, cu, cv;
    float delta, cx, cy, cz;
    // Per-view table in device memory, kViewStride doubles per view:
    //   [0] cos phi  [1] sin phi
    //   [2] delta cos  [3] delta sin  [4] -(delta sin) FDD/du
    //   [5] (delta cos) FDD/du  [6] delta FDD/dv  [7] unused
    const double* view;
    // Reference float tables alpha = (float)cos, beta = (float)sin, 2 per view.
    const float* view_f;
};
constexpr int kViewStride = 8;

}  // namespace tomo

struct tomo_ctx {
    int device = 0;
};

struct tomo_geom {
    tomo_ctx* ctx = nullptr;
    tomo_geom_desc desc{};
    tomo_geom_info info{};
    std::vector<double> phi, view;  // host copies
    std::vector<float> view_f;
    double* d_view = nullptr;
    float* d_view_f = nullptr;
    void* d_vi_boxes = nullptr;  // VI box table (tomo_geom_create), or null: made per call
    tomo::DevGeom dev{};
    size_t nvox() const { return (size_t)info.nx * info.ny * info.nz; }
    size_t nsino() const { return (size_t)info.nu * info.nv * info.nphi; }
};

namespace tomo {

// Element counts of the z-pair copies of a volume for x slabs (qx) and y
// slabs (qy): nA * (nz+1) * (nB+2) pairs.
__host__ __device__ inline size_t zpair_elems_x(int nx, int ny, int nz) { return (size_t)nx * (nz + 1) * (ny + 2) * 2; }
__host__ __device__ inline size_t zpair_elems_y(int nx, int ny, int nz) { return (size_t)ny * (nz + 1) * (nx + 2) * 2; }
// Element counts of the A-pair copies (16-bit volumes, RS): (nA+1) * (nz+1)
// * (nB+2) groups of four elements (zpair.hpp).
__host__ __device__ inline size_t apair_elems_x(int nx, int ny, int nz) { return (size_t)(nx + 1) * (nz + 1) * (ny + 2) * 4; }
__host__ __device__ inline size_t apair_elems_y(int nx, int ny, int nz) { return (size_t)(ny + 1) * (nz + 1) * (nx + 2) * 4; }
// Whether RS reads the A-pair copies through 32-bit buffer offsets: each copy
// below 2^32 - 16 bytes (the offset of a read that returns zeros lies past
// it) and its plane stride below 2^24 bytes (24-bit multiplies); otherwise
// 16-bit RS reads the z-pair copies.
__host__ __device__ inline bool apair_buffer_path(int nx, int ny, int nz) {
    const uint64_t plane_x = (uint64_t)(ny + 2) * 8u * (uint64_t)(nz + 1);
    const uint64_t plane_y = (uint64_t)(nx + 2) * 8u * (uint64_t)(nz + 1);
    return apair_elems_x(nx, ny, nz) * 2u < 0xfffffff0ull && apair_elems_y(nx, ny, nz) * 2u < 0xfffffff0ull &&
           plane_x < (1ull << 24) && plane_y < (1ull << 24);
}
// Bytes of the largest copy any projector form reads for slabs along x (y)
// from a volume of es-byte elements: what a caller that hands copies to every
// projection (ProjExtra) allocates.
inline size_t proj_copy_bytes_x(int nx, int ny, int nz, size_t es) {
    const size_t z = zpair_elems_x(nx, ny, nz) * es;
    return es == 2 ? std::max(z, apair_elems_x(nx, ny, nz) * es) : z;
}
inline size_t proj_copy_bytes_y(int nx, int ny, int nz, size_t es) {
    const size_t z = zpair_elems_y(nx, ny, nz) * es;
    return es == 2 ? std::max(z, apair_elems_y(nx, ny, nz) * es) : z;
}

// Derives the geometry exactly as the reference constructors do.
int derive_geometry(const tomo_geom_desc& d, tomo_geom* g);

// ------------------------------------------------------------- launchers
// All launchers are stream-ordered and check their launch shape against the
// buffers they are given.  Element type codes are TOMO_F32/F16/BF16.
struct Buf {
    const void* p;
    int dtype;
    float scale;
};
struct MutBuf {
    void* p;
    int dtype;
    float scale;
};

// fdk: FDK-weighted form (back_fdk) instead of delta * sum.
int launch_vi(const tomo_geom* g, Buf sino, MutBuf vol, hipStream_t s, int fdk = 0);
int launch_vi_adjoint(const tomo_geom* g, Buf vol, MutBuf sino, hipStream_t s);
// Projector options beyond the FOV / compat mode.  index_reads: read the
// z-pair copies through 64-bit element indices even where 32-bit buffer
// offsets apply (the form every copy of 4 GiB or more takes).  qx, qy: the
// projector's copies in memory the caller provides (proj_copy_bytes_x / _y
// bytes each, for the volume's element size); null: the launcher allocates
// them from the stream-ordered pool for the call.
struct ProjExtra {
    bool index_reads = false;
    bool staged = false;  // RS, Siddon (FP32 volumes): LDS-staged taps instead of per-lane gathers
    void* qx = nullptr;
    void* qy = nullptr;
};
int launch_rs(const tomo_geom* g, int cube, Buf vol, MutBuf sino, hipStream_t s, const ProjExtra& x = ProjExtra{});
int launch_rs_adjoint(const tomo_geom* g, int cube, Buf sino, MutBuf vol, hipStream_t s);
int launch_siddon(const tomo_geom* g, int compat, Buf vol, MutBuf sino, hipStream_t s,
                  const ProjExtra& x = ProjExtra{});
int launch_siddon_adjoint(const tomo_geom* g, Buf sino, MutBuf vol, hipStream_t s);
// Work counters (roofline accounting): sum of L_r (RS) and number of
// positive chords (Siddon).  Synchronous.
int count_rs_samples(tomo_ctx* ctx, const tomo_geom* g, int cube, uint64_t* out, hipStream_t s);
int count_siddon_segments(tomo_ctx* ctx, const tomo_geom* g, uint64_t* out, hipStream_t s);

int launch_fill(float* x, float v, size_t n, hipStream_t s);
int launch_diff(const float* a, const float* b, float* out, size_t n, hipStream_t s);
int launch_scale(float* x, float v, size_t n, hipStream_t s);
int launch_axpy(float* y, const float* x, float a, int positive, size_t n, hipStream_t s);
int launch_laplacian(const float* x, float* y, int nx, int ny, int nz, hipStream_t s);
// SIRT element-wise steps: w = s > 0 ? a / s : 0; x *= w; x = x + c * b (+clamp).
int launch_recip_weight(const float* s_in, float a, float* w, size_t n, hipStream_t s);
int launch_mul(float* x, const float* w, size_t n, hipStream_t s);
int launch_waxpy(float* x, const float* c, const float* b, int positive, size_t n, hipStream_t s);
int reduce_sq_norm(tomo_ctx* ctx, Buf x, size_t n, double* out, hipStream_t s);
int reduce_dot(tomo_ctx* ctx, const float* a, const float* b, size_t n, double* out, hipStream_t s);
// Allocation-free squared norm: the partial sums live in a device array of
// the library, one per GPU (noalloc_scratch_init looks its address up once per
// device, with that device current; tomo_ctx_create calls it).  Synchronous.
int noalloc_scratch_init(int device);
int reduce_sq_norm_noalloc(int device, Buf x, size_t n, double* out, hipStream_t s);
// max |x| over n FP32 elements.  Synchronous.
int reduce_max_abs(const float* x, size_t n, float* out, hipStream_t s);
int launch_convert(Buf in, MutBuf out, size_t n, hipStream_t s);

// Ellipsoid phantoms.  mode 0: Shepp-Logan form (normalised coordinates,
// rotation then offset, division by squared semi-axes); mode 1: random
// ellipsoids (voxel offsets, offset then rotation, inverse squared axes).
// table: 16 doubles per ellipsoid (9 rotation, 3 centre, 3 axis terms, 1
// density).
int launch_ellipsoids(const tomo_geom* g, int mode, const double* d_table, int count, float* vol,
                      hipStream_t s);

}  // namespace tomo
