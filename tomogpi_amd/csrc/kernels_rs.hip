// kernels_rs.hip - ray-driven regular-sampling (RS) projector, its exact
// transpose and its sample counter.
//
// Restates K/Projector/projection_ER_kernel_UM.cuh:25-153 (inscribed-cylinder
// FOV, the reference's device-resident kernel) and, with TOMO_FLAG_RS_FOV_CUBE,
// K/Projector/projection_ER_kernel.cuh:56-312 (circumscribed circle clipped
// to the square of voxel centres, the reference's default host-resident
// path).  One thread per detector cell (m,n) of view p:
//   source S = -FOD(cos,sin,0); cell D = (cos(FDD-FOD) - s sin,
//   sin(FDD-FOD) + s cos, t), s = (m-cu)du, t = (n-cv)dv; unit d = (D-S)/|D-S|
//   entry/exit = circle intersections (entry nearer the source); cube mode
//   clips to [0,N-1]^2; a ray whose entry z is outside [0,Nz) ([0,Nz-1] in
//   cube mode) gives 0; an exit beyond z=Nz-1 or z=0 is moved onto that plane;
//   L_r = floor(|exit-entry|) voxels;  g = delta * sum_{n<L_r} trilinear(f, entry + n d)
// The ray setup is FP64 (the reference parametrises by x, which is singular
// for rays parallel to y; the arc-length form here is well conditioned), the
// samples and the sum are FP32.  The output is overwritten.
#include "device_common.hpp"

namespace tomo {
namespace {

constexpr int BX = 64, BY = 4;

struct RsRay {
    float ex, ey, ez, dx, dy, dz;
    int lr;  // number of samples (0: ray dropped)
};

__device__ __forceinline__ RsRay rs_setup(const DevGeom& g, int cube, int p, int m, int n) {
    RsRay r{0, 0, 0, 0, 0, 0, 0};
    const double* vw = g.view + (size_t)p * kViewStride;
    const double ca = vw[0], sa = vw[1], dl = (double)g.delta;
    const double s = ((double)m - (double)g.cu) * (double)g.du;
    const double t = ((double)n - (double)g.cv) * (double)g.dv;
    const double rx = ca * g.fdd - s * sa, ry = sa * g.fdd + s * ca, rz = t;
    const double L = sqrt((rx * rx + ry * ry) + rz * rz);
    const double d0 = rx / L, d1 = ry / L, d2 = rz / L;
    const double Sx = -g.fod * ca / dl, Sy = -g.fod * sa / dl;
    const double half = 0.5 * (double)g.nx;
    const double R2 = cube ? 2.0 * (half * half) : half * half;
    const double a = d0 * d0 + d1 * d1;
    if (!(a > 0.0)) return r;
    const double b = Sx * d0 + Sy * d1;
    const double c = (Sx * Sx + Sy * Sy) - R2;
    const double disc = b * b - a * c;
    if (!(disc >= 0.0)) return r;
    const double sq = sqrt(disc);
    double l1 = (-b - sq) / a, l2 = (-b + sq) / a;
    const double cx = g.cx, cy = g.cy, cz = g.cz;
    if (cube) {
        // x then y, exactly as the oracle's loop over the two axes.
        {
            const double lo = -cx, hi = (double)(g.nx - 1) - cx;
            if (d0 != 0.0) {
                const double ta = (lo - Sx) / d0, tb = (hi - Sx) / d0;
                l1 = fmax(l1, fmin(ta, tb));
                l2 = fmin(l2, fmax(ta, tb));
            } else if (!(Sx >= lo && Sx <= hi)) {
                return r;
            }
        }
        {
            const double lo = -cy, hi = (double)(g.ny - 1) - cy;
            if (d1 != 0.0) {
                const double ta = (lo - Sy) / d1, tb = (hi - Sy) / d1;
                l1 = fmax(l1, fmin(ta, tb));
                l2 = fmin(l2, fmax(ta, tb));
            } else if (!(Sy >= lo && Sy <= hi)) {
                return r;
            }
        }
        if (!(l1 <= l2)) return r;
    }
    const double ze = l1 * d2 + cz;
    if (cube) {
        if (!(ze >= 0.0 && ze <= (double)(g.nz - 1))) return r;
    } else {
        if (!(ze >= 0.0 && ze < (double)g.nz)) return r;
    }
    const double zx = l2 * d2 + cz;
    if (zx > (double)(g.nz - 1)) l2 = ((double)(g.nz - 1) - cz) / d2;
    else if (zx < 0.0) l2 = -cz / d2;
    const double len = fabs(l2 - l1);
    if (!(len >= 1.0)) return r;
    r.ex = (float)((Sx + l1 * d0) + cx);
    r.ey = (float)((Sy + l1 * d1) + cy);
    r.ez = (float)(l1 * d2 + cz);
    r.dx = (float)d0;
    r.dy = (float)d1;
    r.dz = (float)d2;
    const double lf = floor(len);
    r.lr = lf > 1.0e9 ? 1000000000 : (int)lf;
    return r;
}

template <typename TV>
__device__ __forceinline__ float trilinear(const TV* __restrict__ f, int nx, int ny, int nz, float x, float y, float z,
                                           float inv) {
    const float xf = floorf(x), yf = floorf(y), zf = floorf(z);
    if (!(xf >= -1.0f && xf <= (float)nx && yf >= -1.0f && yf <= (float)ny && zf >= -1.0f && zf <= (float)nz))
        return 0.0f;
    const int ix = (int)xf, iy = (int)yf, iz = (int)zf;
    const float fx = x - xf, fy = y - yf, fz = z - zf;
    const bool x0 = ix >= 0 && ix < nx, x1 = ix + 1 >= 0 && ix + 1 < nx;
    const bool y0 = iy >= 0 && iy < ny, y1 = iy + 1 >= 0 && iy + 1 < ny;
    const bool z0 = iz >= 0 && iz < nz, z1 = iz + 1 >= 0 && iz + 1 < nz;
    const int64_t sy = nx, sz = (int64_t)nx * ny;
    const int64_t o = (int64_t)ix + sy * iy + sz * iz;
    const float t000 = tap(f, o, x0 && y0 && z0, inv), t100 = tap(f, o + 1, x1 && y0 && z0, inv);
    const float t010 = tap(f, o + sy, x0 && y1 && z0, inv), t110 = tap(f, o + sy + 1, x1 && y1 && z0, inv);
    const float t001 = tap(f, o + sz, x0 && y0 && z1, inv), t101 = tap(f, o + sz + 1, x1 && y0 && z1, inv);
    const float t011 = tap(f, o + sy + sz, x0 && y1 && z1, inv), t111 = tap(f, o + sy + sz + 1, x1 && y1 && z1, inv);
    const float c00 = fmaf(fx, t100 - t000, t000), c10 = fmaf(fx, t110 - t010, t010);
    const float c01 = fmaf(fx, t101 - t001, t001), c11 = fmaf(fx, t111 - t011, t011);
    const float c0 = fmaf(fy, c10 - c00, c00), c1 = fmaf(fy, c11 - c01, c01);
    return fmaf(fz, c1 - c0, c0);
}

template <typename TV, typename TS>
__global__ __launch_bounds__(BX* BY) void rs_kernel(DevGeom g, int cube, const TV* __restrict__ vol, float vol_inv,
                                                  TS* __restrict__ sino, float sino_scale) {
    const int m = blockIdx.x * BX + threadIdx.x;
    const int n = blockIdx.y * BY + threadIdx.y;
    const int p = blockIdx.z;
    if (m >= g.nu || n >= g.nv) return;
    const RsRay r = rs_setup(g, cube, p, m, n);
    float acc = 0.0f;
    for (int s = 0; s < r.lr; ++s) {
        const float sf = (float)s;
        acc += trilinear(vol, g.nx, g.ny, g.nz, fmaf(sf, r.dx, r.ex), fmaf(sf, r.dy, r.ey), fmaf(sf, r.dz, r.ez),
                         vol_inv);
    }
    const size_t o = (size_t)m + (size_t)g.nu * ((size_t)n + (size_t)g.nv * p);
    Storage<TS>::store(sino, o, r.lr > 0 ? acc * g.delta : 0.0f, sino_scale);
}

// RS^T: f[tap] += delta * w_tap * g(ray) (FP32 atomics; output zeroed first).
template <typename TS>
__global__ __launch_bounds__(BX* BY) void rs_adjoint_kernel(DevGeom g, int cube, const TS* __restrict__ sino,
                                                          float sino_inv, float* __restrict__ vol) {
    const int m = blockIdx.x * BX + threadIdx.x;
    const int n = blockIdx.y * BY + threadIdx.y;
    const int p = blockIdx.z;
    if (m >= g.nu || n >= g.nv) return;
    const size_t oi = (size_t)m + (size_t)g.nu * ((size_t)n + (size_t)g.nv * p);
    const float gv = Storage<TS>::load(sino, oi, sino_inv);
    if (gv == 0.0f) return;
    const RsRay r = rs_setup(g, cube, p, m, n);
    const float val = g.delta * gv;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const int64_t sy = nx, sz = (int64_t)nx * ny;
    for (int s = 0; s < r.lr; ++s) {
        const float sf = (float)s;
        const float x = fmaf(sf, r.dx, r.ex), y = fmaf(sf, r.dy, r.ey), z = fmaf(sf, r.dz, r.ez);
        const float xf = floorf(x), yf = floorf(y), zf = floorf(z);
        if (!(xf >= -1.0f && xf <= (float)nx && yf >= -1.0f && yf <= (float)ny && zf >= -1.0f && zf <= (float)nz))
            continue;
        const int ix = (int)xf, iy = (int)yf, iz = (int)zf;
        const float fx = x - xf, fy = y - yf, fz = z - zf;
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            const int ox = c & 1, oy = (c >> 1) & 1, oz = (c >> 2) & 1;
            const int X = ix + ox, Y = iy + oy, Z = iz + oz;
            if (X < 0 || X >= nx || Y < 0 || Y >= ny || Z < 0 || Z >= nz) continue;
            const float w = (ox ? fx : 1.0f - fx) * (oy ? fy : 1.0f - fy) * (oz ? fz : 1.0f - fz);
            unsafeAtomicAdd(vol + (X + sy * Y + sz * Z), w * val);
        }
    }
}

__global__ __launch_bounds__(BX* BY) void rs_count_kernel(DevGeom g, int cube, unsigned long long* total) {
    const int m = blockIdx.x * BX + threadIdx.x;
    const int n = blockIdx.y * BY + threadIdx.y;
    const int p = blockIdx.z;
    unsigned long long c = 0;
    if (m < g.nu && n < g.nv) c = (unsigned long long)rs_setup(g, cube, p, m, n).lr;
    // wave-level sum, one atomic per wave
    for (int off = 32; off > 0; off >>= 1) c += __shfl_down(c, off, 64);
    if ((threadIdx.x & 63) == 0) atomicAdd(total, c);
}

}  // namespace

int launch_rs(const tomo_geom* g, int cube, Buf vol, MutBuf sino, hipStream_t s) {
    const DevGeom& d = g->dev;
    dim3 grid(ceil_div(d.nu, BX), ceil_div(d.nv, BY), d.nphi);
    const float inv = 1.0f / vol.scale;
    TOMO_DISPATCH_STORAGE(vol.dtype, TV, {
        TOMO_DISPATCH_STORAGE(sino.dtype, TS, {
            hipLaunchKernelGGL((rs_kernel<TV, TS>), grid, dim3(BX, BY), 0, s, d, cube, (const TV*)vol.p, inv,
                               (TS*)sino.p, sino.scale);
        });
    });
    TOMO_HIP(hipGetLastError());
    return TOMO_OK;
}

int launch_rs_adjoint(const tomo_geom* g, int cube, Buf sino, MutBuf vol, hipStream_t s) {
    if (vol.dtype != TOMO_F32) return fail(TOMO_ERR_UNSUPPORTED, "adjoint output must be F32");
    const DevGeom& d = g->dev;
    int rc = launch_fill((float*)vol.p, 0.0f, g->nvox(), s);
    if (rc) return rc;
    dim3 grid(ceil_div(d.nu, BX), ceil_div(d.nv, BY), d.nphi);
    const float inv = 1.0f / sino.scale;
    TOMO_DISPATCH_STORAGE(sino.dtype, TS, {
        hipLaunchKernelGGL((rs_adjoint_kernel<TS>), grid, dim3(BX, BY), 0, s, d, cube, (const TS*)sino.p, inv,
                           (float*)vol.p);
    });
    TOMO_HIP(hipGetLastError());
    return TOMO_OK;
}

int count_rs_samples(tomo_ctx* ctx, const tomo_geom* g, int cube, uint64_t* out, hipStream_t s) {
    const DevGeom& d = g->dev;
    unsigned long long* dt = nullptr;
    TOMO_HIP(hipMallocAsync((void**)&dt, sizeof(unsigned long long), s));
    TOMO_HIP(hipMemsetAsync(dt, 0, sizeof(unsigned long long), s));
    dim3 grid(ceil_div(d.nu, BX), ceil_div(d.nv, BY), d.nphi);
    hipLaunchKernelGGL(rs_count_kernel, grid, dim3(BX, BY), 0, s, d, cube, dt);
    TOMO_HIP(hipGetLastError());
    unsigned long long h = 0;
    TOMO_HIP(hipMemcpyAsync(&h, dt, sizeof(h), hipMemcpyDeviceToHost, s));
    TOMO_HIP(hipStreamSynchronize(s));
    TOMO_HIP(hipFreeAsync(dt, s));
    *out = h;
    (void)ctx;
    return TOMO_OK;
}

}  // namespace tomo
