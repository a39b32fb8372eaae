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
#include <atomic>
#include <type_traits>

#include "zpair.hpp"
#include "batch.hpp"
#include "stage_box.hpp"

namespace tomo {
namespace {

constexpr int BX = kProjTileU, BY = kProjTileV;

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

// The trilinear blend of a sample's eight taps, given as the (B, z) quads of
// its two A planes: u on plane floor(A), w on floor(A) + 1, each ordered
// (b, z), (b, z+1), (b+1, z), (b+1, z+1).  Interpolated along x, then y, then
// z with explicit fmaf.
template <int A>
__device__ __forceinline__ float trilinear_taps(float u0, float u1, float u2, float u3, float w0, float w1, float w2,
                                                float w3, float fx, float fy, float fz) {
    float t000, t100, t010, t110, t001, t101, t011, t111;
    if (A == 0) {
        t000 = u0; t001 = u1; t010 = u2; t011 = u3;
        t100 = w0; t101 = w1; t110 = w2; t111 = w3;
    } else {
        t000 = u0; t001 = u1; t100 = u2; t101 = u3;
        t010 = w0; t011 = w1; t110 = w2; t111 = w3;
    }
    const float c00 = fmaf(fx, t100 - t000, t000), c10 = fmaf(fx, t110 - t010, t010);
    const float c01 = fmaf(fx, t101 - t001, t001), c11 = fmaf(fx, t111 - t011, t011);
    const float c0 = fmaf(fy, c10 - c00, c00), c1 = fmaf(fy, c11 - c01, c01);
    return fmaf(fz, c1 - c0, c0);
}

// Trilinear interpolation of f at (x, y, z) (voxel units), taps outside the
// grid 0, interpolated along x, then y, then z with explicit fmaf.  The eight
// taps come from two z-pair quads of the slab axis A: the (B, z) block of the
// planes A = floor(A) and floor(A) + 1.  A = 0 reads qx (B = y), A = 1 reads
// qy (B = x).  Outside the copy's corner range (B or z corner = n) the block
// is entirely outside the grid and the quad is 0.
template <int A, typename TV>
__device__ __forceinline__ float trilinear_quads(const TV* __restrict__ q, int nx, int ny, int nz, float x, float y,
                                                 float z, float inv) {
    const float xf = floorf(x), yf = floorf(y), zf = floorf(z);
    if (!(xf >= -1.0f && xf <= (float)nx && yf >= -1.0f && yf <= (float)ny && zf >= -1.0f && zf <= (float)nz))
        return 0.0f;
    const int ix = (int)xf, iy = (int)yf, iz = (int)zf;
    const float fx = x - xf, fy = y - yf, fz = z - zf;
    const int ia = A == 0 ? ix : iy, ib = A == 0 ? iy : ix;
    const int nA = A == 0 ? nx : ny, nB = A == 0 ? ny : nx;
    const bool blk = ib < nB && iz < nz;  // ib, iz >= -1 from the range test
    const bool p0 = blk && ia >= 0 && ia < nA, p1 = blk && ia + 1 < nA;
    const int bq = min(ib, nB - 1) + 1, zq = min(iz, nz - 1) + 1;
    const float4 u = Storage<TV>::load4(q, zpair_index(min(max(ia, 0), nA - 1), zq, bq, nz, nB), inv);
    const float4 w = Storage<TV>::load4(q, zpair_index(min(ia + 1, nA - 1), zq, bq, nz, nB), inv);
    // u, w = (b, z), (b, z+1), (b+1, z), (b+1, z+1) on planes ia, ia+1
    const float u0 = p0 ? u.x : 0.0f, u1 = p0 ? u.y : 0.0f, u2 = p0 ? u.z : 0.0f, u3 = p0 ? u.w : 0.0f;
    const float w0 = p1 ? w.x : 0.0f, w1 = p1 ? w.y : 0.0f, w2 = p1 ? w.z : 0.0f, w3 = p1 ? w.w : 0.0f;
    return trilinear_taps<A>(u0, u1, u2, u3, w0, w1, w2, w3, fx, fy, fz);
}

// Buffer-path sampler, for copies below 4 GiB whose A-plane stride is below
// 2^24 bytes: 32-bit byte offsets from 24-bit multiply-adds, loads
// issued unconditionally.  A buffer load at an offset past the copy's end
// returns 0 and never faults, so an absent plane (outside the grid along A),
// a sample whose cell has no voxel in the grid, or an inactive one (act
// false) reads from kOutside: its taps are 0 and the interpolation gives +0,
// the value the index forms return for it.  Corners are clamped before the integer conversion, which
// keeps every offset well defined.  Same taps and arithmetic as above.
constexpr uint32_t kOutside = 0xfffffff0u;

struct CopyBuf {
    __amdgpu_buffer_rsrc_t r;
    uint32_t sa, sz;  // bytes per A plane and per z row
};

template <int A, typename TV>
__device__ __forceinline__ float trilinear_quads_buf(const CopyBuf& q, int nx, int ny, int nz, float x, float y,
                                                     float z, bool act) {
    constexpr uint32_t es = sizeof(TV);  // a z-pair element is 2 es bytes
    const float xf = floorf(x), yf = floorf(y), zf = floorf(z);
    const float xc = __builtin_amdgcn_fmed3f(xf, -1.0f, (float)nx);
    const float yc = __builtin_amdgcn_fmed3f(yf, -1.0f, (float)ny);
    const float zc = __builtin_amdgcn_fmed3f(zf, -1.0f, (float)nz);
    const bool in = act & (xc == xf) & (yc == yf) & (zc == zf);
    const int ix = (int)xc, iy = (int)yc, iz = (int)zc;
    const float fx = x - xf, fy = y - yf, fz = z - zf;
    const int ia = A == 0 ? ix : iy, ib = A == 0 ? iy : ix;
    const int nA = A == 0 ? nx : ny, nB = A == 0 ? ny : nx;
    const bool blk = in & (ib < nB) & (iz < nz);
    const bool p0 = blk & (ia >= 0) & (ia < nA), p1 = blk & (ia + 1 < nA);
    const uint32_t bq = (uint32_t)(min(ib, nB - 1) + 1), zq = (uint32_t)(min(iz, nz - 1) + 1);
    const uint32_t inner = __umul24(zq, q.sz) + bq * (2u * es);
    const uint32_t o0 = p0 ? __umul24((uint32_t)max(ia, 0), q.sa) + inner : kOutside;
    const uint32_t o1 = p1 ? __umul24((uint32_t)(ia + 1), q.sa) + inner : kOutside;
    // u, w = (b, z), (b, z+1), (b+1, z), (b+1, z+1) on planes ia, ia+1
    float u0, u1, u2, u3, w0, w1, w2, w3;
    if constexpr (es == 4) {
        const auto u = __builtin_amdgcn_raw_buffer_load_b128(q.r, (int)o0, 0, 0);
        const auto w = __builtin_amdgcn_raw_buffer_load_b128(q.r, (int)o1, 0, 0);
        u0 = __uint_as_float(u[0]); u1 = __uint_as_float(u[1]); u2 = __uint_as_float(u[2]); u3 = __uint_as_float(u[3]);
        w0 = __uint_as_float(w[0]); w1 = __uint_as_float(w[1]); w2 = __uint_as_float(w[2]); w3 = __uint_as_float(w[3]);
    } else {
        using U = Unpack16<TV>;
        const auto u = __builtin_amdgcn_raw_buffer_load_b64(q.r, (int)o0, 0, 0);
        const auto w = __builtin_amdgcn_raw_buffer_load_b64(q.r, (int)o1, 0, 0);
        u0 = U::lo(u[0]); u1 = U::hi(u[0]); u2 = U::lo(u[1]); u3 = U::hi(u[1]);
        w0 = U::lo(w[0]); w1 = U::hi(w[0]); w2 = U::lo(w[1]); w3 = U::hi(w[1]);
    }
    return trilinear_taps<A>(u0, u1, u2, u3, w0, w1, w2, w3, fx, fy, fz);
}

// A-pair sampler through 64-bit element indices (16-bit volumes whose A-pair
// copies are 4 GiB or more): one 16-byte load per sample, as below; a sample
// outside the grid or whose block is returns 0, as trilinear_quads does.
template <typename TV>
__device__ __forceinline__ void apair_unpack(const uint32_t (&d)[4], float& u0, float& u1, float& u2, float& u3,
                                             float& w0, float& w1, float& w2, float& w3) {
    using U = Unpack16<TV>;
    // d[0], d[1] = (a, b, z), (a, b, z+1), (a+1, b, z), (a+1, b, z+1); d[2], d[3]: the same at b + 1
    u0 = U::lo(d[0]); u1 = U::hi(d[0]); w0 = U::lo(d[1]); w1 = U::hi(d[1]);
    u2 = U::lo(d[2]); u3 = U::hi(d[2]); w2 = U::lo(d[3]); w3 = U::hi(d[3]);
    if constexpr (std::is_same<TV, __half>::value) {
        // Keep the FP16 taps as converted FP32 registers: otherwise the
        // compiler folds their conversions into v_fma_mix_f32 operands, the
        // same numbers at more issue time (RS at 512^3 / 512 views: 81.4 ms
        // folded, 76.3 ms not).  An empty statement: it emits no instruction.
        asm("" : "+v"(u0), "+v"(u1), "+v"(u2), "+v"(u3), "+v"(w0), "+v"(w1), "+v"(w2), "+v"(w3));
    }
}

template <int A, typename TV>
__device__ __forceinline__ float trilinear_apair(const TV* __restrict__ q, int nx, int ny, int nz, float x, float y,
                                                 float z) {
    static_assert(sizeof(TV) == 2, "A-pair copies hold 16-bit volumes");
    const float xf = floorf(x), yf = floorf(y), zf = floorf(z);
    if (!(xf >= -1.0f && xf <= (float)nx && yf >= -1.0f && yf <= (float)ny && zf >= -1.0f && zf <= (float)nz))
        return 0.0f;
    const int ix = (int)xf, iy = (int)yf, iz = (int)zf;
    const float fx = x - xf, fy = y - yf, fz = z - zf;
    const int ia = A == 0 ? ix : iy, ib = A == 0 ? iy : ix;
    const int nA = A == 0 ? nx : ny, nB = A == 0 ? ny : nx;
    if (!(ib < nB && iz < nz && ia < nA)) return 0.0f;  // every tap 0: the blend gives +0
    const unsigned char* p =
        (const unsigned char*)__builtin_assume_aligned(q + apair_index(ia + 1, iz + 1, ib + 1, nz, nB), 8);
    uint32_t d[4];
    __builtin_memcpy(d, p, 16);
    float u0, u1, u2, u3, w0, w1, w2, w3;
    apair_unpack<TV>(d, u0, u1, u2, u3, w0, w1, w2, w3);
    return trilinear_taps<A>(u0, u1, u2, u3, w0, w1, w2, w3, fx, fy, fz);
}

// A-pair sampler (16-bit volumes, copies below 4 GiB): the eight taps are two
// neighbouring groups of the A-pair copy (zpair.hpp), one 16-byte load where
// the z-pair copy takes two.  Its planes -1 and nA are zeros of the copy, so
// only a block outside the grid (B or z corner = n, A corner = nA) or an
// inactive sample reads from kOutside.  Same taps and arithmetic as above.
struct ApairShape {
    uint64_t bytes_x, bytes_y;
    uint32_t sa_x, sz_x, sa_y, sz_y;
    bool buf;
};
__host__ __device__ __forceinline__ ApairShape rs_apair_shape(int nx, int ny, int nz) {
    ApairShape c;
    c.sz_x = (uint32_t)(ny + 2) * 8u;
    c.sz_y = (uint32_t)(nx + 2) * 8u;
    c.bytes_x = apair_elems_x(nx, ny, nz) * 2u;
    c.bytes_y = apair_elems_y(nx, ny, nz) * 2u;
    c.sa_x = c.sz_x * (uint32_t)(nz + 1);
    c.sa_y = c.sz_y * (uint32_t)(nz + 1);
    c.buf = apair_buffer_path(nx, ny, nz);
    return c;
}

template <int A, typename TV>
__device__ __forceinline__ float trilinear_apair_buf(const CopyBuf& q, int nx, int ny, int nz, float x, float y,
                                                     float z, bool act) {
    static_assert(sizeof(TV) == 2, "A-pair copies hold 16-bit volumes");
    const float xf = floorf(x), yf = floorf(y), zf = floorf(z);
    const float xc = __builtin_amdgcn_fmed3f(xf, -1.0f, (float)nx);
    const float yc = __builtin_amdgcn_fmed3f(yf, -1.0f, (float)ny);
    const float zc = __builtin_amdgcn_fmed3f(zf, -1.0f, (float)nz);
    const bool in = act & (xc == xf) & (yc == yf) & (zc == zf);
    const int ix = (int)xc, iy = (int)yc, iz = (int)zc;
    const float fx = x - xf, fy = y - yf, fz = z - zf;
    const int ia = A == 0 ? ix : iy, ib = A == 0 ? iy : ix;
    const int nA = A == 0 ? nx : ny, nB = A == 0 ? ny : nx;
    const bool blk = in & (ib < nB) & (iz < nz) & (ia < nA);  // ia >= -1 from the clamp
    const uint32_t bq = (uint32_t)(min(ib, nB - 1) + 1), zq = (uint32_t)(min(iz, nz - 1) + 1);
    const uint32_t ap = (uint32_t)(min(ia, nA - 1) + 1);
    // kOutside unless blk, as a mask: a select of the offset would let the
    // compiler move its arithmetic into a branch, which splits the unrolled
    // samples and serialises their loads
    const uint32_t keep = 0u - (uint32_t)blk;
    const uint32_t o = ((__umul24(ap, q.sa) + __umul24(zq, q.sz) + bq * 8u) & keep) | (kOutside & ~keep);
    const auto v = __builtin_amdgcn_raw_buffer_load_b128(q.r, (int)o, 0, 0);
    const uint32_t d[4] = {v[0], v[1], v[2], v[3]};
    float u0, u1, u2, u3, w0, w1, w2, w3;
    apair_unpack<TV>(d, u0, u1, u2, u3, w0, w1, w2, w3);
    return trilinear_taps<A>(u0, u1, u2, u3, w0, w1, w2, w3, fx, fy, fz);
}

__device__ __forceinline__ int wave_min(int v) {
    for (int off = 32; off > 0; off >>= 1) v = min(v, __shfl_xor(v, off, 64));
    return v;
}
__device__ __forceinline__ int wave_max(int v) {
    for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off, 64));
    return v;
}

__device__ __forceinline__ float wave_minf(float v) {
    for (int off = 32; off > 0; off >>= 1) v = fminf(v, __shfl_xor(v, off, 64));
    return v;
}

// Samples of one lane, taken one per iteration k of the wave's loop: sample
// s = k - k0 for k0 <= k < k0 + lr.  Same positions (fma(s, d, entry)) and
// the same order of the sum as the per-ray loop.
// q is the z-pair copy of A (FP32 or 16-bit, as stored); BUF reads it
// through qb (see the buffer-path samplers); AP (16-bit volumes) reads the
// A-pair copy of A instead, through qb or 64-bit indices.  The buffer path
// has no branch: RS_UNROLL samples per iteration, their loads in flight
// together, an inactive sample (s outside [0, lr)) adding +0 in its place,
// which leaves the sum unchanged (it starts at +0 and is never -0).
#ifndef RS_UNROLL
#define RS_UNROLL 4
#endif
template <int A, typename TV, bool BUF, bool AP>
__device__ __forceinline__ float rs_samples(const RsRay& r, const void* __restrict__ q, const CopyBuf& qb, int nx,
                                            int ny, int nz, int k0, int kend) {
    float acc = 0.0f;
    if constexpr (BUF) {
        for (int k = 0; k < kend; k += RS_UNROLL) {
            float v[RS_UNROLL];
#pragma unroll
            for (int j = 0; j < RS_UNROLL; ++j) {
                const int s = k + j - k0;
                const float sf = (float)s;
                const float x = fmaf(sf, r.dx, r.ex), y = fmaf(sf, r.dy, r.ey), z = fmaf(sf, r.dz, r.ez);
                const bool act = (s >= 0) & (s < r.lr);
                if constexpr (AP) v[j] = trilinear_apair_buf<A, TV>(qb, nx, ny, nz, x, y, z, act);
                else v[j] = trilinear_quads_buf<A, TV>(qb, nx, ny, nz, x, y, z, act);
            }
#pragma unroll
            for (int j = 0; j < RS_UNROLL; ++j) acc += v[j];
        }
    } else {
        for (int k = 0; k < kend; ++k) {
            const int s = k - k0;
            if (s >= 0 && s < r.lr) {
                const float sf = (float)s;
                const float x = fmaf(sf, r.dx, r.ex), y = fmaf(sf, r.dy, r.ey), z = fmaf(sf, r.dz, r.ez);
                if constexpr (AP) acc += trilinear_apair<A, TV>((const TV*)q, nx, ny, nz, x, y, z);
                else acc += trilinear_quads<A>((const TV*)q, nx, ny, nz, x, y, z, 1.0f);
            }
        }
    }
    return acc;
}

// Byte strides of the copy read by the buffer path, and whether it applies.
struct CopyShape {
    uint64_t bytes_x, bytes_y;
    uint32_t sa_x, sz_x, sa_y, sz_y;
    bool buf;
};
__host__ __device__ __forceinline__ CopyShape rs_copy_shape(int nx, int ny, int nz, uint32_t es) {
    CopyShape c;
    c.sz_x = (uint32_t)(ny + 2) * 2u * es;
    c.sz_y = (uint32_t)(nx + 2) * 2u * es;
    c.bytes_x = zpair_elems_x(nx, ny, nz) * es;
    c.bytes_y = zpair_elems_y(nx, ny, nz) * es;
    const uint64_t sax = (uint64_t)c.sz_x * (uint64_t)(nz + 1), say = (uint64_t)c.sz_y * (uint64_t)(nz + 1);
    c.sa_x = (uint32_t)sax;
    c.sa_y = (uint32_t)say;
    c.buf = c.bytes_x < kOutside && c.bytes_y < kOutside && sax < (1ull << 24) && say < (1ull << 24);
    return c;
}

// Forward RS projector.  Samples, taps and the order of the sum are those of
// the per-ray loop (sample s at fma(s, d, entry), trilinear interpolation,
// acc += in s order), so results are unchanged; what changes is the
// schedule.  When every ray of the wave has the same dominant in-plane axis
// A (|d_A| >= the other two) and direction, each lane starts its samples k0
// iterations after the wave's first entry along A, k0 = its entry's distance
// behind that one in samples (|d_A| per sample): at iteration k every lane
// samples near the same A plane, one sample per lane per iteration.  Its 64
// rays (a 32 x 2 block of detector cells) then read neighbouring quads of the
// z-pair copy of A, two loads per sample (16 bytes each for FP32 storage,
// 8 for 16-bit), or (AP, 16-bit volumes) the A-pair copy of A, one 16-byte
// load per sample.  Other waves sample from k = 0 on the y-slab copy.  Grid: x = column group, y = view, z = row group
// (every view of a row band before the next band).
template <typename TV, typename TS, bool BUF, bool AP = false>
__device__ __forceinline__ void rs_tile(const DevGeom& g, int cube, const void* __restrict__ qx,
                                        const void* __restrict__ qy, float vol_inv, TS* __restrict__ sino,
                                        float sino_scale, int bx, int p, int bz) {
    const int m = bx * BX + threadIdx.x;
    const int n = bz * BY + threadIdx.y;
    const bool inside = m < g.nu && n < g.nv;
    RsRay r{0, 0, 0, 0, 0, 0, 0};
    if (inside) r = rs_setup(g, cube, p, m, n);
    const bool ok = r.lr > 0;
    int code = -1;
    if (ok) {
        const float ax = fabsf(r.dx), ay = fabsf(r.dy), az = fabsf(r.dz);
        if (ax > 0.0f && ax >= ay && ax >= az) code = r.dx > 0.0f ? 0 : 1;
        else if (ay > 0.0f && ay >= ax && ay >= az) code = r.dy > 0.0f ? 2 : 3;
    }
    const unsigned long long okmask = __ballot(ok);
    int mode = -1;
    if (okmask != 0ull) {
#pragma unroll
        for (int c = 0; c < 4; ++c)
            if (__ballot(ok && code == c) == okmask) mode = c;
    }
    int k0 = 0;
    if (mode >= 0) {
        const bool xs = mode < 2;
        const float pa = (mode & 1) ? -(xs ? r.ex : r.ey) : (xs ? r.ex : r.ey);  // entry along the travel direction
        const float da = fabsf(xs ? r.dx : r.dy);                                  // >= 1/sqrt(3) for admitted rays
        const float pmin = wave_minf(ok ? pa : 3.0e38f);
        // Entries are inside the volume's bounding box, so k0 <= sqrt(3) N.
        if (ok) k0 = (int)((pa - pmin) / da + 0.5f);
    }
    const int kend = wave_max(ok ? k0 + r.lr : 0);
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    float acc;
    // 16-bit volumes are read unscaled; the sum is linear in them, so the
    // storage scale applies once to it (FP32 storage has no scale).
    CopyBuf qb{};
    if constexpr (BUF) {
        const bool xs = mode >= 0 && mode < 2;
        if constexpr (AP) {
            const ApairShape cs = rs_apair_shape(nx, ny, nz);
            qb.r = buffer_rsrc(xs ? qx : qy, (uint32_t)(xs ? cs.bytes_x : cs.bytes_y));
            qb.sa = xs ? cs.sa_x : cs.sa_y;
            qb.sz = xs ? cs.sz_x : cs.sz_y;
        } else {
            const CopyShape cs = rs_copy_shape(nx, ny, nz, sizeof(TV));
            qb.r = buffer_rsrc(xs ? qx : qy, (uint32_t)(xs ? cs.bytes_x : cs.bytes_y));
            qb.sa = xs ? cs.sa_x : cs.sa_y;
            qb.sz = xs ? cs.sz_x : cs.sz_y;
        }
    }
    if (mode >= 0 && mode < 2) acc = rs_samples<0, TV, BUF, AP>(r, qx, qb, nx, ny, nz, k0, kend);
    else acc = rs_samples<1, TV, BUF, AP>(r, qy, qb, nx, ny, nz, k0, kend);
    if (sizeof(TV) == 2) acc *= vol_inv;
    if (inside) {
        const size_t o = (size_t)m + (size_t)g.nu * ((size_t)n + (size_t)g.nv * p);
        Storage<TS>::store(sino, o, r.lr > 0 ? acc * g.delta : 0.0f, sino_scale);
    }
}

template <typename TV, typename TS, bool BUF, bool AP>
__global__ __launch_bounds__(BX* BY) void rs_kernel(DevGeom g, int cube, const void* __restrict__ qx,
                                                  const void* __restrict__ qy, float vol_inv, TS* __restrict__ sino,
                                                  float sino_scale) {
    rs_tile<TV, TS, BUF, AP>(g, cube, qx, qy, vol_inv, sino, sino_scale, blockIdx.x, blockIdx.y, blockIdx.z);
}

// The gather form's launch: A-pair copies (16-bit volumes) when ap, else the
// z-pair copies; through the buffer path when buf, else 64-bit indices.
template <typename TV, typename TS>
void launch_rs_gathers(bool ap, bool buf, dim3 grid, hipStream_t s, const DevGeom& d, int cube, const void* cx,
                       const void* cy, float inv, TS* sino, float sino_scale) {
    if constexpr (sizeof(TV) == 2) {
        if (ap) {
            if (buf)
                hipLaunchKernelGGL((rs_kernel<TV, TS, true, true>), grid, dim3(BX, BY), 0, s, d, cube, cx, cy, inv,
                                   sino, sino_scale);
            else
                hipLaunchKernelGGL((rs_kernel<TV, TS, false, true>), grid, dim3(BX, BY), 0, s, d, cube, cx, cy, inv,
                                   sino, sino_scale);
            return;
        }
    }
    if (buf)
        hipLaunchKernelGGL((rs_kernel<TV, TS, true, false>), grid, dim3(BX, BY), 0, s, d, cube, cx, cy, inv, sino,
                           sino_scale);
    else
        hipLaunchKernelGGL((rs_kernel<TV, TS, false, false>), grid, dim3(BX, BY), 0, s, d, cube, cx, cy, inv, sino,
                           sino_scale);
}

// ---------------------------------------------------------------------------
// LDS-staged RS (FP32 volumes).
//
// A workgroup is 32 x 8 detector cells of one view.  When all its rays share
// the dominant in-plane axis A (|d_A| >= the other two) and its direction,
// it walks A in windows: for the window of samples with floor(A) in
// [wlo, whi] it copies the voxels those samples interpolate (planes
// wlo .. whi + 1, the window's box of z rows and B columns, stage_box.hpp)
// into LDS, and every lane takes its samples of the window with the eight
// taps read from there (four ds_read2_b32 per sample).  The staging is double
// buffered: the copy of window i + 1 is in flight while the workgroup samples
// window i, and one barrier per window separates them.  A window holds up to
// kRsWmax sample planes, fewer (6, 4, 2, then 1) when its box is too large
// for a buffer.  The box is not clipped to the grid: a sample outside the grid (a
// cylinder wider than a y extent smaller than x, or a ray whose clipped exit
// lies behind its entry) reads zero taps there and adds +0, as the range
// test of the per-ray loop makes it.  Samples, taps, arithmetic and the order
// of the sum are those of the per-ray loop, so results equal rs_kernel's bit
// for bit.  A window whose box exceeds a buffer even at one plane, and a
// workgroup whose rays do not share A and its direction, read their taps
// from the volume directly.
constexpr int kRsWmax = 8;  // sample planes per window at most

// The eight taps straight from the volume: the range and border rules and the
// arithmetic of trilinear_quads.
__device__ __forceinline__ float trilinear_direct(const float* __restrict__ f, int nx, int ny, int nz, float x,
                                                  float y, float z) {
    const float xf = floorf(x), yf = floorf(y), zf = floorf(z);
    if (!(xf >= -1.0f && xf <= (float)nx && yf >= -1.0f && yf <= (float)ny && zf >= -1.0f && zf <= (float)nz))
        return 0.0f;
    const int ix = (int)xf, iy = (int)yf, iz = (int)zf;
    const float fx = x - xf, fy = y - yf, fz = z - zf;
    const size_t sy = (size_t)nx, sz = (size_t)nx * ny;
    auto t = [&](int X, int Y, int Z) {
        const bool ok = (unsigned)X < (unsigned)nx && (unsigned)Y < (unsigned)ny && (unsigned)Z < (unsigned)nz;
        return ok ? f[(size_t)X + sy * (size_t)Y + sz * (size_t)Z] : 0.0f;
    };
    const float t000 = t(ix, iy, iz), t100 = t(ix + 1, iy, iz), t010 = t(ix, iy + 1, iz);
    const float t110 = t(ix + 1, iy + 1, iz), t001 = t(ix, iy, iz + 1), t101 = t(ix + 1, iy, iz + 1);
    const float t011 = t(ix, iy + 1, iz + 1), t111 = t(ix + 1, iy + 1, iz + 1);
    const float c00 = fmaf(fx, t100 - t000, t000), c10 = fmaf(fx, t110 - t010, t010);
    const float c01 = fmaf(fx, t101 - t001, t001), c11 = fmaf(fx, t111 - t011, t011);
    const float c0 = fmaf(fy, c10 - c00, c00), c1 = fmaf(fy, c11 - c01, c01);
    return fmaf(fz, c1 - c0, c0);
}

// A window: samples with floor(A) in [wlo, whi], staged planes wlo .. whi + 1
// (s.lo = wlo).  All fields are wave-uniform.
struct RsWindow {
    int wlo, whi;
    StageWin s;
};

// The window that starts at sample plane `from` in the travel direction: the
// largest of kRsWmax, 6, 4, 2, 1 planes whose box fits a buffer.
__device__ __forceinline__ RsWindow rs_plan(const WgCorners& wc, int from, bool inc, int max_nzr) {
    RsWindow w{};
    for (int W = kRsWmax;; W = W > 4 ? W - 2 : W >> 1) {
        const int wlo = inc ? from : from - W + 1, whi = wlo + W - 1;
        const StageBox b = stage_box(wc, (float)wlo, (float)(whi + 1), 0.0f, 0, 1);
        w = RsWindow{wlo, whi, stage_window(b, wlo, W + 1, max_nzr)};
        if (w.s.fits || W == 1) break;
    }
    return w;
}

#ifndef RS_STAGED_UNROLL
#define RS_STAGED_UNROLL 4
#endif

// The window walk of one workgroup along A (0: x, 1: y); inc: A increasing.
// src: rows of nBp floats along B (A = 0: the y-fastest copy, A = 1: the
// volume or its x-padded copy).  first, last: the workgroup's first and last
// sample planes.  wc: the corner rays (every wave plans the same windows).
template <int A>
__device__ __forceinline__ float rs_staged_walk(const DevGeom& g, const RsRay& r, int lr, bool inc, int first, int last,
                                                const float* __restrict__ src, int nBp, const float* __restrict__ vol,
                                                const WgCorners& wc, float* __restrict__ st0,
                                                float* __restrict__ st1, int lane, int wv) {
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const int nA = A == 0 ? nx : ny;
    const size_t src_elems = (size_t)nBp * (size_t)nA * (size_t)nz;
    float acc = 0.0f;
    int s = 0;
    const int max_nzr = stage_max_rows(nA, nBp);
    RsWindow cur = rs_plan(wc, first, inc, max_nzr);
    if (cur.s.fits) stage_copy(src, src_elems, nA, nBp, nz, cur.s, st0, lane, wv);
    // One window: window `cur` is in `buf`, the next one is copied into
    // `other`; false after the last window.  The windows alternate between
    // the two LDS arrays and each call names them directly, so the compiler
    // knows that the copy in flight does not write what the reads read and
    // does not wait for it before them.
    auto window = [&](const float* __restrict__ buf, float* __restrict__ other) __attribute__((always_inline)) {
        // this wave's copies of window `cur` have landed; every wave is done with the other buffer
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        const int from = inc ? cur.whi + 1 : cur.wlo - 1;
        const bool more = inc ? from <= last : from >= last;
        RsWindow nxt = cur;
        if (more) {
            nxt = rs_plan(wc, from, inc, max_nzr);
            if (nxt.s.fits) stage_copy(src, src_elems, nA, nBp, nz, nxt.s, other, lane, wv);
        }
        const float wlof = (float)cur.wlo, whif = (float)cur.whi;
        // offset of tap (a, z, b) in the buffer, ((a - wlo) nzr + z - zlo) P + b - blo4 with P = 4 nch
        // floats per row, as one exact FP32 expression of integers below 2^24 (the launcher admits
        // grids with nA kSlotFloats < 2^24)
        const int P = 4 * cur.s.nch;
        const int pstride = cur.s.nzr * P;
        const float ps = (float)pstride, pf = (float)P;
        const float c0 = -((float)cur.wlo * ps + (float)(cur.s.zlo * P + cur.s.blo4));
        for (;;) {
            float px[RS_STAGED_UNROLL], py[RS_STAGED_UNROLL], pz[RS_STAGED_UNROLL];
            bool act[RS_STAGED_UNROLL];
#pragma unroll
            for (int j = 0; j < RS_STAGED_UNROLL; ++j) {
                const float sf = (float)(s + j);
                px[j] = fmaf(sf, r.dx, r.ex);
                py[j] = fmaf(sf, r.dy, r.ey);
                pz[j] = fmaf(sf, r.dz, r.ez);
                const float fa = floorf(A == 0 ? px[j] : py[j]);
                act[j] = (s + j < lr) & (fa >= wlof) & (fa <= whif);
            }
            if (__ballot(act[0]) == 0ull) break;
            float v[RS_STAGED_UNROLL];
            if (cur.s.fits) {
                float fx[RS_STAGED_UNROLL], fy[RS_STAGED_UNROLL], fz[RS_STAGED_UNROLL];
                int off[RS_STAGED_UNROLL];
#pragma unroll
                for (int j = 0; j < RS_STAGED_UNROLL; ++j) {
                    const float xf = floorf(px[j]), yf = floorf(py[j]), zf = floorf(pz[j]);
                    fx[j] = px[j] - xf;
                    fy[j] = py[j] - yf;
                    fz[j] = pz[j] - zf;
                    const float af = A == 0 ? xf : yf, bf = A == 0 ? yf : xf;
                    const int o = (int)fmaf(af, ps, fmaf(zf, pf, bf + c0));
                    off[j] = act[j] ? o : 0;
                }
                float u[RS_STAGED_UNROLL][4], w[RS_STAGED_UNROLL][4];
#pragma unroll
                for (int j = 0; j < RS_STAGED_UNROLL; ++j) {
                    const float* q0 = buf + off[j];
                    const float* q1 = q0 + pstride;
                    // (b, z), (b, z+1), (b+1, z), (b+1, z+1) on planes ia, ia + 1
                    u[j][0] = q0[0];
                    u[j][1] = q0[P];
                    u[j][2] = q0[1];
                    u[j][3] = q0[P + 1];
                    w[j][0] = q1[0];
                    w[j][1] = q1[P];
                    w[j][2] = q1[1];
                    w[j][3] = q1[P + 1];
                }
#pragma unroll
                for (int j = 0; j < RS_STAGED_UNROLL; ++j) {
                    float t000, t100, t010, t110, t001, t101, t011, t111;
                    if (A == 0) {
                        t000 = u[j][0]; t001 = u[j][1]; t010 = u[j][2]; t011 = u[j][3];
                        t100 = w[j][0]; t101 = w[j][1]; t110 = w[j][2]; t111 = w[j][3];
                    } else {
                        t000 = u[j][0]; t001 = u[j][1]; t100 = u[j][2]; t101 = u[j][3];
                        t010 = w[j][0]; t011 = w[j][1]; t110 = w[j][2]; t111 = w[j][3];
                    }
                    const float c00 = fmaf(fx[j], t100 - t000, t000), c10 = fmaf(fx[j], t110 - t010, t010);
                    const float c01 = fmaf(fx[j], t101 - t001, t001), c11 = fmaf(fx[j], t111 - t011, t011);
                    const float e0 = fmaf(fy[j], c10 - c00, c00), e1 = fmaf(fy[j], c11 - c01, c01);
                    const float val = fmaf(fz[j], e1 - e0, e0);
                    v[j] = act[j] ? val : 0.0f;
                }
            } else {
#pragma unroll
                for (int j = 0; j < RS_STAGED_UNROLL; ++j)
                    v[j] = act[j] ? trilinear_direct(vol, nx, ny, nz, px[j], py[j], pz[j]) : 0.0f;
            }
            int adv = 0;
#pragma unroll
            for (int j = 0; j < RS_STAGED_UNROLL; ++j) {
                acc += v[j];
                adv += act[j] ? 1 : 0;
            }
            s += adv;
        }
        cur = nxt;
        return more;
    };
    while (window(st0, st1) && window(st1, st0)) {
    }
    return acc;
}

// Forward RS with LDS-staged taps (FP32 volume `vol`).  volT: its y-fastest
// copy with rows of nyp floats, volP: rows of nxp floats along x (the volume
// itself when nxp = nx); nxp, nyp are nx, ny rounded up to multiples of four.
// Grid: x = column group, y = view, z = row group.
template <typename TS>
__global__ __launch_bounds__(BX* BY, 5) void rs_staged_kernel(DevGeom g, int cube, const float* __restrict__ vol,
                                                             const float* __restrict__ volT,
                                                             const float* __restrict__ volP, int nxp, int nyp,
                                                             TS* __restrict__ sino, float sino_scale) {
    __shared__ float st0[kSlotFloats];
    __shared__ float st1[kSlotFloats];
    __shared__ int wsh[4][3];
    const int tid = threadIdx.y * BX + threadIdx.x;
    const int lane = tid & 63, wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int bx = blockIdx.x, p = blockIdx.y, bz = blockIdx.z;
    const int m = bx * BX + threadIdx.x;
    const int n = bz * BY + threadIdx.y;
    const bool inside = m < g.nu && n < g.nv;
    RsRay r{0, 0, 0, 0, 0, 0, 0};
    if (inside) r = rs_setup(g, cube, p, m, n);
    const bool ok = r.lr > 0;
    int code = -1;
    if (ok) {
        const float ax = fabsf(r.dx), ay = fabsf(r.dy), az = fabsf(r.dz);
        if (ax > 0.0f && ax >= ay && ax >= az) code = r.dx > 0.0f ? 0 : 1;
        else if (ay > 0.0f && ay >= ax && ay >= az) code = r.dy > 0.0f ? 2 : 3;
    }
    // The workgroup's mode: 0..3 (axis x / y, increasing / decreasing), 4: no
    // ray, -1: the rays differ.
    const unsigned long long okmask = __ballot(ok);
    int wmode = 4;
    if (okmask != 0ull) {
        wmode = -1;
#pragma unroll
        for (int c = 0; c < 4; ++c)
            if (__ballot(ok && code == c) == okmask) wmode = c;
    }
    if (lane == 0) wsh[wv][0] = wmode;
    __syncthreads();
    int mode = 4;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int mk = wsh[k][0];
        if (mk != 4) mode = (mode == 4 || mode == mk) ? mk : -1;
    }
    mode = __builtin_amdgcn_readfirstlane(mode);
    float acc = 0.0f;
    if (mode == -1) {
        for (int s = 0; s < r.lr; ++s) {
            const float sf = (float)s;
            acc += trilinear_direct(vol, g.nx, g.ny, g.nz, fmaf(sf, r.dx, r.ex), fmaf(sf, r.dy, r.ey),
                                    fmaf(sf, r.dz, r.ez));
        }
    } else if (mode != 4) {
        const int A = mode < 2 ? 0 : 1;
        const bool inc = (mode & 1) == 0;
        // the workgroup's first and last sample planes along the travel direction
        int first = inc ? INT_MAX : INT_MIN, last = inc ? INT_MIN : INT_MAX;
        if (ok) {
            const float e = A == 0 ? r.ex : r.ey, d = A == 0 ? r.dx : r.dy;
            first = (int)floorf(e);
            last = (int)floorf(fmaf((float)(r.lr - 1), d, e));
        }
        first = inc ? wave_min(first) : wave_max(first);
        last = inc ? wave_max(last) : wave_min(last);
        __syncthreads();  // every wave has read wsh[*][0]
        if (lane == 0) {
            wsh[wv][1] = first;
            wsh[wv][2] = last;
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            first = inc ? min(first, wsh[k][1]) : max(first, wsh[k][1]);
            last = inc ? max(last, wsh[k][2]) : min(last, wsh[k][2]);
        }
        first = __builtin_amdgcn_readfirstlane(first);
        last = __builtin_amdgcn_readfirstlane(last);
        const int mlo = bx * BX, mhi = min(bx * BX + BX - 1, g.nu - 1);
        const int nlo = bz * BY, nhi = min(bz * BY + BY - 1, g.nv - 1);
        const WgCorners wc = wg_corners(g, A, p, mlo, mhi, nlo, nhi);
        if (A == 0)
            acc = rs_staged_walk<0>(g, r, ok ? r.lr : 0, inc, first, last, volT, nyp, vol, wc, st0, st1, lane, wv);
        else
            acc = rs_staged_walk<1>(g, r, ok ? r.lr : 0, inc, first, last, volP, nxp, vol, wc, st0, st1, lane, wv);
    }
    if (inside) {
        const size_t o = (size_t)m + (size_t)g.nu * ((size_t)n + (size_t)g.nv * p);
        Storage<TS>::store(sino, o, r.lr > 0 ? acc * g.delta : 0.0f, sino_scale);
    }
}

// Shared launch (FP32 or BF16 (kItemBF16) z-pair copies, or BF16 A-pair
// copies (kItemApair), FP32 sinograms); item flags bit 0 selects the cube
// FOV.  No minimum-waves bound: under one
// (8 or 6 waves per SIMD, 52 or 74 VGPRs) the compiler waits for each
// sample's two gathers before it issues the next sample's, where unbounded
// (86 VGPRs, 5 waves per SIMD) it issues the eight gathers of four samples
// together, as rs_kernel does; the serialised form took 27% longer per
// 512^3 projection.  Its grid is what the chip holds at once
// (batch_grids_init), so every workgroup is resident from the start.
__global__ __launch_bounds__(BX* BY) void rs_batch_kernel(const BatchTable* __restrict__ t, uint32_t* __restrict__ next) {
    TOMO_BATCH_CLAIMED(t, next) {
        const BatchItem& it = t->item[batch_find(t, tile)];
        const uint32_t local = tile - it.begin;
        const uint32_t bx = local % it.gx, rest = local / it.gx;
        const int by = (int)(rest % it.gy), bz = (int)(rest / it.gy), cube = it.flags & 1;
        const bool bf = (it.flags & kItemBF16) != 0;
        const bool buf = rs_copy_shape(it.g.nx, it.g.ny, it.g.nz, bf ? 2u : 4u).buf;
        using BF = __hip_bfloat16;
        if (it.flags & kItemApair)
            rs_tile<BF, float, true, true>(it.g, cube, it.in, it.in2, 1.0f, (float*)it.out, 1.0f, (int)bx, by, bz);
        else if (bf && buf)
            rs_tile<BF, float, true>(it.g, cube, it.in, it.in2, 1.0f, (float*)it.out, 1.0f, (int)bx, by, bz);
        else if (bf)
            rs_tile<BF, float, false>(it.g, cube, it.in, it.in2, 1.0f, (float*)it.out, 1.0f, (int)bx, by, bz);
        else if (buf)
            rs_tile<float, float, true>(it.g, cube, it.in, it.in2, 1.0f, (float*)it.out, 1.0f, (int)bx, by, bz);
        else
            rs_tile<float, float, false>(it.g, cube, it.in, it.in2, 1.0f, (float*)it.out, 1.0f, (int)bx, by, bz);
    }
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
    if (((threadIdx.y * BX + threadIdx.x) & 63) == 0) atomicAdd(total, c);  // first lane of the wave
}

}  // namespace

int launch_rs(const tomo_geom* g, int cube, Buf vol, MutBuf sino, hipStream_t s, const ProjExtra& x) {
    const DevGeom& d = g->dev;
    const dim3 grid(ceil_div(d.nu, BX), d.nphi, ceil_div(d.nv, BY));
    // LDS-staged taps: FP32 volumes whose tap offsets stay exact in FP32
    // ((max(nx, ny) + 16) kSlotFloats < 2^24)
    if (x.staged && vol.dtype == TOMO_F32 && !x.index_reads && std::max(d.nx, d.ny) <= 4096) {
        StageSources src;
        hipError_t le = src.make((const float*)vol.p, d.nx, d.ny, d.nz, s);
        if (le == hipSuccess) {
            TOMO_DISPATCH_STORAGE(sino.dtype, TS, {
                hipLaunchKernelGGL((rs_staged_kernel<TS>), grid, dim3(BX, BY), 0, s, d, cube, (const float*)vol.p,
                                   (const float*)src.vt, src.rows_x, src.nxp, src.nyp, (TS*)sino.p, sino.scale);
            });
            le = hipGetLastError();
        }
        const hipError_t fe = src.release(s);
        TOMO_HIP(le);
        TOMO_HIP(fe);
        return TOMO_OK;
    }
    const float inv = 1.0f / vol.scale;
    const uint32_t es = vol.dtype == TOMO_F32 ? 4u : 2u;
    // 16-bit volumes take one load per sample from A-pair copies (the taps are
    // the same numbers), through the buffer path where those fit it
    const bool ap = es == 2;
    ZPairs zp;
    hipError_t le = zp.alloc(vol.p, es, d.nx, d.ny, d.nz, s, x, ap);
    const void* cx = zp.qx;
    const void* cy = zp.qy;
    if (le == hipSuccess) {
        const bool buf = !x.index_reads && (ap ? apair_buffer_path(d.nx, d.ny, d.nz)
                                               : rs_copy_shape(d.nx, d.ny, d.nz, es).buf);
        TOMO_DISPATCH_STORAGE(vol.dtype, TV, {
            TOMO_DISPATCH_STORAGE(sino.dtype, TS, {
                launch_rs_gathers<TV, TS>(ap, buf, grid, s, d, cube, cx, cy, inv, (TS*)sino.p, sino.scale);
            });
        });
        le = hipGetLastError();
    }
    const hipError_t fe = zp.release(s);
    TOMO_HIP(le);
    TOMO_HIP(fe);
    return TOMO_OK;
}

namespace {
std::atomic<int> g_rs_batch_grid{kBatchGrid};
}

int batch_grids_init(int device) {
    int per_cu = 0, cus = 0;
    TOMO_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, rs_batch_kernel, BX * BY, 0));
    TOMO_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
    int grid = per_cu * cus;
    grid -= grid % kBatchXcds;
    g_rs_batch_grid.store(std::min(std::max(grid, kBatchXcds), kBatchGrid));
    return TOMO_OK;
}

hipError_t launch_batch_rs(BatchTable* t, hipStream_t s) {
    hipLaunchKernelGGL(rs_batch_kernel, dim3(g_rs_batch_grid.load()), dim3(BX, BY), 0, s, t, t->next);
    return hipGetLastError();
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
