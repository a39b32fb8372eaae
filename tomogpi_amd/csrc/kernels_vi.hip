// kernels_vi.hip - voxel-driven bilinear backprojector (VI) and its exact
// transpose.
//
// Restates K/BackProjector/backprojection3D_kernel_16reg_UM.cuh:23-151 (the
// reference's device-resident VI kernel): for voxel (i,j,k), x'=i-cx,
// y'=j-cy, z'=k-cz and view p,
//   w = delta(cos x' + sin y') + FOD
//   u = (FDD/du) delta (-sin x' + cos y') / w + cu
//   v = A z' + cv,   A = (FDD/dv) delta / w
//   f = delta * sum_p bilinear(g_p, u, v)      (taps outside the detector = 0)
// The output is overwritten.  As in the reference's "16reg" blocking, a
// thread owns a column of 16 voxels that share u and A; the per-(column,view)
// terms w, 1/w, u and A are evaluated in FP64 (the geometry of the FP32 form)
// and the per-voxel v, bilinear weights and sum in FP32.  A BF16 sinogram is
// widened to FP32 (value = stored / scale) before the same kernel; an FP16
// sinogram is read as u-pairs of FP16 taps (vi_pair_kernel) and interpolated
// with v_dot2_f32_f16 (FP16 u-weights in steps of 1/2048, FP32 sums).
#include <algorithm>
#include <climits>
#include <cmath>
#include <type_traits>

#include "device_common.hpp"
#include "batch.hpp"

namespace tomo {
namespace {

constexpr int TX = kViTileX, TY = kViTileY, ZB = kViTileZ;  // VI^T workgroup

struct ColumnView {
    float u, A;
    float num, base;  // FDK weight terms (FDK only)
};

// Per (column, view) terms in FP64.  FDK (back_fdk,
// K/BackProjector/FDK3D_kernel_16reg.cuh:71-88; constants C/BackProjector.cu:
// 34-35) adds num = alphaC (FDD/du) / w^2 and base = betaC + u'^2, u' = u - cu
// in pixels, alphaC = (float)(pi FOD FDD / Nphi), betaC = (FDD/du)^2.  A
// column-view with a non-finite u or A contributes 0: it reads off the
// detector (u = -8, A = 0) with a zero weight.
template <bool FDK>
__device__ __forceinline__ ColumnView column_view(const DevGeom& g, const double* c, double xp, double yp) {
    const double w = fma(c[2], xp, fma(c[3], yp, g.fod));
    const double inv = 1.0 / w;
    const double un = fma(c[4], xp, c[5] * yp);
    ColumnView r;
    r.u = (float)fma(un, inv, (double)g.cu);
    r.A = (float)(c[6] * inv);
    r.num = 0.0f;
    r.base = 1.0f;
    const bool fin = __builtin_isfinite(r.u) && __builtin_isfinite(r.A);
    if (FDK && fin) {
        const double ku = g.fdd / (double)g.du;
        const double alpha_c = (double)(float)(M_PI * g.fod * g.fdd / (double)g.nphi);
        const double up = un * inv;
        r.num = (float)(alpha_c * ku / (w * w));
        r.base = (float)(ku * ku + up * up);
    }
    if (!fin) {
        r.u = -8.0f;
        r.A = 0.0f;
    }
    return r;
}

// ---------------------------------------------------------------------------
// LDS-staged VI for FP32 sinograms.
//
// A workgroup owns SX x SY voxel columns of ZS voxels.  For each view it
// copies the detector taps its voxels read into LDS with coalesced loads
// (taps off the detector as 0), and each voxel-view update reads its four
// taps from there with two ds_read2_b32 instead of one 16-byte gather from
// the tap-quad copy; the arithmetic is vi_tile's.  The taps are the box of
// rows [v0, v1 + 1] and columns [u0, u1 + 1] with (u0, u1) the extreme
// clamped cell corners floor(u) of the tile's columns and (v0, v1) the
// extreme floor(v) of its voxels.  A table holds, per (column tile, view),
// u0, u1 and the extreme A of the tile's columns (vi_box_kernel, the same
// FP64 column terms); since v = fmaf(z', A, cv) is monotone in z' and in A,
// the row extremes are the floors at the four (z', A) corners, so the box is
// exact.  The rows are not clamped as in vi_tile: a row off the detector
// holds zeros in the box, as the clamped cell it stands for.  View p is
// staged in buffer p & 1, so one barrier per view separates the copy from
// the reads.  A view whose box exceeds the buffer reads the taps directly.
constexpr int SX = kViStageX, SY = kViStageY;  // voxel columns per workgroup; a wave is 8 x 8 columns
// Floats per staged row.  A wave's 64 lanes read taps (u, r) at word
// r * pitch + u; lanes spread over about 20 columns and a few rows, so two
// lanes share a bank when their row and column offsets satisfy
// dr * pitch + du = 0 mod 32.  With 8 x 8 waves a pitch of 75 (11 mod 32)
// keeps such pairs rare at 512^3 / 512 views; with 32 x 2 waves (40 columns,
// past the 32 banks) and a pitch of 65 (1 mod 32) a ds_read2_b32 took 2.6
// extra cycles on bank conflicts.
constexpr int kStagePitch = 75;

// Staging buffer of a workgroup of 16-voxel columns: up to 38 rows of 64
// taps (at 512^3 / 512 views about 0.1% of the boxes need 39 or 40 and read
// their taps directly); two of them are 22.8 KB of LDS, 7 workgroups per CU
// (FDK: 6, its VGPRs).
template <int ZS>
struct StageShape {
    static_assert(ZS == 16, "staging buffer sized for 16-voxel columns");
    static_assert(SX == 32 && SY == 8, "four 8 x 8 waves per workgroup");
    static constexpr int rows = 38;                  // staged rows per view
    static constexpr int taps = rows * kStagePitch;  // floats per buffer
};

__device__ __forceinline__ int wave_min(int v) { return __reduce_min_sync(~0ull, v); }
__device__ __forceinline__ int wave_max(int v) { return __reduce_max_sync(~0ull, v); }

// Order-preserving map of a finite float to int: x < y  <=>  key(x) < key(y).
__device__ __forceinline__ int fkey(float v) {
    const int b = __float_as_int(v);
    return b >= 0 ? b : (int)(0x80000000u - (uint32_t)b);
}
__device__ __forceinline__ float fkey_inv(int k) {
    return __int_as_float(k >= 0 ? k : (int)(0x80000000u - (uint32_t)k));
}

// The ViBox of column tile `tile` (row-major over ntx tiles per row) in view
// p, one workgroup of SX x SY threads (one per column).
__device__ __forceinline__ void vi_box_tile(const DevGeom& g, ViBox* __restrict__ box, int ntx, uint32_t tile,
                                            uint32_t p) {
    const int i = (int)(tile % ntx) * SX + (int)(threadIdx.x % SX);
    const int j = (int)(tile / ntx) * SY + (int)(threadIdx.x / SX);
    int k[4] = {INT_MAX, INT_MIN, INT_MAX, INT_MIN};
    if (i < g.nx && j < g.ny) {
        const double xp = (double)i - (double)g.cx, yp = (double)j - (double)g.cy;
        const ColumnView cvw = column_view<false>(g, g.view + (size_t)p * kViewStride, xp, yp);
        const int uc = (int)__builtin_amdgcn_fmed3f(floorf(cvw.u), -2.0f, (float)g.nu);
        k[0] = k[1] = uc;
        k[2] = k[3] = fkey(cvw.A);
    }
    __shared__ int part[SX * SY / 64][4];
    const int w = threadIdx.x >> 6;
    const int r0 = wave_min(k[0]), r1 = wave_max(k[1]), r2 = wave_min(k[2]), r3 = wave_max(k[3]);
    __syncthreads();  // the previous tile's reads of part are done
    if ((threadIdx.x & 63) == 0) {
        part[w][0] = r0;
        part[w][1] = r1;
        part[w][2] = r2;
        part[w][3] = r3;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int m[4] = {part[0][0], part[0][1], part[0][2], part[0][3]};
        for (int q = 1; q < SX * SY / 64; ++q) {
            m[0] = min(m[0], part[q][0]);
            m[1] = max(m[1], part[q][1]);
            m[2] = min(m[2], part[q][2]);
            m[3] = max(m[3], part[q][3]);
        }
        ViBox b;
        b.u0 = m[0];
        b.u1 = m[1];
        b.amin = fkey_inv(m[2]);
        b.amax = fkey_inv(m[3]);
        box[(size_t)tile * g.nphi + p] = b;
    }
}

// One workgroup per (column tile, view).
__global__ __launch_bounds__(SX* SY) void vi_box_kernel(DevGeom g, ViBox* __restrict__ box, int ntx) {
    vi_box_tile(g, box, ntx, blockIdx.x / (uint32_t)g.nphi, blockIdx.x % (uint32_t)g.nphi);
}

// Per-view staging plan of a workgroup: the box of taps (rows from v0,
// columns from u0) and whether it fits the buffer.  Wave w copies box rows
// w, w + 4, ..., one column per lane, so a box is at most 64 columns wide.
struct StagePlan {
    int u0, v0, cols, rows;
    bool fits;
};

// The box rows from the four (z', A) corners; with A >= 0 (a wave-uniform
// sign test on the table's bits) the lowest row is at z' = zp0 with the
// largest A when zp0 < 0 (the smallest otherwise) and the highest at z' = zp1
// with the largest A when zp1 > 0, the same values with two fmas.
template <int ZS>
__device__ __forceinline__ StagePlan stage_plan(const ViBox& bb, float zp0, float cv, bool lo_amax, bool hi_amax) {
    const float zp1 = zp0 + (float)(ZS - 1);
    float vlo, vhi;
    if (__float_as_int(bb.amin) >= 0) {
        vlo = fmaf(zp0, lo_amax ? bb.amax : bb.amin, cv);
        vhi = fmaf(zp1, hi_amax ? bb.amax : bb.amin, cv);
    } else {
        const float va = fmaf(zp0, bb.amin, cv), vb = fmaf(zp0, bb.amax, cv);
        const float vc = fmaf(zp1, bb.amin, cv), vd = fmaf(zp1, bb.amax, cv);
        vlo = fminf(fminf(va, vb), fminf(vc, vd));
        vhi = fmaxf(fmaxf(va, vb), fmaxf(vc, vd));
    }
    vlo = floorf(vlo);
    vhi = floorf(vhi);
    StagePlan s;
    s.u0 = bb.u0;
    s.cols = bb.u1 - bb.u0 + 2;
    const bool fits = vlo >= -32768.0f && vhi <= 32768.0f && vhi - vlo + 2.0f <= (float)StageShape<ZS>::rows &&
                      s.cols <= 64;
    // wave-uniform: keep the plan in scalar registers
    s.fits = __builtin_amdgcn_readfirstlane((int)fits) != 0;
    s.v0 = __builtin_amdgcn_readfirstlane(s.fits ? (int)vlo : 0);
    s.rows = __builtin_amdgcn_readfirstlane(s.fits ? (int)(vhi - vlo) + 2 : 0);
    return s;
}

// FP16 u-weights (1 - w, w) of a column-view, w = fu rounded to a multiple of
// 1/2048: both are exact FP16 values and sum to 1, so a constant row
// interpolates to itself.
__device__ __forceinline__ uint32_t pair_weights(float fu) {
    typedef _Float16 h2 __attribute__((ext_vector_type(2)));
    const float w1 = rintf(fu * 2048.0f) * (1.0f / 2048.0f);
    return __builtin_bit_cast(uint32_t, h2{(_Float16)(1.0f - w1), (_Float16)w1});
}

// Bilinear sample from two u-pair words (rows v and v + 1): r0 = w . P0 and
// the row difference w . P1 - r0, each one v_dot2_f32_f16 (products of FP16
// values are exact in FP32), then r0 + fv (r1 - r0).  The builtins leave the
// wait states a dot2 result needs before its use to the compiler.
__device__ __forceinline__ float pair_sample(uint32_t p0, uint32_t p1, uint32_t w, float fv) {
    typedef _Float16 h2 __attribute__((ext_vector_type(2)));
    const h2 hw = __builtin_bit_cast(h2, w);
    const float r0 = __builtin_amdgcn_fdot2(__builtin_bit_cast(h2, p0), hw, 0.0f, false);
    const float dr = __builtin_amdgcn_fdot2(__builtin_bit_cast(h2, p1), hw, -r0, false);
    return fmaf(fv, dr, r0);
}

// Four bilinear samples at once (the staged path's inner loop): the four
// r0 = w . P0 first, then the four w . P1 - r0, so each dot2 result is read
// three or more instructions after it is written (the wait a VALU read of a
// dot2 result needs), and s_nop 2 before any later read of the last one.
__device__ __forceinline__ void pair_sample4(const uint32_t (&p0)[4], const uint32_t (&p1)[4], uint32_t w,
                                             float (&r0)[4], float (&dr)[4]) {
    asm volatile(
        "v_dot2_f32_f16 %0, %8, %16, 0\n\t"
        "v_dot2_f32_f16 %1, %9, %16, 0\n\t"
        "v_dot2_f32_f16 %2, %10, %16, 0\n\t"
        "v_dot2_f32_f16 %3, %11, %16, 0\n\t"
        "v_dot2_f32_f16 %4, %12, %16, %0 neg_lo:[0,0,1]\n\t"
        "v_dot2_f32_f16 %5, %13, %16, %1 neg_lo:[0,0,1]\n\t"
        "v_dot2_f32_f16 %6, %14, %16, %2 neg_lo:[0,0,1]\n\t"
        "v_dot2_f32_f16 %7, %15, %16, %3 neg_lo:[0,0,1]\n\t"
        "s_nop 2"
        : "=&v"(r0[0]), "=&v"(r0[1]), "=&v"(r0[2]), "=&v"(r0[3]), "=&v"(dr[0]), "=&v"(dr[1]), "=&v"(dr[2]),
          "=&v"(dr[3])
        : "v"(p0[0]), "v"(p0[1]), "v"(p0[2]), "v"(p0[3]), "v"(p1[0]), "v"(p1[1]), "v"(p1[2]), "v"(p1[3]), "v"(w));
}

// PAIR: `sino` holds u-pair words (vi_pair_kernel; nu + 1 per row) instead of
// FP32 taps; the box copy and the tap offsets are the same, one word per row.
template <typename TV, bool FDK, int ZS, bool PAIR = false>
__device__ __forceinline__ void vi_staged_tile(const DevGeom& g, const float* __restrict__ sino,
                                               const ViBox* __restrict__ boxes, TV* __restrict__ vol,
                                               float vol_scale, float post, int bx, int by, int bz,
                                               float* __restrict__ stage) {
    using SS = StageShape<ZS>;
    constexpr int NT = SX * SY;
    constexpr int NW = NT / 64;                             // waves per workgroup
    constexpr int kSlots = (SS::rows + NW - 1) / NW;        // box rows per wave at most
    const int tid = threadIdx.x;
    const int lane = tid & 63, wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    // wave wv covers columns x in [8 wv, 8 wv + 8), y in [0, 8)
    const int i = bx * SX + wv * 8 + (lane & 7);
    const int j = by * SY + (lane >> 3);
    const int k0 = bz * ZS;
    const bool active = i < g.nx && j < g.ny;
    const double xp = (double)i - (double)g.cx, yp = (double)j - (double)g.cy;
    const float zp0 = (float)((double)k0 - (double)g.cz);
    const float cv = g.cv;
    const float fnu = (float)g.nu;
    const int nu = g.nu, nv = g.nv;
    const int rl = PAIR ? nu + 1 : nu;  // words per detector row
    const size_t plane = (size_t)rl * nv;
    // The box and view tables are read-only here and indexed by wave-uniform
    // values: read through the constant address space they come in with scalar
    // loads, and the staging plan's branches are uniform.
    typedef const __attribute__((address_space(4))) ViBox* ConstBox;
    typedef const __attribute__((address_space(4))) double* ConstView;
    const ConstBox bt = (ConstBox)(boxes + (size_t)(by * ((g.nx + SX - 1) / SX) + bx) * g.nphi);
    float acc[ZS];
#pragma unroll
    for (int k = 0; k < ZS; ++k) acc[k] = 0.0f;
    auto add = [&](const ColumnView& cvw, int k, float r) {
        if (FDK) {
            const float vp = (zp0 + (float)k) * cvw.A;
            const float wt = cvw.num / sqrtf(fmaf(vp, vp, cvw.base));
            acc[k] += wt * r;
        } else {
            acc[k] += r;
        }
    };
    auto term = [&](const ColumnView& cvw, float fu, int k, float v, float vf, float t00, float t01, float t10,
                    float t11) {
        // t00 = (u, v), t01 = (u, v+1), t10 = (u+1, v), t11 = (u+1, v+1)
        const float fv = v - vf;
        const float a = fmaf(fu, t10 - t00, t00);
        const float b = fmaf(fu, t11 - t01, t01);
        add(cvw, k, fmaf(fv, b - a, a));
    };
    // Copy of view p's box into buffer b: wave wv copies box rows wv + q NW,
    // lane l column l, straight from memory into LDS (buffer_load ... lds).
    // Each row is read through a buffer resource of that detector row (of no
    // bytes for a row off the detector), so a tap off the detector reads past
    // the resource's end, which writes 0; the lane offsets are the same for
    // every row.
    auto copy_box = [&](const StagePlan& s, int p, int b) {
        if (!s.fits) return;
        const float* view = sino + (size_t)p * plane;
        // past the row's end unless 0 <= u < nu (FP32) or the pair index u + 1 <= nu
        const uint32_t loff = (uint32_t)(s.u0 + (PAIR ? 1 : 0) + lane) * 4u;
        float* dst = stage + b * SS::taps + wv * kStagePitch;
#pragma unroll
        for (int q = 0; q < kSlots; ++q) {
            if (wv + q * NW < s.rows) {
                const int v = s.v0 + wv + q * NW;
                const bool vin = (unsigned)v < (unsigned)nv;
                const __amdgpu_buffer_rsrc_t rr =
                    buffer_rsrc(view + (size_t)(vin ? v : 0) * rl, vin ? (uint32_t)rl * 4u : 0u);
                __builtin_amdgcn_raw_ptr_buffer_load_lds(
                    rr, (__attribute__((address_space(3))) void*)(dst + q * NW * kStagePitch), 4, loff, 0, 0, 0);
            }
        }
    };
    const bool lo_amax = zp0 < 0.0f, hi_amax = zp0 + (float)(ZS - 1) > 0.0f;
    const uint32_t view_bytes = (uint32_t)(plane * sizeof(float));
    auto box_of = [&](int p) { return ViBox{bt[p].u0, bt[p].u1, bt[p].amin, bt[p].amax}; };
    StagePlan nxt = stage_plan<ZS>(box_of(0), zp0, cv, lo_amax, hi_amax);
    copy_box(nxt, 0, 0);
    for (int p = 0; p < g.nphi; ++p) {
        // this wave's copies of view p have landed; every wave is done with view p - 1
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        const StagePlan now = nxt;
        if (p + 1 < g.nphi) {
            nxt = stage_plan<ZS>(box_of(p + 1), zp0, cv, lo_amax, hi_amax);
            copy_box(nxt, p + 1, (p + 1) & 1);
        }
        if (active) {
            const ConstView vt = (ConstView)(g.view + (size_t)p * kViewStride);
            double c[kViewStride];
#pragma unroll
            for (int e = 0; e < kViewStride; ++e) c[e] = vt[e];
            const ColumnView cvw = column_view<FDK>(g, c, xp, yp);
            const float uf = floorf(cvw.u);
            const float fu = cvw.u - uf;
            const float uc = __builtin_amdgcn_fmed3f(uf, -2.0f, fnu);
            const uint32_t pw = PAIR ? pair_weights(fu) : 0u;
            if (now.fits) {
                // byte offset in the buffer of tap (uc, vf): ((vf - v0) pitch + uc - u0) 4,
                // one exact fma of small integers
                const float colb =
                    (float)(((int)uc - now.u0 - now.v0 * kStagePitch + (p & 1) * SS::taps) * 4);
                if constexpr (PAIR) {
                    static_assert(ZS % 4 == 0, "u-pair samples go four voxels at a time");
#pragma unroll
                    for (int k4 = 0; k4 < ZS; k4 += 4) {
                        uint32_t p0[4], p1[4];
                        float fv[4], r0[4], dr[4];
#pragma unroll
                        for (int t = 0; t < 4; ++t) {
                            const float v = fmaf(zp0 + (float)(k4 + t), cvw.A, cv);
                            const float vf = floorf(v);
                            const uint32_t off = (uint32_t)fmaf(vf, (float)(kStagePitch * 4), colb);
                            const uint32_t* qw = (const uint32_t*)((const char*)stage + off);
                            p0[t] = qw[0];
                            p1[t] = qw[kStagePitch];
                            fv[t] = v - vf;
                        }
                        pair_sample4(p0, p1, pw, r0, dr);
#pragma unroll
                        for (int t = 0; t < 4; ++t) add(cvw, k4 + t, fmaf(fv[t], dr[t], r0[t]));
                    }
                } else {
#pragma unroll
                    for (int k = 0; k < ZS; ++k) {
                        const float v = fmaf(zp0 + (float)k, cvw.A, cv);
                        const float vf = floorf(v);
                        const uint32_t off = (uint32_t)fmaf(vf, (float)(kStagePitch * 4), colb);
                        const float* q = (const float*)((const char*)stage + off);
                        term(cvw, fu, k, v, vf, q[0], q[kStagePitch], q[1], q[kStagePitch + 1]);
                    }
                }
            } else {
                // box larger than the buffer: the four taps straight from the view, one voxel at a time
                const __amdgpu_buffer_rsrc_t rsrc = buffer_rsrc(sino + (size_t)p * plane, view_bytes);
                const int iu = (int)uc;
                const bool u0ok = (unsigned)iu < (unsigned)nu, u1ok = (unsigned)(iu + 1) < (unsigned)nu;
                const bool pok = (unsigned)(iu + 1) <= (unsigned)nu;  // pair word iu + 1 exists
                ColumnView cw = cvw;
                asm volatile("" : "+v"(cw.A));  // this path's rows are its own, not hoisted above the branch
                auto rd = [&](bool ok, int v, int u) {
                    const uint32_t off = ok ? (uint32_t)(v * rl + u) * 4u : 0xfffffff0u;
                    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rsrc, (int)off, 0, 0));
                };
#pragma unroll
                for (int k = 0; k < ZS; ++k) {
                    const float v = fmaf(zp0 + (float)k, cw.A, cv);
                    const float vf = floorf(v);
                    const int iv = (int)__builtin_amdgcn_fmed3f(vf, -2.0f, (float)nv);
                    const bool v0ok = (unsigned)iv < (unsigned)nv, v1ok = (unsigned)(iv + 1) < (unsigned)nv;
                    if constexpr (PAIR) {
                        const uint32_t p0 = __builtin_bit_cast(uint32_t, rd(pok && v0ok, iv, iu + 1));
                        const uint32_t p1 = __builtin_bit_cast(uint32_t, rd(pok && v1ok, iv + 1, iu + 1));
                        add(cw, k, pair_sample(p0, p1, pw, v - vf));
                    } else {
                        term(cw, fu, k, v, vf, rd(u0ok && v0ok, iv, iu), rd(u0ok && v1ok, iv + 1, iu),
                             rd(u1ok && v0ok, iv, iu + 1), rd(u1ok && v1ok, iv + 1, iu + 1));
                    }
                    __builtin_amdgcn_sched_barrier(0);
                }
            }
        }
    }
    if (!active) return;
    const size_t sxy = (size_t)g.nx * g.ny;
    const size_t base = (size_t)i + (size_t)g.nx * j;
#pragma unroll
    for (int k = 0; k < ZS; ++k)
        if (k0 + k < g.nz)
            Storage<TV>::store(vol, base + (size_t)(k0 + k) * sxy, FDK ? acc[k] * post : (acc[k] * g.delta) * post,
                               vol_scale);
}

// post: 1 for FP32 taps, 1 / storage scale for FP16 u-pairs.
template <typename TV, bool FDK, int ZS, bool PAIR>
__global__ __launch_bounds__(SX* SY, FDK ? 6 : 7) void vi_staged_kernel(DevGeom g, const float* __restrict__ sino,
                                                         const ViBox* __restrict__ boxes, TV* __restrict__ vol,
                                                         float vol_scale, float post) {
    __shared__ float stage[2 * StageShape<ZS>::taps];
    vi_staged_tile<TV, FDK, ZS, PAIR>(g, sino, boxes, vol, vol_scale, post, blockIdx.x, blockIdx.y, blockIdx.z,
                                      stage);
}

// u-pair words of an FP16 sinogram, one workgroup per detector row: word j in
// [0, nu] of the row holds taps (j - 1, j) as its low and high halves, a tap
// off the row as 0.
__global__ __launch_bounds__(256) void vi_pair_kernel(const uint16_t* __restrict__ sino,
                                                      uint32_t* __restrict__ pairs, int nu) {
    const size_t row = blockIdx.x;
    const uint16_t* src = sino + row * (size_t)nu;
    uint32_t* dst = pairs + row * (size_t)(nu + 1);
    for (int j = threadIdx.x; j <= nu; j += blockDim.x) {
        const uint32_t lo = j >= 1 ? src[j - 1] : 0u;
        const uint32_t hi = j < nu ? src[j] : 0u;
        dst[j] = lo | (hi << 16);
    }
}

// Shared launch (FP32 sinograms and volumes): tile grid (ceil(nx/SX),
// ceil(ny/SY), ceil(nz/16)) on the sinogram `in` with the problem's box table
// `aux` (launch_vi_boxes); the stored value is (delta * sum) * a.
__global__ __launch_bounds__(SX* SY, 7) void vi_batch_kernel(const BatchTable* __restrict__ t, uint32_t* __restrict__ next) {
    __shared__ float stage[2 * StageShape<kViTileZ>::taps];
    TOMO_BATCH_CLAIMED(t, next) {
        const BatchItem& it = t->item[batch_find(t, tile)];
        const uint32_t local = tile - it.begin;
        const uint32_t bx = local % it.gx, rest = local / it.gx;
        const int by = (int)(rest % it.gy), bz = (int)(rest / it.gy);
        vi_staged_tile<float, false, kViTileZ>(it.g, (const float*)it.in, (const ViBox*)it.aux, (float*)it.out, 1.0f,
                                               it.a, (int)bx, by, bz, stage);
    }
}

// VI^T: g_p[tap] += delta * w_tap * f(i,j,k) (FP32 atomics, output zeroed by
// the caller).  Same geometry and weights as vi_kernel.
template <typename TV>
__global__ __launch_bounds__(TX* TY) void vi_adjoint_kernel(DevGeom g, const TV* __restrict__ vol, float vol_inv,
                                                          float* __restrict__ sino) {
    const int i = blockIdx.x * TX + (threadIdx.x % TX);
    const int j = blockIdx.y * TY + (threadIdx.x / TX);
    const int k0 = blockIdx.z * ZB;
    if (i >= g.nx || j >= g.ny) return;
    const double xp = (double)i - (double)g.cx, yp = (double)j - (double)g.cy;
    const float zp0 = (float)((double)k0 - (double)g.cz);
    const size_t sxy = (size_t)g.nx * g.ny;
    const size_t base = (size_t)i + (size_t)g.nx * j;
    float fv[ZB];
#pragma unroll
    for (int k = 0; k < ZB; ++k)
        fv[k] = (k0 + k < g.nz) ? Storage<TV>::load(vol, base + (size_t)(k0 + k) * sxy, vol_inv) * g.delta : 0.0f;
    const size_t plane = (size_t)g.nu * g.nv;
    const int nu = g.nu, nv = g.nv;
    for (int p = 0; p < g.nphi; ++p) {
        const ColumnView cvw = column_view<false>(g, g.view + (size_t)p * kViewStride, xp, yp);
        float* gp = sino + (size_t)p * plane;
        const float uf = floorf(cvw.u);
        if (!(uf >= -1.0f && uf <= (float)nu)) continue;
        const int iu = (int)uf;
        const float fu = cvw.u - uf;
        const bool u0 = iu >= 0 && iu < nu, u1 = iu + 1 >= 0 && iu + 1 < nu;
        for (int k = 0; k < ZB; ++k) {
            if (fv[k] == 0.0f) continue;
            const float v = fmaf(zp0 + (float)k, cvw.A, g.cv);
            const float vf = floorf(v);
            if (!(vf >= -1.0f && vf <= (float)nv)) continue;
            const int iv = (int)vf;
            const float fvv = v - vf;
            const bool v0 = iv >= 0 && iv < nv, v1 = iv + 1 >= 0 && iv + 1 < nv;
            const int64_t o = (int64_t)iv * nu + iu;
            const float val = fv[k];
            if (u0 && v0) unsafeAtomicAdd(gp + o, (1.0f - fu) * (1.0f - fvv) * val);
            if (u1 && v0) unsafeAtomicAdd(gp + o + 1, fu * (1.0f - fvv) * val);
            if (u0 && v1) unsafeAtomicAdd(gp + o + nu, (1.0f - fu) * fvv * val);
            if (u1 && v1) unsafeAtomicAdd(gp + o + nu + 1, fu * fvv * val);
        }
    }
}

}  // namespace

// Voxels per thread column of the staged kernel.
constexpr int kStagedZ = kViTileZ;

// The staged kernel on an FP32 sinogram, with the geometry's box table
// (`boxes`) or, when null, one made for this call from the stream-ordered
// pool.
template <typename TV, bool PAIR>
static hipError_t launch_vi_staged(const DevGeom& d, int fdk, const float* sino, const void* boxes, void* vol,
                                   float vol_scale, float post, hipStream_t s) {
    const int ntx = ceil_div(d.nx, SX), nty = ceil_div(d.ny, SY);
    void* own = nullptr;
    hipError_t e = hipSuccess;
    if (!boxes) {
        if ((e = hipMallocAsync(&own, vi_box_count(d) * sizeof(ViBox), s)) != hipSuccess) return e;
        hipLaunchKernelGGL(vi_box_kernel, dim3((unsigned)vi_box_count(d)), dim3(SX * SY), 0, s, d, (ViBox*)own, ntx);
        e = hipGetLastError();
        boxes = own;
    }
    if (e == hipSuccess) {
        const dim3 grid(ntx, nty, ceil_div(d.nz, kStagedZ));
        if (fdk)
            hipLaunchKernelGGL((vi_staged_kernel<TV, true, kStagedZ, PAIR>), grid, dim3(SX * SY), 0, s, d, sino,
                               (const ViBox*)boxes, (TV*)vol, vol_scale, post);
        else
            hipLaunchKernelGGL((vi_staged_kernel<TV, false, kStagedZ, PAIR>), grid, dim3(SX * SY), 0, s, d, sino,
                               (const ViBox*)boxes, (TV*)vol, vol_scale, post);
        e = hipGetLastError();
    }
    if (own) {
        const hipError_t f = hipFreeAsync(own, s);
        if (e == hipSuccess) e = f;
    }
    return e;
}

int launch_vi(const tomo_geom* g, Buf sino, MutBuf vol, hipStream_t s, int fdk) {
    const DevGeom& d = g->dev;
    if (sino.dtype != TOMO_F32 && sino.dtype != TOMO_F16 && sino.dtype != TOMO_BF16)
        return fail(TOMO_ERR_UNSUPPORTED, "unsupported element type");
    if (vol.dtype != TOMO_F32 && vol.dtype != TOMO_F16 && vol.dtype != TOMO_BF16)
        return fail(TOMO_ERR_UNSUPPORTED, "unsupported element type");
    // byte offsets within one view are 32-bit
    if ((size_t)(d.nu + 1) * d.nv >= (1ull << 30))
        return fail(TOMO_ERR_UNSUPPORTED, "VI needs views of fewer than 2^30 cells");
    // An FP16 sinogram becomes u-pair words, a BF16 one is widened (value =
    // stored / scale) to FP32; either copy comes from the stream-ordered pool
    // and is released after the backprojection on the same stream.
    const bool pair = sino.dtype == TOMO_F16;
    const float* taps = (const float*)sino.p;
    void* wide = nullptr;
    if (pair) {
        const size_t rows = (size_t)d.nv * d.nphi;
        if (rows >= (1ull << 31)) return fail(TOMO_ERR_UNSUPPORTED, "VI needs fewer than 2^31 detector rows");
        TOMO_HIP(hipMallocAsync(&wide, rows * (size_t)(d.nu + 1) * sizeof(uint32_t), s));
        if (rows > 0) hipLaunchKernelGGL(vi_pair_kernel, dim3((unsigned)rows), dim3(256), 0, s,
                                         (const uint16_t*)sino.p, (uint32_t*)wide, d.nu);
        const hipError_t le = hipGetLastError();
        if (le != hipSuccess) {
            (void)hipFreeAsync(wide, s);
            TOMO_HIP(le);
        }
        taps = (const float*)wide;
    } else if (sino.dtype != TOMO_F32) {
        TOMO_HIP(hipMallocAsync(&wide, g->nsino() * sizeof(float), s));
        const int rc = launch_convert(sino, MutBuf{wide, TOMO_F32, 1.0f}, g->nsino(), s);
        if (rc) {
            (void)hipFreeAsync(wide, s);
            return rc;
        }
        taps = (const float*)wide;
    }
    const float post = pair ? 1.0f / sino.scale : 1.0f;
    hipError_t e = hipSuccess;
#define TOMO_VI_STAGED(TV)                                                                              \
    e = pair ? launch_vi_staged<TV, true>(d, fdk, taps, g->d_vi_boxes, vol.p, vol.scale, post, s)     \
             : launch_vi_staged<TV, false>(d, fdk, taps, g->d_vi_boxes, vol.p, vol.scale, post, s)
    switch (vol.dtype) {
        case TOMO_F32: TOMO_VI_STAGED(float); break;
        case TOMO_F16: TOMO_VI_STAGED(__half); break;
        default: TOMO_VI_STAGED(__hip_bfloat16); break;
    }
#undef TOMO_VI_STAGED
    if (wide) TOMO_HIP(hipFreeAsync(wide, s));
    TOMO_HIP(e);
    return TOMO_OK;
}

int launch_vi_boxes(const DevGeom& d, void* boxes, hipStream_t s) {
    const int ntx = ceil_div(d.nx, SX);
    hipLaunchKernelGGL(vi_box_kernel, dim3((unsigned)vi_box_count(d)), dim3(SX * SY), 0, s, d, (ViBox*)boxes, ntx);
    TOMO_HIP(hipGetLastError());
    return TOMO_OK;
}

hipError_t launch_batch_vi(BatchTable* t, hipStream_t s) {
    hipLaunchKernelGGL(vi_batch_kernel, dim3(kBatchGrid), dim3(SX * SY), 0, s, t, t->next);
    return hipGetLastError();
}

int launch_vi_adjoint(const tomo_geom* g, Buf vol, MutBuf sino, hipStream_t s) {
    if (sino.dtype != TOMO_F32) return fail(TOMO_ERR_UNSUPPORTED, "adjoint output must be F32");
    const DevGeom& d = g->dev;
    int rc = launch_fill((float*)sino.p, 0.0f, g->nsino(), s);
    if (rc) return rc;
    dim3 grid(ceil_div(d.nx, TX), ceil_div(d.ny, TY), ceil_div(d.nz, ZB));
    const float inv = 1.0f / vol.scale;
    TOMO_DISPATCH_STORAGE(vol.dtype, TV, {
        hipLaunchKernelGGL((vi_adjoint_kernel<TV>), grid, dim3(TX * TY), 0, s, d, (const TV*)vol.p, inv,
                           (float*)sino.p);
    });
    TOMO_HIP(hipGetLastError());
    return TOMO_OK;
}

}  // namespace tomo
