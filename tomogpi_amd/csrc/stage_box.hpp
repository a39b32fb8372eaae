// stage_box.hpp - LDS staging of the voxels a projector workgroup's rays
// read over a window of slabs, shared by the RS and Siddon projectors.
//
// A workgroup is 32 x 8 detector cells of one view whose rays share the
// dominant in-plane axis A.  For a window of consecutive A planes the
// workgroup copies the box of voxels (planes x z rows x B columns) its rays
// can touch into LDS, laid out [plane][z row][B column] with kStagePitchP
// floats per row, and reads its taps from there.  The box comes from the
// workgroup's four corner rays: at an A plane a, the B of a ray is monotone
// in its detector column and its z in its column and row, so over a window
// the extremes of B and z are those of the corner rays at the window's two
// end planes.  Rows are copied straight from memory into LDS, one
// buffer_load_dword ... lds per row and wave: along x from the volume when
// A = y, along y from its y-fastest copy when A = x.  A row off the grid
// reads through a buffer resource of no bytes and a column off the grid
// past the row's end, so those taps are 0, the border value.
#pragma once

#include "device_common.hpp"

namespace tomo {

constexpr int kStageCols = 64;   // B columns per staged row (one lane each)
constexpr int kStagePitchP = 68;  // floats per staged row: 64 columns and 4 of padding
constexpr int kStageRows = 72;    // rows per staging buffer (planes x z rows)
constexpr int kStageBufF = kStageRows * kStagePitchP;  // floats per buffer; two are 39168 bytes

// A float's wave-uniform copy (the builtin converts to int, so the bits go
// through it).
__device__ __forceinline__ float uniform_f(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, v)));
}

// The rays of the workgroup's four corner cells: at the A plane a (voxel
// index units), ray c has B = sB + (a - sA) kB[c] and z = sz + (a - sA) kz[c]
// (one source, so sA, sB do not depend on c).  Kept as the ranges of the
// slopes over the corners, wave-uniform (every lane computes the same).
struct WgCorners {
    float sA, sB, sz;
    float kBlo, kBhi, kzlo, kzhi;
};
__device__ __forceinline__ WgCorners wg_corners(const DevGeom& g, int A, int p, int mlo, int mhi, int nlo, int nhi) {
    const double* vw = g.view + (size_t)p * kViewStride;
    const double ca = vw[0], sa = vw[1], dl = (double)g.delta;
    const double Sx = -g.fod * ca / dl + (double)g.cx, Sy = -g.fod * sa / dl + (double)g.cy;
    WgCorners w;
    w.sA = uniform_f((float)(A == 0 ? Sx : Sy));
    w.sB = uniform_f((float)(A == 0 ? Sy : Sx));
    w.sz = uniform_f(g.cz);
    float kBlo = 3.0e38f, kBhi = -3.0e38f, kzlo = 3.0e38f, kzhi = -3.0e38f;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const int m = (c & 1) ? mhi : mlo, n = (c & 2) ? nhi : nlo;
        const double s = ((double)m - (double)g.cu) * (double)g.du;
        const double t = ((double)n - (double)g.cv) * (double)g.dv;
        const double rx = ca * g.fdd - s * sa, ry = sa * g.fdd + s * ca, rz = t;
        const float kB = (float)(A == 0 ? ry / rx : rx / ry), kz = (float)(A == 0 ? rz / rx : rz / ry);
        kBlo = fminf(kBlo, kB);
        kBhi = fmaxf(kBhi, kB);
        kzlo = fminf(kzlo, kz);
        kzhi = fmaxf(kzhi, kz);
    }
    w.kBlo = uniform_f(kBlo);
    w.kBhi = uniform_f(kBhi);
    w.kzlo = uniform_f(kzlo);
    w.kzhi = uniform_f(kzhi);
    return w;
}

struct StageBox {
    int blo, bhi, zlo, zhi;  // staged columns [blo, blo + 63], rows [zlo, zhi]
    bool fits;
};

// Box of the window whose rays are read between the A planes a0 and a1 (the
// corner rays at both planes; every lane computes the same values).  B and z
// bounds are floor(extreme + shift -+ margin) widened by lo_pad below and
// hi_pad above (RS: shift 0, taps floor and floor + 1; Siddon: shift 1/2,
// the voxel round(B) and its neighbours).  The margin, 1/32 voxel, is far
// above the FP32 error of the positions the kernels compute along the rays
// (below 1e-3 voxel).
__device__ __forceinline__ StageBox stage_box(const WgCorners& w, float a0, float a1, float shift, int lo_pad,
                                              int hi_pad, int max_rows) {
    constexpr float kMargin = 1.0f / 32.0f;
    // each bound is affine in the slope, so its extremes over the corners are at the slopes' extremes
    const float d0 = a0 - w.sA, d1 = a1 - w.sA;
    const float b00 = fmaf(d0, w.kBlo, w.sB), b01 = fmaf(d0, w.kBhi, w.sB);
    const float b10 = fmaf(d1, w.kBlo, w.sB), b11 = fmaf(d1, w.kBhi, w.sB);
    const float z00 = fmaf(d0, w.kzlo, w.sz), z01 = fmaf(d0, w.kzhi, w.sz);
    const float z10 = fmaf(d1, w.kzlo, w.sz), z11 = fmaf(d1, w.kzhi, w.sz);
    const float bmn = fminf(fminf(b00, b01), fminf(b10, b11)), bmx = fmaxf(fmaxf(b00, b01), fmaxf(b10, b11));
    const float zmn = fminf(fminf(z00, z01), fminf(z10, z11)), zmx = fmaxf(fmaxf(z00, z01), fmaxf(z10, z11));
    StageBox b{0, -1, 0, -1, false};
    const bool finite = bmn > -1.0e6f && bmx < 1.0e6f && zmn > -1.0e6f && zmx < 1.0e6f;
    if (finite) {
        b.blo = __builtin_amdgcn_readfirstlane((int)floorf(bmn + shift - kMargin) - lo_pad);
        b.bhi = __builtin_amdgcn_readfirstlane((int)floorf(bmx + shift + kMargin) + hi_pad);
        b.zlo = __builtin_amdgcn_readfirstlane((int)floorf(zmn + shift - kMargin) - lo_pad);
        b.zhi = __builtin_amdgcn_readfirstlane((int)floorf(zmx + shift + kMargin) + hi_pad);
        b.fits = b.bhi - b.blo < kStageCols && b.zhi - b.zlo < max_rows;
    }
    return b;
}

// Copies planes [lo, lo + planes) x rows [zlo, zlo + nzr) x columns
// [blo, blo + 63] into the buffer as rows (plane, z row) = pi nzr + zi,
// kStagePitchP floats apart (planes * nzr <= kStageRows); wave wv of the
// workgroup's 4 copies the rows zi = wv, wv + 4, ... of every plane.  src
// (src_elems floats): rows along B (A = 0: the y-fastest copy, element
// (y, x, z) at y + ny (x + nx z); A = 1: the volume).  One buffer resource
// covers the window's rows from its first row in the grid; a lane's byte
// offset is its row's plus its column's, added with unsigned saturation, and
// a column off the grid has an offset of 0xffffffff, past the resource's end,
// so it reads 0.  When every row of the window is in the grid (the usual
// case) the row loop only steps two offsets; otherwise a row off the grid
// also gets the offset 0xffffffff.  Every argument but lane is wave-uniform.
// The caller waits for the copies (s_waitcnt) and synchronises the workgroup
// before reading.
__device__ __forceinline__ void stage_rows(const float* __restrict__ src, size_t src_elems, int nA, int nB, int nz,
                                           int lo, int planes, int zlo, int nzr, int blo, float* __restrict__ stage,
                                           int lane, int wv) {
    const int zc = max(zlo, 0), ac = max(lo, 0);
    const size_t first = ((size_t)zc * (size_t)nA + (size_t)ac) * (size_t)nB;
    const size_t left = first < src_elems ? (src_elems - first) * 4 : 0;
    const __amdgpu_buffer_rsrc_t rr = buffer_rsrc(src + (first < src_elems ? first : 0),
                                                  (uint32_t)(left < 0xfffffff0ull ? left : 0xfffffff0ull));
    const int col = blo + lane;
    const uint32_t coff = (unsigned)col < (unsigned)nB ? (uint32_t)col * 4u : 0xffffffffu;
    const int64_t rstep = (int64_t)nA * nB * 4;  // bytes per z row
    const uint32_t step4 = (uint32_t)(4 * rstep);
    const bool inside = zlo >= 0 && zlo + nzr <= nz && lo >= 0 && lo + planes <= nA;
    if (inside) {
        // rows (lo + pi, zlo + zi) at byte offset (zi nA + pi) nB 4 from the resource's base
        for (int pi = 0; pi < planes; ++pi) {
            uint32_t ro = (uint32_t)(((int64_t)wv * nA + pi) * nB * 4);
            float* dst = stage + (pi * nzr + wv) * kStagePitchP;
            for (int zi = wv; zi < nzr; zi += 4) {
                __builtin_amdgcn_raw_ptr_buffer_load_lds(rr, (__attribute__((address_space(3))) void*)dst, 4,
                                                         __builtin_elementwise_add_sat(ro, coff), 0, 0, 0);
                ro += step4;
                dst += 4 * kStagePitchP;
            }
        }
        return;
    }
    for (int pi = 0; pi < planes; ++pi) {
        const int a = lo + pi;
        int64_t roff = ((int64_t)(zlo + wv - zc) * nA + (a - ac)) * (int64_t)nB * 4;
        float* dst = stage + (pi * nzr + wv) * kStagePitchP;
        for (int zi = wv; zi < nzr; zi += 4) {
            const bool in = (unsigned)a < (unsigned)nA && (unsigned)(zlo + zi) < (unsigned)nz;
            const uint32_t ro = in ? (uint32_t)roff : 0xffffffffu;
            __builtin_amdgcn_raw_ptr_buffer_load_lds(rr, (__attribute__((address_space(3))) void*)dst, 4,
                                                     __builtin_elementwise_add_sat(ro, coff), 0, 0, 0);
            roff += 4 * rstep;
            dst += 4 * kStagePitchP;
        }
    }
}

}  // namespace tomo
