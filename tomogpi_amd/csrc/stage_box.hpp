// stage_box.hpp - LDS staging of the voxels a projector workgroup's rays
// read over a window of slabs, shared by the RS and Siddon projectors.
//
// A workgroup is 32 x 8 detector cells of one view whose rays share the
// dominant in-plane axis A.  For a window of consecutive A planes the
// workgroup copies the box of voxels (planes x z rows x B columns) its rays
// can touch into LDS, laid out [plane][z row][B column], and reads its taps
// from there.  The box comes from the workgroup's four corner rays: at an A
// plane a, the B of a ray is monotone in its detector column and its z in
// its column and row, so over a window the extremes of B and z are those of
// the corner rays at the window's two end planes.
//
// The copy goes straight from memory into LDS in 16-byte chunks of four B
// columns (buffer_load_dwordx4 ... lds, which writes lane l's chunk at
// 16 l bytes past the instruction's LDS address): a staged row holds the
// box's columns rounded out to whole chunks, rows follow each other with
// no gap, and one instruction copies 64 consecutive chunks of the buffer,
// whatever rows they belong to.  The rows come from a copy of the volume
// whose rows run along B and are padded with zeros to a multiple of four
// columns (A = x: the y-fastest copy; A = y: the volume itself when nx is a
// multiple of four).  A chunk whose row lies off the grid, or whose columns
// lie past the padded row, reads through an offset past the buffer
// resource's end, so its taps are 0, the border value.
#pragma once

#include "device_common.hpp"

namespace tomo {

// A staging buffer: kSlotChunks 16-byte chunks.  Two of them (32256 bytes)
// and the workgroup's small tables fit in 32 KB, five workgroups per CU.
constexpr int kSlotChunks = 1008;
constexpr int kSlotFloats = 4 * kSlotChunks;
constexpr int kMaxChunks = 17;  // chunks per staged row at most (68 columns)
// A byte offset that reads zeros through any buffer resource of the copy (at
// most this many bytes), also for 16-byte loads.
constexpr uint32_t kStageOutside = 0xfffffff0u;

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
    int blo, bhi, zlo, zhi;  // columns [blo, bhi], rows [zlo, zhi]
    bool ok;                 // the corner rays' positions are finite
};

// Box of the window whose rays are read between the A planes a0 and a1 (the
// corner rays at both planes; every lane computes the same values).  B and z
// bounds are floor(extreme + shift -+ margin) widened by lo_pad below and
// hi_pad above (RS: shift 0, taps floor and floor + 1; Siddon: shift 1/2,
// the voxel round(B) and its neighbours).  The margin, 1/32 voxel, is far
// above the FP32 error of the positions the kernels compute along the rays
// (below 1e-3 voxel).
__device__ __forceinline__ StageBox stage_box(const WgCorners& w, float a0, float a1, float shift, int lo_pad,
                                              int hi_pad) {
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
        b.ok = true;
    }
    return b;
}

// A staged window: planes [lo, lo + planes), z rows [zlo, zlo + nzr),
// columns [blo4, blo4 + 4 nch) with blo4 a multiple of four.  Tap (a, z, b)
// is float ((a - lo) nzr + z - zlo) (4 nch) + b - blo4 of the buffer.
// Every field is wave-uniform.
struct StageWin {
    int lo, planes, zlo, nzr, blo4, nch;
    bool fits;
};

// z rows a window may span so that stage_copy's byte offsets from the
// window's first row, below (rows nA + planes) nBp 4, stay below
// kStageOutside (planes < 16).
__device__ __forceinline__ int stage_max_rows(int nA, int nBp) {
    const uint64_t per_row = (uint64_t)nA * (uint64_t)nBp * 4u;
    const uint64_t m = ((uint64_t)kStageOutside - 16u * (uint64_t)nBp * 4u) / per_row;
    return (int)(m < (uint64_t)kSlotChunks ? m : (uint64_t)kSlotChunks);
}

// The window of `planes` planes from lo with box b; fits when its chunks fit
// a buffer, its rows are at most kMaxChunks wide and it spans at most
// max_nzr z rows.
__device__ __forceinline__ StageWin stage_window(const StageBox& b, int lo, int planes, int max_nzr) {
    StageWin w;
    w.lo = lo;
    w.planes = planes;
    w.zlo = b.zlo;
    w.nzr = b.zhi - b.zlo + 1;
    w.blo4 = b.blo & ~3;  // rounds down, also below 0
    w.nch = ((b.bhi - w.blo4) >> 2) + 1;
    w.fits = b.ok && w.nch <= kMaxChunks && w.nzr <= max_nzr && planes * w.nzr * w.nch <= kSlotChunks;
    return w;
}

// Copies window w into buf: chunk q = (row, c) of the buffer, row =
// plane nzr + z row, is chunk c of that row; instruction i of the window
// copies chunks 64 i .. 64 i + 63, and wave wv of the workgroup's 4 issues
// instructions wv, wv + 4, ....  src (src_elems floats): rows of nBp floats
// along B (nBp a multiple of four, columns [nB, nBp) zero), row (a, z) at
// (z nA + a) nBp.  One buffer resource covers the window's rows from its
// first row in the grid.  Every argument but lane is wave-uniform.  The
// caller waits for the copies (s_waitcnt) and synchronises the workgroup
// before reading.
__device__ __forceinline__ void stage_copy(const float* __restrict__ src, size_t src_elems, int nA, int nBp, int nz,
                                           const StageWin& w, float* __restrict__ buf, int lane, int wv) {
    const int zc = max(w.zlo, 0), ac = max(w.lo, 0);
    const size_t first = ((size_t)zc * (size_t)nA + (size_t)ac) * (size_t)nBp;
    const size_t left = first < src_elems ? (src_elems - first) * 4 : 0;
    const __amdgpu_buffer_rsrc_t rr = buffer_rsrc(src + (first < src_elems ? first : 0),
                                                  (uint32_t)(left < kStageOutside ? left : kStageOutside));
    const int total = w.planes * w.nzr * w.nch;
    // floor(q / d) = (int)((q + 1/2) (1/d)) for q < 2^11: with 1/d and the
    // product correctly rounded the result is within 2^-12 / d of
    // (q + 1/2) / d, which lies 1/(2d) or more from an integer
    const float rc = 1.0f / (float)w.nch, rz = 1.0f / (float)w.nzr;
    const uint32_t rowb = (uint32_t)nBp * 4u;
    for (int i = wv; i * 64 < total; i += 4) {
        const int q = i * 64 + lane;
        const int row = (int)(((float)q + 0.5f) * rc);
        const int c = q - row * w.nch;
        const int pl = (int)(((float)row + 0.5f) * rz);
        const int zi = row - pl * w.nzr;
        const int z = w.zlo + zi, a = w.lo + pl, col = w.blo4 + 4 * c;
        const bool ok = (unsigned)z < (unsigned)nz && (unsigned)a < (unsigned)nA && (unsigned)col < (unsigned)nBp;
        const uint32_t rowi = (uint32_t)(z - zc) * (uint32_t)nA + (uint32_t)(a - ac);
        const uint32_t off = ok ? rowi * rowb + (uint32_t)col * 4u : kStageOutside;
        if (q < total)
            __builtin_amdgcn_raw_ptr_buffer_load_lds(rr, (__attribute__((address_space(3))) void*)(buf + i * 256), 16,
                                                     off, 0, 0, 0);
    }
}

}  // namespace tomo
