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
// thread owns a column of ZB voxels that share u and A; the per-(column,view)
// terms w, 1/w, u and A are evaluated in FP64 (the geometry of the FP32 form)
// and the per-voxel v, bilinear weights and sum in FP32.
#include <algorithm>
#include <climits>
#include <cmath>
#include <type_traits>

#include "device_common.hpp"
#include "batch.hpp"

namespace tomo {
namespace {

constexpr int TX = kViTileX, TY = kViTileY, ZB = kViTileZ;

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

// Tap-quad copy of an FP32 sinogram: element (p, vv, uu), vv in [0, nv+2],
// uu in [0, nu+3], holds the detector values (u, v) and (u, v+1) with
// u = uu-2, v = vv-2 (0 off the detector).  Two neighbouring elements are the
// four bilinear taps of the cell with lower corner (u, v), so one 16-byte
// load serves one voxel-view update.  A 16-bit sinogram gets the tap-column
// copy instead: element (p, vv, uu) holds (u, v..v+3), and two neighbouring
// elements (16 bytes) are the taps of columns u, u+1 on rows v..v+3, enough
// for two voxels of a column whose rows differ by at most 2.  Cells with a
// corner in [-2, nu] x [-2, nv] are in the copies; a cell whose corner is
// outside [-1, n-1] on either axis has no tap on the detector, and clamping
// its corner into [-2, n] lands on another such cell, all zeros, which
// replaces the per-tap range tests.  Values are copied as stored.
struct QuadShape {
    uint32_t row;        // elements per row of the copy: per (nu + 4)
    uint32_t view;       // elements per view: row (nv + 3)
};
__host__ __device__ __forceinline__ QuadShape quad_shape(int nu, int nv, uint32_t per) {
    QuadShape q;
    q.row = per * (uint32_t)(nu + 4);
    q.view = q.row * (uint32_t)(nv + 3);
    return q;
}

template <typename TS>
__device__ __forceinline__ void vi_quads_tile(const TS* __restrict__ in, TS* __restrict__ out, int nu, int nv, int bx,
                                              int by, int p) {
    const int uu = bx * 64 + (threadIdx.x & 63);
    const int vv = by * 4 + (threadIdx.x >> 6);
    if (uu > nu + 3 || vv > nv + 2) return;
    const int u = uu - 2, v = vv - 2;
    const bool uok = u >= 0 && u < nu;
    const TS* gp = in + (size_t)p * nu * nv;
    TS a, b;
    __builtin_memset(&a, 0, sizeof(a));
    __builtin_memset(&b, 0, sizeof(b));
    if (uok && v >= 0 && v < nv) a = gp[(size_t)v * nu + u];
    if (uok && v + 1 >= 0 && v + 1 < nv) b = gp[(size_t)(v + 1) * nu + u];
    const QuadShape qs = quad_shape(nu, nv, 2);
    const size_t o = (size_t)p * qs.view + (size_t)vv * qs.row + 2u * (uint32_t)uu;
    out[o] = a;
    out[o + 1] = b;
}

// Tap-column copy of a 16-bit sinogram (same grid as the tap quads).
__device__ __forceinline__ void vi_columns_tile(const uint16_t* __restrict__ in, uint16_t* __restrict__ out, int nu,
                                                int nv, int bx, int by, int p) {
    const int uu = bx * 64 + (threadIdx.x & 63);
    const int vv = by * 4 + (threadIdx.x >> 6);
    if (uu > nu + 3 || vv > nv + 2) return;
    const int u = uu - 2, v = vv - 2;
    const bool uok = u >= 0 && u < nu;
    const uint16_t* gp = in + (size_t)p * nu * nv;
    uint32_t t[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) t[r] = (uok && v + r >= 0 && v + r < nv) ? gp[(size_t)(v + r) * nu + u] : 0u;
    const QuadShape qs = quad_shape(nu, nv, 4);
    const uint2 w = make_uint2(t[0] | (t[1] << 16), t[2] | (t[3] << 16));
    __builtin_memcpy(out + (size_t)p * qs.view + (size_t)vv * qs.row + 4u * (uint32_t)uu, &w, sizeof(w));
}

// Raw taps of a later voxel of a group from the group's tap-column load:
// rows d, d+1 of columns u (words c0 = rows 0,1 and c1 = rows 2,3) and u+1
// (e0, e1), d in {0, 1, 2} its clamped row less the first voxel's.  When the
// voxels' v differ by less than 2 (A below 1.99 for neighbours, below 0.995
// two apart), their floors (and the clamped rows) differ by 0, 1 or 2.
__device__ __forceinline__ void pair_rows(float d, uint32_t c0, uint32_t c1, uint32_t e0, uint32_t e1, uint32_t& b0,
                                          uint32_t& b1) {
    const uint32_t cm = __builtin_amdgcn_alignbyte(c1, c0, 2), em = __builtin_amdgcn_alignbyte(e1, e0, 2);
    b0 = d == 0.0f ? c0 : (d == 1.0f ? cm : c1);
    b1 = d == 0.0f ? e0 : (d == 1.0f ? em : e1);
}

// Four words of a tap-column load (the non-buffer path's stand-in for the
// buffer load's vector).
struct w4 {
    uint32_t v[4];
    __device__ __forceinline__ uint32_t operator[](int i) const { return v[i]; }
};

// The 16 voxels of a column for one view on 16-bit tap columns, GROUP
// voxels per load: the group's first voxel reads rows 0,1 of the load, the
// others rows d, d+1 with d their clamped row less the first's (0..2).
// col4 / col3 load 4 / 3 words at a clamped row; term16 adds one term.
template <int GROUP, typename C4, typename C3, typename T16, typename CV>
__device__ __forceinline__ void vi_column16(const C4& col4, const C3& col3, const T16& term16, const CV& cvw,
                                            __half2 fu2, float fu, float zp0, float cv, float fnv) {
    constexpr int kFull = GROUP == 1 ? 0 : ZB / GROUP * GROUP;
#pragma unroll
    for (int k = 0; k < kFull; k += GROUP) {
        const float va = fmaf(zp0 + (float)k, cvw.A, cv);
        const float vfa = floorf(va);
        const float vca = __builtin_amdgcn_fmed3f(vfa, -2.0f, fnv);
        const auto w = col4(vca);
        term16(cvw, fu2, fu, k, va, vfa, w[0], w[2]);
#pragma unroll
        for (int j = 1; j < GROUP; ++j) {
            const float vb = fmaf(zp0 + (float)(k + j), cvw.A, cv);
            const float vfb = floorf(vb);
            const float vcb = __builtin_amdgcn_fmed3f(vfb, -2.0f, fnv);
            uint32_t b0, b1;
            pair_rows(vcb - vca, w[0], w[1], w[2], w[3], b0, b1);
            term16(cvw, fu2, fu, k + j, vb, vfb, b0, b1);
        }
    }
#pragma unroll
    for (int k = kFull; k < ZB; ++k) {
        const float v = fmaf(zp0 + (float)k, cvw.A, cv);
        const float vf = floorf(v);
        const auto w = col3(__builtin_amdgcn_fmed3f(vf, -2.0f, fnv));
        term16(cvw, fu2, fu, k, v, vf, w[0], w[2]);
    }
}

// Forward VI on the tap-quad copy (FP32) or the tap-column copy (16-bit).  Bilinear interpolation at (u, v): the
// taps (u,v), (u,v+1), (u+1,v), (u+1,v+1) of the cell with lower corner
// (floor u, floor v), 0 off the detector, interpolated along u, then v, with
// explicit fmaf.  The cell corner is clamped into the copy (a cell off the
// detector reads four zeros and adds +0).  FAST (copy below 4 GiB, one view
// below 16 MiB): buffer loads at a byte offset computed as one exact FP32
// fma of small integers; otherwise element indices in integers.
// VI stores delta * sum; FDK weights each term by W = num / sqrt(fma(v', v',
// base)), v' = z' A, and stores the sum itself.  16-bit storage reads the
// taps of GROUP consecutive voxels of the column with one tap-column load:
// GROUP 2 when the host has shown A in [0, 1.99] for every column and view
// (rows of voxels k, k+1 differ by at most 2), GROUP 3 when A is in
// [0, 0.995] (rows of k, k+2 differ by at most 2), else 1; the voxels left
// after the last full group load alone.  The taps and the arithmetic of each
// term are the same for every GROUP.
template <typename TS, typename TV, bool FDK, int GROUP = 1>
__device__ __forceinline__ void vi_tile(const DevGeom& g, const TS* __restrict__ quads, float sino_inv,
                                        TV* __restrict__ vol, float vol_scale, float post, int bx, int by, int bz) {
    const int i = bx * TX + (threadIdx.x % TX);
    const int j = by * TY + (threadIdx.x / TX);
    const int k0 = bz * ZB;
    if (i >= g.nx || j >= g.ny) return;
    const double xp = (double)i - (double)g.cx, yp = (double)j - (double)g.cy;
    const float zp0 = (float)((double)k0 - (double)g.cz);
    const float cv = g.cv;
    float acc[ZB];
#pragma unroll
    for (int k = 0; k < ZB; ++k) acc[k] = 0.0f;
    constexpr bool kF16 = std::is_same<TS, __half>::value;
    constexpr bool kBF16 = std::is_same<TS, __hip_bfloat16>::value;
    constexpr uint32_t kPer = (kF16 || kBF16) ? 4u : 2u;  // values per element: tap columns / tap quads
    const QuadShape qs = quad_shape(g.nu, g.nv, kPer);
    const float fnu = (float)g.nu, fnv = (float)g.nv;
    constexpr uint32_t es = sizeof(TS);
    const uint64_t view_bytes = (uint64_t)qs.view * es;
    const bool fast = view_bytes < (1ull << 24) && view_bytes * (uint64_t)g.nphi < (1ull << 32);
    // one voxel-view term from its interpolated value r
    auto add = [&](const ColumnView& cvw, int k, float r) {
        if (FDK) {
            const float vp = (zp0 + (float)k) * cvw.A;
            const float wt = cvw.num / sqrtf(fmaf(vp, vp, cvw.base));
            acc[k] += wt * r;
        } else {
            acc[k] += r;
        }
    };
    // FP32 taps (FP32 storage: the oracle's arithmetic)
    auto term = [&](const ColumnView& cvw, float fu, int k, float v, float vf, const float4& t) {
        const float fv = v - vf;
        const float a = fmaf(fu, t.z - t.x, t.x);
        const float b = fmaf(fu, t.w - t.y, t.y);
        add(cvw, k, fmaf(fv, b - a, a));
    };
    // 16-bit storage, raw taps (g(u,v), g(u,v+1)) in w0 and (g(u+1,v),
    // g(u+1,v+1)) in w1; the storage scale is applied once at the store.
    // FP16: both rows interpolated along u in one packed FP16 fma, the v
    // step in FP32 (error well inside the mixed-form bound; accumulation
    // stays FP32).  BF16: FP32 interpolation of the widened taps.
    auto term16 = [&](const ColumnView& cvw, __half2 fu2, float fu, int k, float v, float vf, uint32_t w0,
                      uint32_t w1) {
        const float fv = v - vf;
        if constexpr (kF16) {
            const __half2 p0 = __builtin_bit_cast(__half2, w0), p1 = __builtin_bit_cast(__half2, w1);
            const __half2 ab = __hfma2(fu2, __hsub2(p1, p0), p0);
            const __half a = __low2half(ab);
            add(cvw, k, fmaf(fv, __half2float(__hsub(__high2half(ab), a)), __half2float(a)));
        } else {
            const float t00 = __uint_as_float(w0 << 16), t10 = __uint_as_float(w0 & 0xffff0000u);
            const float t01 = __uint_as_float(w1 << 16), t11 = __uint_as_float(w1 & 0xffff0000u);
            const float a = fmaf(fu, t01 - t00, t00);
            const float b = fmaf(fu, t11 - t10, t10);
            add(cvw, k, fmaf(fv, b - a, a));
        }
    };
    if (fast) {
        const __amdgpu_buffer_rsrc_t rsrc = buffer_rsrc(quads, (uint32_t)(view_bytes * (uint64_t)g.nphi));
        const float row_b = (float)(qs.row * es);
        for (int p = 0; p < g.nphi; ++p) {
            const ColumnView cvw = column_view<FDK>(g, g.view + (size_t)p * kViewStride, xp, yp);
            const float uf = floorf(cvw.u);
            const float fu = cvw.u - uf;
            const float uc = __builtin_amdgcn_fmed3f(uf, -2.0f, fnu);
            // byte offset of cell (uc, vc): ((vc + 2) row + per (uc + 2)) es
            const float base = fmaf(uc + 2.0f, (float)(kPer * es), 2.0f * row_b);
            const uint32_t soff = (uint32_t)p * (uint32_t)view_bytes;
            if constexpr (kF16 || kBF16) {
                const __half2 fu2 = __float2half2_rn(fu);
                const auto col4 = [&](float vc) {
                    return __builtin_amdgcn_raw_buffer_load_b128(rsrc, (int)(uint32_t)fmaf(vc, row_b, base),
                                                                 (int)soff, 0);
                };
                const auto col3 = [&](float vc) {
                    return __builtin_amdgcn_raw_buffer_load_b96(rsrc, (int)(uint32_t)fmaf(vc, row_b, base),
                                                                (int)soff, 0);
                };
                vi_column16<GROUP>(col4, col3, term16, cvw, fu2, fu, zp0, cv, fnv);
            } else {
#pragma unroll
                for (int k = 0; k < ZB; ++k) {
                    const float v = fmaf(zp0 + (float)k, cvw.A, cv);
                    const float vf = floorf(v);
                    const float vc = __builtin_amdgcn_fmed3f(vf, -2.0f, fnv);
                    const uint32_t off = (uint32_t)fmaf(vc, row_b, base);
                    term(cvw, fu, k, v, vf, BufLoad4<TS>::load(rsrc, off, soff, sino_inv));
                }
            }
        }
    } else {
        for (int p = 0; p < g.nphi; ++p) {
            const ColumnView cvw = column_view<FDK>(g, g.view + (size_t)p * kViewStride, xp, yp);
            const float uf = floorf(cvw.u);
            const float fu = cvw.u - uf;
            const int uc = (int)__builtin_amdgcn_fmed3f(uf, -2.0f, fnu);
            const TS* gp = quads + (size_t)p * qs.view + kPer * (uint32_t)(uc + 2);
            if constexpr (kF16 || kBF16) {
                const __half2 fu2 = __float2half2_rn(fu);
                const auto col4 = [&](float vc) {
                    uint32_t w[4];
                    __builtin_memcpy(w, gp + (uint64_t)(uint32_t)((int)vc + 2) * qs.row, sizeof(w));
                    return w4{w[0], w[1], w[2], w[3]};
                };
                const auto col3 = [&](float vc) {
                    uint32_t w[3];
                    __builtin_memcpy(w, gp + (uint64_t)(uint32_t)((int)vc + 2) * qs.row, sizeof(w));
                    return w4{w[0], w[1], w[2], 0u};
                };
                vi_column16<GROUP>(col4, col3, term16, cvw, fu2, fu, zp0, cv, fnv);
            } else {
#pragma unroll
                for (int k = 0; k < ZB; ++k) {
                    const float v = fmaf(zp0 + (float)k, cvw.A, cv);
                    const float vf = floorf(v);
                    const int vc = (int)__builtin_amdgcn_fmed3f(vf, -2.0f, fnv);
                    const uint64_t e = (uint64_t)(uint32_t)(vc + 2) * qs.row;
                    term(cvw, fu, k, v, vf, Storage<TS>::load4(gp, e, sino_inv));
                }
            }
        }
    }
    const size_t sxy = (size_t)g.nx * g.ny;
    const size_t base = (size_t)i + (size_t)g.nx * j;
    const float in_scale = (kF16 || kBF16) ? sino_inv : 1.0f;  // 16-bit taps are read unscaled
#pragma unroll
    for (int k = 0; k < ZB; ++k) {
        const float a = (kF16 || kBF16) ? acc[k] * in_scale : acc[k];
        if (k0 + k < g.nz)
            Storage<TV>::store(vol, base + (size_t)(k0 + k) * sxy, FDK ? a : (a * g.delta) * post, vol_scale);
    }
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
constexpr int SX = kViStageX, SY = kViStageY;  // voxel columns per workgroup; a wave is 32 x 2 columns
constexpr int kStagePitch = 65;  // floats per staged row (odd: rows of one column fall in different banks)

template <int ZS>
struct StageShape {
    static constexpr int rows = ZS == 32 ? 64 : 44;  // staged rows per view
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

template <int ZS>
__device__ __forceinline__ StagePlan stage_plan(const ViBox& bb, float zp0, float cv) {
    const float zp1 = zp0 + (float)(ZS - 1);
    const float va = fmaf(zp0, bb.amin, cv), vb = fmaf(zp0, bb.amax, cv);
    const float vc = fmaf(zp1, bb.amin, cv), vd = fmaf(zp1, bb.amax, cv);
    const float vlo = floorf(fminf(fminf(va, vb), fminf(vc, vd)));
    const float vhi = floorf(fmaxf(fmaxf(va, vb), fmaxf(vc, vd)));
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

template <typename TV, bool FDK, int ZS>
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
    const int i = bx * SX + (tid % SX);
    const int j = by * SY + (tid / SX);
    const int k0 = bz * ZS;
    const bool active = i < g.nx && j < g.ny;
    const double xp = (double)i - (double)g.cx, yp = (double)j - (double)g.cy;
    const float zp0 = (float)((double)k0 - (double)g.cz);
    const float cv = g.cv;
    const float fnu = (float)g.nu;
    const int nu = g.nu, nv = g.nv;
    const size_t plane = (size_t)nu * nv;
    const ViBox* bt = boxes + (size_t)(by * ((g.nx + SX - 1) / SX) + bx) * g.nphi;
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
    // A tap off the detector or beyond the box is read at an offset past the
    // view's end, which writes 0.
    const uint32_t view_bytes = (uint32_t)(plane * sizeof(float));
    auto copy_box = [&](const StagePlan& s, int p, int b) {
        if (!s.fits) return;
        const __amdgpu_buffer_rsrc_t rsrc = buffer_rsrc(sino + (size_t)p * plane, view_bytes);
        const int u = s.u0 + lane;
        const bool uin = lane < s.cols && (unsigned)u < (unsigned)nu;
        const int vw = s.v0 + wv;
        const uint32_t base = ((uint32_t)vw * (uint32_t)nu + (uint32_t)u) * 4u;  // mod 2^32, exact when in range
        float* dst = stage + b * SS::taps + wv * kStagePitch;
#pragma unroll
        for (int q = 0; q < kSlots; ++q) {
            if (wv + q * NW < s.rows) {
                const bool in = uin && (unsigned)(vw + q * NW) < (unsigned)nv;
                const uint32_t off = in ? base + (uint32_t)(q * NW) * (uint32_t)nu * 4u : 0xfffffff0u;
                __builtin_amdgcn_raw_ptr_buffer_load_lds(
                    rsrc, (__attribute__((address_space(3))) void*)(dst + q * NW * kStagePitch), 4, off, 0, 0, 0);
            }
        }
    };
    StagePlan nxt = stage_plan<ZS>(bt[0], zp0, cv);
    copy_box(nxt, 0, 0);
    for (int p = 0; p < g.nphi; ++p) {
        // this wave's copies of view p have landed; every wave is done with view p - 1
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        const StagePlan now = nxt;
        if (p + 1 < g.nphi) {
            nxt = stage_plan<ZS>(bt[p + 1], zp0, cv);
            copy_box(nxt, p + 1, (p + 1) & 1);
        }
        if (active) {
            const ColumnView cvw = column_view<FDK>(g, g.view + (size_t)p * kViewStride, xp, yp);
            const float uf = floorf(cvw.u);
            const float fu = cvw.u - uf;
            const float uc = __builtin_amdgcn_fmed3f(uf, -2.0f, fnu);
            if (now.fits) {
                // byte offset in the buffer of tap (uc, vf): ((vf - v0) pitch + uc - u0) 4,
                // one exact fma of small integers
                const float colb =
                    (float)(((int)uc - now.u0 - now.v0 * kStagePitch + (p & 1) * SS::taps) * 4);
#pragma unroll
                for (int k = 0; k < ZS; ++k) {
                    const float v = fmaf(zp0 + (float)k, cvw.A, cv);
                    const float vf = floorf(v);
                    const uint32_t off = (uint32_t)fmaf(vf, (float)(kStagePitch * 4), colb);
                    const float* q = (const float*)((const char*)stage + off);
                    term(cvw, fu, k, v, vf, q[0], q[kStagePitch], q[1], q[kStagePitch + 1]);
                }
            } else {
                // box larger than the buffer: the four taps straight from the view, one voxel at a time
                const __amdgpu_buffer_rsrc_t rsrc = buffer_rsrc(sino + (size_t)p * plane, view_bytes);
                const int iu = (int)uc;
                const bool u0ok = (unsigned)iu < (unsigned)nu, u1ok = (unsigned)(iu + 1) < (unsigned)nu;
                ColumnView cw = cvw;
                asm volatile("" : "+v"(cw.A));  // this path's rows are its own, not hoisted above the branch
                auto rd = [&](bool ok, int v, int u) {
                    const uint32_t off = ok ? (uint32_t)(v * nu + u) * 4u : 0xfffffff0u;
                    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rsrc, (int)off, 0, 0));
                };
#pragma unroll
                for (int k = 0; k < ZS; ++k) {
                    const float v = fmaf(zp0 + (float)k, cw.A, cv);
                    const float vf = floorf(v);
                    const int iv = (int)__builtin_amdgcn_fmed3f(vf, -2.0f, (float)nv);
                    const bool v0ok = (unsigned)iv < (unsigned)nv, v1ok = (unsigned)(iv + 1) < (unsigned)nv;
                    term(cw, fu, k, v, vf, rd(u0ok && v0ok, iv, iu), rd(u0ok && v1ok, iv + 1, iu),
                         rd(u1ok && v0ok, iv, iu + 1), rd(u1ok && v1ok, iv + 1, iu + 1));
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
            Storage<TV>::store(vol, base + (size_t)(k0 + k) * sxy, FDK ? acc[k] : (acc[k] * g.delta) * post,
                               vol_scale);
}

template <typename TV, bool FDK, int ZS>
__global__ __launch_bounds__(SX* SY, 6) void vi_staged_kernel(DevGeom g, const float* __restrict__ sino,
                                                         const ViBox* __restrict__ boxes, TV* __restrict__ vol,
                                                         float vol_scale) {
    __shared__ float stage[2 * StageShape<ZS>::taps];
    vi_staged_tile<TV, FDK, ZS>(g, sino, boxes, vol, vol_scale, 1.0f, blockIdx.x, blockIdx.y, blockIdx.z, stage);
}

template <typename TS>
__global__ __launch_bounds__(256) void vi_quads_kernel(const TS* __restrict__ in, TS* __restrict__ out, int nu,
                                                     int nv) {
    if constexpr (sizeof(TS) == 2)
        vi_columns_tile((const uint16_t*)in, (uint16_t*)out, nu, nv, blockIdx.x, blockIdx.y, blockIdx.z);
    else
        vi_quads_tile<TS>(in, out, nu, nv, blockIdx.x, blockIdx.y, blockIdx.z);
}

template <typename TS, typename TV, bool FDK, int GROUP>
__global__ __launch_bounds__(TX* TY) void vi_kernel(DevGeom g, const TS* __restrict__ quads, float sino_inv,
                                                  TV* __restrict__ vol, float vol_scale) {
    vi_tile<TS, TV, FDK, GROUP>(g, quads, sino_inv, vol, vol_scale, 1.0f, blockIdx.x, blockIdx.y, blockIdx.z);
}

// Shared launch (FP32 sinograms and volumes): tile grid (ceil(nx/SX),
// ceil(ny/SY), ceil(nz/16)) on the sinogram `in` with the problem's box table
// `aux` (launch_vi_boxes); the stored value is (delta * sum) * a.
__global__ __launch_bounds__(SX* SY, 6) void vi_batch_kernel(const BatchTable* __restrict__ t, uint32_t* __restrict__ next) {
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

// Voxels per 16-bit tap-column load the geometry allows: 3 when every
// column-view's row step A = delta FDD/dv / w is in [0, 0.99], 2 when it is
// in [0, 1.98], else 1 (the kernel's bounds 0.995 and 1.99 less a margin far
// above A's FP32 rounding).  w = delta (cos x' + sin y') + FOD is linear in
// (x', y'), so its minimum over the voxel centres is at a corner of their
// rectangle.
int vi_row_group(const tomo_geom* g) {
    const DevGeom& d = g->dev;
    const double xs[2] = {-(double)d.cx, (double)(d.nx - 1) - (double)d.cx};
    const double ys[2] = {-(double)d.cy, (double)(d.ny - 1) - (double)d.cy};
    double amax = 0.0;
    for (int p = 0; p < d.nphi; ++p) {
        const double* c = &g->view[(size_t)p * kViewStride];
        double wmin = 1.0e300;
        for (double x : xs)
            for (double y : ys) wmin = std::min(wmin, std::fma(c[2], x, std::fma(c[3], y, d.fod)));
        if (!(wmin > 0.0) || !(c[6] >= 0.0)) return 1;
        amax = std::max(amax, c[6] / wmin);
    }
    return amax <= 0.99 ? 3 : (amax <= 1.98 ? 2 : 1);
}

template <typename TS, typename TV, bool FDK>
static void launch_vi_kernel(const DevGeom& d, dim3 grid, int group, const void* quads, float inv, void* vol,
                             float vol_scale, hipStream_t s) {
    constexpr bool k16 = sizeof(TS) == 2;
    if (k16 && group == 3)
        hipLaunchKernelGGL((vi_kernel<TS, TV, FDK, k16 ? 3 : 1>), grid, dim3(TX * TY), 0, s, d, (const TS*)quads,
                           inv, (TV*)vol, vol_scale);
    else if (k16 && group == 2)
        hipLaunchKernelGGL((vi_kernel<TS, TV, FDK, k16 ? 2 : 1>), grid, dim3(TX * TY), 0, s, d, (const TS*)quads,
                           inv, (TV*)vol, vol_scale);
    else
        hipLaunchKernelGGL((vi_kernel<TS, TV, FDK, 1>), grid, dim3(TX * TY), 0, s, d, (const TS*)quads, inv,
                           (TV*)vol, vol_scale);
}

// Voxels per thread column of the staged kernel.
constexpr int kStagedZ = kViTileZ;

template <typename TV>
static hipError_t launch_vi_staged(const DevGeom& d, int fdk, const float* sino, void* vol, float vol_scale,
                                   hipStream_t s) {
    const int ntx = ceil_div(d.nx, SX), nty = ceil_div(d.ny, SY);
    void* boxes = nullptr;
    hipError_t e = hipMallocAsync(&boxes, vi_box_count(d) * sizeof(ViBox), s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(vi_box_kernel, dim3((unsigned)vi_box_count(d)), dim3(SX * SY), 0, s, d, (ViBox*)boxes, ntx);
    e = hipGetLastError();
    if (e == hipSuccess) {
        const dim3 grid(ntx, nty, ceil_div(d.nz, kStagedZ));
        if (fdk)
            hipLaunchKernelGGL((vi_staged_kernel<TV, true, kStagedZ>), grid, dim3(SX * SY), 0, s, d, sino,
                               (const ViBox*)boxes, (TV*)vol, vol_scale);
        else
            hipLaunchKernelGGL((vi_staged_kernel<TV, false, kStagedZ>), grid, dim3(SX * SY), 0, s, d, sino,
                               (const ViBox*)boxes, (TV*)vol, vol_scale);
        e = hipGetLastError();
    }
    const hipError_t f = hipFreeAsync(boxes, s);
    return e != hipSuccess ? e : f;
}

int launch_vi(const tomo_geom* g, Buf sino, MutBuf vol, hipStream_t s, int fdk) {
    const DevGeom& d = g->dev;
    if (sino.dtype == TOMO_F32) {
        // FP32 sinograms: LDS-staged taps read from the sinogram itself.
        hipError_t e = hipSuccess;
        TOMO_DISPATCH_STORAGE(vol.dtype, TV, { e = launch_vi_staged<TV>(d, fdk, (const float*)sino.p, vol.p,
                                                                        vol.scale, s); });
        TOMO_HIP(e);
        return TOMO_OK;
    }
    // 16-bit sinograms: tap-column copy from the stream-ordered pool,
    // released after the backprojection on the same stream.
    void* quads = nullptr;
    TOMO_HIP(hipMallocAsync(&quads, vi_quad_elems(d.nu, d.nv, d.nphi) * 4, s));
    const dim3 qgrid(ceil_div(d.nu + 4, 64), ceil_div(d.nv + 3, 4), d.nphi);
    const dim3 grid(ceil_div(d.nx, TX), ceil_div(d.ny, TY), ceil_div(d.nz, ZB));
    const float inv = 1.0f / sino.scale;
    const int group = vi_row_group(g);
    hipError_t le = hipSuccess;
    TOMO_DISPATCH_STORAGE(sino.dtype, TS, {
        hipLaunchKernelGGL((vi_quads_kernel<TS>), qgrid, dim3(256), 0, s, (const TS*)sino.p, (TS*)quads, d.nu, d.nv);
        le = hipGetLastError();
        if (le == hipSuccess) {
            TOMO_DISPATCH_STORAGE(vol.dtype, TV, {
                if (fdk)
                    launch_vi_kernel<TS, TV, true>(d, grid, group, quads, inv, vol.p, vol.scale, s);
                else
                    launch_vi_kernel<TS, TV, false>(d, grid, group, quads, inv, vol.p, vol.scale, s);
            });
            le = hipGetLastError();
        }
    });
    TOMO_HIP(hipFreeAsync(quads, s));
    TOMO_HIP(le);
    return TOMO_OK;
}

int launch_vi_boxes(const DevGeom& d, void* boxes, hipStream_t s) {
    const int ntx = ceil_div(d.nx, SX);
    hipLaunchKernelGGL(vi_box_kernel, dim3((unsigned)vi_box_count(d)), dim3(SX * SY), 0, s, d, (ViBox*)boxes, ntx);
    TOMO_HIP(hipGetLastError());
    return TOMO_OK;
}

hipError_t launch_batch_vi(BatchTable* t, hipStream_t s) {
    hipLaunchKernelGGL(vi_batch_kernel, dim3(kBatchGrid), dim3(SX * SY), 0, s, t, &t->next);
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
