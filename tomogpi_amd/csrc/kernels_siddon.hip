// kernels_siddon.hip - Siddon projector (clean and reference-compatible
// forms), the exact transpose of the clean form, and its segment counter.
//
// Clean form (default).  Exact chord lengths (mm) through the voxels the ray
// crosses inside the FOV = inscribed cylinder of radius N delta/2 (the
// reference Siddon's FOV, SiddonProjector_compute_C_mem_CPU.cpp:99-111) cut by
// the z extent of the volume.  The source sits on the RS/VI side
// (-FOD(cos,sin)), so RS, VI and the clean Siddon share one geometry.  Ray
// setup (FOV entry, first voxel, first plane crossing on each axis, distance
// between crossings) is FP64; the traversal runs in FP32 relative to the FOV
// entry with the n-th crossing of an axis at fma(n, inc, first), so no
// crossing is lost to accumulated rounding; the chord-weighted sum is FP32.
//
// Compat form (TOMO_FLAG_SIDDON_COMPAT).  The reference C path
// (SiddonProjector_compute_C_mem_CPU.cpp:20-276) with its float/double
// evaluation order and quirks: source on the +FOD side (mirror of RS/VI),
// first voxel half a voxel behind the entry, last chord not clipped, indices
// % 256 with unsigned wrap, no z clip.  Defined for 256^3 volumes only.
#include "zpair.hpp"
#include "batch.hpp"
#include "stage_box.hpp"

namespace tomo {
namespace {

constexpr int BX = kProjTileU, BY = kProjTileV;

struct SidRay {
    int idx[3], step[3];
    float first[3], inc[3], tmax;
    bool ok;
};

// FP64 ray setup (first crossings at +0 or later: "+ 0.0f" turns a -0 into
// +0).  tmax is the FOV chord, shortened to the first crossing
// that takes a voxel index out of the grid (the n-th crossing of axis a is
// at fma(n, inc_a, first_a), the value the traversal computes).  A traversal
// that stops at that crossing takes exactly the segments of one that steps
// out of the grid there and then stops, so no index range test is needed in
// the loops below.
__device__ __forceinline__ SidRay sid_setup(const DevGeom& g, int p, int m, int n) {
    SidRay r;
    r.ok = false;
    const double* vw = g.view + (size_t)p * kViewStride;
    const double ca = vw[0], sa = vw[1], dl = (double)g.delta;
    const double s = ((double)m - (double)g.cu) * (double)g.du;
    const double t = ((double)n - (double)g.cv) * (double)g.dv;
    const double rx = ca * g.fdd - s * sa, ry = sa * g.fdd + s * ca, rz = t;
    const double L = sqrt((rx * rx + ry * ry) + rz * rz);
    const double d[3] = {rx / L, ry / L, rz / L};
    const double S[3] = {-g.fod * ca, -g.fod * sa, 0.0};
    const double Rm = 0.5 * (double)g.nx * dl;
    const double a = d[0] * d[0] + d[1] * d[1];
    if (!(a > 0.0)) return r;
    const double b = S[0] * d[0] + S[1] * d[1];
    const double c = (S[0] * S[0] + S[1] * S[1]) - Rm * Rm;
    const double disc = b * b - a * c;
    if (!(disc > 0.0)) return r;
    const double sq = sqrt(disc);
    double l0 = (-b - sq) / a, l1 = (-b + sq) / a;
    const double zh = 0.5 * (double)g.nz * dl;
    if (d[2] != 0.0) {
        const double ta = -zh / d[2], tb = zh / d[2];
        l0 = fmax(l0, fmin(ta, tb));
        l1 = fmin(l1, fmax(ta, tb));
    }
    if (!(l1 > l0)) return r;
    const double cen[3] = {(double)g.cx, (double)g.cy, (double)g.cz};
    const int nn[3] = {g.nx, g.ny, g.nz};
    float tmax = (float)(l1 - l0);
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        const double pe = (S[ax] + l0 * d[ax]) / dl + cen[ax];
        double fi = floor(pe + 0.5);
        if (fi < 0.0) fi = 0.0;
        if (fi > (double)(nn[ax] - 1)) fi = (double)(nn[ax] - 1);
        r.idx[ax] = (int)fi;
        if (d[ax] > 0.0) {
            r.step[ax] = 1;
            r.first[ax] = (float)fmax(((fi + 0.5 - cen[ax]) * dl - S[ax]) / d[ax] - l0, 0.0) + 0.0f;
            r.inc[ax] = (float)(dl / d[ax]);
            tmax = fminf(tmax, fmaf((float)(nn[ax] - 1 - r.idx[ax]), r.inc[ax], r.first[ax]));
        } else if (d[ax] < 0.0) {
            r.step[ax] = -1;
            r.first[ax] = (float)fmax(((fi - 0.5 - cen[ax]) * dl - S[ax]) / d[ax] - l0, 0.0) + 0.0f;
            r.inc[ax] = (float)(-dl / d[ax]);
            tmax = fminf(tmax, fmaf((float)r.idx[ax], r.inc[ax], r.first[ax]));
        } else {
            r.step[ax] = 0;
            r.first[ax] = 3.0e38f;
            r.inc[ax] = 0.0f;
        }
    }
    r.tmax = tmax;
    r.ok = true;
    return r;
}

// Per-ray traversal used by the transpose (MODE 1, scatter chord*gval) and
// the segment counter (MODE 2): t from 0 to tmax; the next segment ends at
// the nearest crossing; ties step every tied axis.  The index range test is
// kept as a second guard for the scatter addresses.
template <int MODE>
__device__ __forceinline__ void sid_trace(const DevGeom& g, const SidRay& r, float* __restrict__ scatter, float gval,
                                          unsigned* nseg) {
    float tx = r.first[0], ty = r.first[1], tz = r.first[2];
    float cx = 0.0f, cy = 0.0f, cz = 0.0f;
    int i = r.idx[0], j = r.idx[1], k = r.idx[2];
    const int64_t sy = g.nx, sz = (int64_t)g.nx * g.ny;
    float t = 0.0f;
    int guard = g.nx + g.ny + g.nz + 4;
    unsigned cnt = 0;
    while (t < r.tmax && guard-- > 0) {
        const float tn = fminf(tx, fminf(ty, tz));
        const float te = fminf(tn, r.tmax);
        const float chord = te - t;
        const int64_t o = (int64_t)i + sy * j + sz * k;
        if (MODE == 1) unsafeAtomicAdd(scatter + o, chord * gval);
        if (MODE == 2) cnt += chord > 0.0f ? 1u : 0u;
        t = te;
        if (tx == tn) { i += r.step[0]; cx += 1.0f; tx = fmaf(cx, r.inc[0], r.first[0]); }
        if (ty == tn) { j += r.step[1]; cy += 1.0f; ty = fmaf(cy, r.inc[1], r.first[1]); }
        if (tz == tn) { k += r.step[2]; cz += 1.0f; tz = fmaf(cz, r.inc[2], r.first[2]); }
        if ((unsigned)i >= (unsigned)g.nx || (unsigned)j >= (unsigned)g.ny || (unsigned)k >= (unsigned)g.nz) break;
    }
    if (MODE == 2) *nseg = cnt;
}

// Wave-wide minimum / maximum of an int (wave64 butterfly).
__device__ __forceinline__ int wave_min(int v) {
    for (int off = 32; off > 0; off >>= 1) v = min(v, __shfl_xor(v, off, 64));
    return v;
}
__device__ __forceinline__ int wave_max(int v) {
    for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off, 64));
    return v;
}

// Volume reads by linear element index.  BUF: a buffer resource with a
// 32-bit byte offset (volumes below 4 GiB); an offset past the end reads 0.
// Otherwise a plain global load.
template <typename TV, bool BUF> struct VolReader;
template <typename TV> struct VolReader<TV, false> {
    const TV* p;
    float inv;
    __device__ __forceinline__ float operator()(unsigned lin) const { return Storage<TV>::load(p, lin, inv); }
};
template <> struct VolReader<float, true> {
    __amdgpu_buffer_rsrc_t r;
    float inv;
    __device__ __forceinline__ float operator()(unsigned lin) const {
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, lin * 4u, 0, 0));
    }
};
template <> struct VolReader<__half, true> {
    __amdgpu_buffer_rsrc_t r;
    float inv;
    __device__ __forceinline__ float operator()(unsigned lin) const {
        const unsigned short b = __builtin_amdgcn_raw_buffer_load_b16(r, lin * 2u, 0, 0);
        return __half2float(__builtin_bit_cast(__half, b)) * inv;
    }
};
template <> struct VolReader<__hip_bfloat16, true> {
    __amdgpu_buffer_rsrc_t r;
    float inv;
    __device__ __forceinline__ float operator()(unsigned lin) const {
        const unsigned short b = __builtin_amdgcn_raw_buffer_load_b16(r, lin * 2u, 0, 0);
        return __uint_as_float((unsigned)b << 16) * inv;
    }
};

template <typename TV, bool BUF>
__device__ __forceinline__ VolReader<TV, BUF> make_reader(const TV* p, unsigned bytes, float inv) {
    VolReader<TV, BUF> v;
    if constexpr (BUF) {
        v.r = buffer_rsrc(p, bytes);
    } else {
        v.p = p;
        (void)bytes;
    }
    v.inv = inv;
    return v;
}

// Minimum and maximum of two times of the walk.  Every such time is +0 or
// positive (the first crossings are clamped at +0, increments and tmax are
// positive), and non-negative floats order as their bit patterns do as
// unsigned integers, so one integer min / max gives the exact result without
// the NaN canonicalisation of fminf / fmaxf.
__device__ __forceinline__ float min_nn(float a, float b) {
    return __builtin_bit_cast(float, min(__builtin_bit_cast(uint32_t, a), __builtin_bit_cast(uint32_t, b)));
}
__device__ __forceinline__ float max_nn(float a, float b) {
    return __builtin_bit_cast(float, max(__builtin_bit_cast(uint32_t, a), __builtin_bit_cast(uint32_t, b)));
}

// Per-lane state of the general forward traversal: next crossing and
// crossings taken per axis, the linear index of the current voxel, the
// running sum.
struct SidState {
    float tn[3], cnt[3];
    unsigned lin;
    unsigned dlin[3];
    float t, acc;
    bool live;
};

template <int AX>
__device__ __forceinline__ void sid_advance(SidState& s, const SidRay& r) {
    s.cnt[AX] += 1.0f;
    s.tn[AX] = fmaf(s.cnt[AX], r.inc[AX], r.first[AX]);
    s.lin += s.dlin[AX];
}

// General step (waves without a common slab axis): the next segment ends at
// the nearest crossing; ties step every tied axis.
template <typename RD>
__device__ __forceinline__ void sid_step(SidState& s, const SidRay& r, const RD& rd) {
    const float tn = min_nn(s.tn[0], min_nn(s.tn[1], s.tn[2]));
    const float te = min_nn(tn, r.tmax);
    s.acc = fmaf(te - s.t, rd(s.lin), s.acc);
    s.t = te;
    if (s.tn[0] == tn) sid_advance<0>(s, r);
    if (s.tn[1] == tn) sid_advance<1>(s, r);
    if (s.tn[2] == tn) sid_advance<2>(s, r);
    s.live = s.t < r.tmax;
}

// Slab walk on the z-pair copy.  Per slab one load gives the current voxel
// V0, its B neighbour Vb, its z neighbour Vz and Vbz (in the step directions).
// The wave was admitted only if, for each of its rays, consecutive crossings
// of B and of z are spaced more than two ulps of tmax beyond the slab
// length, so a slab holds at most one crossing of each; its segments are
//   [t, e1] V0,  [e1, e2] Vb or Vz (whichever comes first),  [e2, tA] Vb, Vz or Vbz
// with e1 <= e2 the interior crossings (tA when absent), every end clipped to
// tmax.  These are the segments of the general step in the same order; an
// absent one has zero length and adds exactly 0.  ZNEG: the wave's rays go
// down in z (rows on one side of the detector centre, one sign).
// Quad reads of a z-pair copy: a buffer resource with a 32-bit byte offset
// (copies below 4 GiB; the walk carries the offset in bytes) or a plain
// global load by element index.
template <typename TV, bool BUF> struct QuadReader;
// A byte offset past every copy the buffer path reads (copies stay below it):
// a read there returns zeros without touching memory.
constexpr uint32_t kSidOutside = 0xfffffff0u;
template <typename TV> struct QuadReader<TV, false> {
    static constexpr uint32_t kUnit = 1;  // element index
    static constexpr bool kAnyOffset = false;  // only a live lane's offset may be read
    const TV* p;
    float inv;
    __device__ __forceinline__ float4 operator()(uint32_t e) const { return Storage<TV>::load4(p, e, inv); }
};
template <typename TV> struct QuadReader<TV, true> {
    static constexpr uint32_t kUnit = sizeof(TV);  // byte offset
    static constexpr bool kAnyOffset = true;  // an offset past the copy reads zeros
    __amdgpu_buffer_rsrc_t r;
    float inv;
    __device__ __forceinline__ float4 operator()(uint32_t off) const { return BufLoad4<TV>::load(r, off, 0u, inv); }
};
template <typename TV, bool BUF>
__device__ __forceinline__ QuadReader<TV, BUF> make_quad_reader(const TV* p, uint32_t bytes, float inv) {
    QuadReader<TV, BUF> q;
    if constexpr (BUF) q.r = buffer_rsrc(p, bytes);
    else q.p = p;
    q.inv = inv;
    return q;
}

// BNEG: 0 / 1 when every ray of the wave steps up / down along B, 2 when
// the wave mixes both (per-lane selects).
template <int A, bool ZNEG, int BNEG, typename QR>
__device__ __forceinline__ void sid_walk_quads(const SidRay& r, const QR& rd, int nA, int nB, int nz, bool live,
                                               float& acc_out) {
    constexpr int B = 1 - A;
    const int sA = r.step[A], sB = r.step[B];
    const bool bneg = BNEG == 2 ? sB < 0 : BNEG == 1;
    float tA = r.first[A], tB = r.first[B], tz = r.first[2];
    float cA = 0.0f, cB = 0.0f, cz = 0.0f;
    int iA = r.idx[A];
    float t = 0.0f, acc = 0.0f;
    // Offset of the block corner (A = iA, B = iB or iB - 1, z = iz or
    // iz - 1), carried by the steps.  A live lane's voxel is inside the grid
    // (tmax stops at the first crossing that leaves it), so the corner is in
    // the copy; the offset of a finished lane is never used.
    constexpr uint32_t u = QR::kUnit;
    const uint32_t sz_e = 2u * (uint32_t)(nB + 2), sa_e = sz_e * (uint32_t)(nz + 1);
    uint32_t e = 0u;
    if (live) {
        const int bb = (bneg ? r.idx[B] - 1 : r.idx[B]) + 1;
        const int zz = (ZNEG ? r.idx[2] - 1 : r.idx[2]) + 1;
        e = ((uint32_t)iA * sa_e + (uint32_t)zz * sz_e + 2u * (uint32_t)bb) * u;
    }
    const uint32_t dA = (uint32_t)sA * sa_e * u, dB = 2u * (uint32_t)sB * u, dZ = (uint32_t)r.step[2] * sz_e * u;
    // Lockstep: the wave's slab starts at the first slab of its lanes (in
    // each lane's direction along A) and advances one slab per iteration; a
    // lane joins at iteration lag = |iA - I| and then takes one slab per
    // iteration while it is live.
    const int I = sA > 0 ? wave_min(live ? iA : 0x7fffffff) : wave_max(live ? iA : -0x7fffffff);
    const int lag = abs(iA - I);
    // The wave's iterations, counted once (a scalar loop instead of a vote
    // per slab): a lane's slabs are at iterations lag .. lag + K, K the first
    // count with fma(K, inc, first) >= tmax (its last slab); the estimate
    // below, a rounded quotient of two FP32 times, is off by less than one,
    // and the count adds a margin of two.  Extra iterations find no live lane
    // and skip the body.
    int span = 0;
    if (live) {
        const float kf = ceilf((r.tmax - r.first[A]) * __builtin_amdgcn_rcpf(r.inc[A]));
        span = lag + (int)fminf(fmaxf(kf, 0.0f), 1.0e6f) + 3;
    }
    const int trips = __builtin_amdgcn_readfirstlane(min(wave_max(span), nA + 2));
    // One slab in two halves: advance() takes the slab's segment ends and
    // steps the walk to the next slab (its block offset follows from the
    // crossing times alone, not from the voxels); accumulate() adds the
    // slab's segments, given its block.
    struct Slab {
        bool act, stepB, stepZ, bfirst;
        float t0, c1, c2, c3;
    };
    auto advance = [&](int outer) __attribute__((always_inline)) {
        Slab sl;
        sl.act = live && outer >= lag;
        // Segment ends: c3 = min(tA, tmax) closes the slab; the interior
        // crossings are c1 = min(tB, tz, c3) and c2 = min(max(tB, tz), c3)
        // (min(x, tA) of each crossing, in order, clipped to tmax).  The
        // middle segment's voxel is Vb when B is crossed first and Vz when z
        // is; the last one's is the voxel after the crossings taken (tB <= tA:
        // B stepped).  Where these choices differ from the general step's (a
        // tie, or a crossing exactly at tA) the segment has zero length and
        // adds exactly 0.
        sl.stepB = tB <= tA;
        sl.stepZ = tz <= tA;
        sl.bfirst = tB < tz;
        sl.c3 = min_nn(tA, r.tmax);
        sl.c1 = min_nn(min_nn(tB, tz), sl.c3);
        sl.c2 = min_nn(max_nn(tB, tz), sl.c3);
        sl.t0 = t;
        if (sl.act) {
            t = sl.c3;
            // the next crossing times follow from the counts (an axis not
            // stepped recomputes the same time)
            e += (sl.stepB ? dB : 0u) + (sl.stepZ ? dZ : 0u) + dA;
            cB += sl.stepB ? 1.0f : 0.0f;
            cz += sl.stepZ ? 1.0f : 0.0f;
            tB = fmaf(cB, r.inc[B], r.first[B]);
            tz = fmaf(cz, r.inc[2], r.first[2]);
            cA += 1.0f;
            tA = fmaf(cA, r.inc[A], r.first[A]);
            live = t < r.tmax;
        }
        return sl;
    };
    auto accumulate = [&](const Slab& sl, const float4& v) __attribute__((always_inline)) {
        if (!sl.act) return;
        // v = (b, z), (b, z+1), (b+1, z), (b+1, z+1) of the block's corner
        const float V0 = bneg ? (ZNEG ? v.w : v.z) : (ZNEG ? v.y : v.x);
        const float Vb = bneg ? (ZNEG ? v.y : v.x) : (ZNEG ? v.w : v.z);
        const float Vz = bneg ? (ZNEG ? v.z : v.w) : (ZNEG ? v.x : v.y);
        const float Vbz = bneg ? (ZNEG ? v.x : v.y) : (ZNEG ? v.z : v.w);
        const float V1 = sl.bfirst ? Vb : Vz;
        const float V2 = sl.stepB ? (sl.stepZ ? Vbz : Vb) : (sl.stepZ ? Vz : V0);
        acc = fmaf(sl.c1 - sl.t0, V0, acc);
        acc = fmaf(sl.c2 - sl.c1, V1, acc);
        acc = fmaf(sl.c3 - sl.c2, V2, acc);
    };
    if constexpr (QR::kAnyOffset) {
        // Copies read through the buffer path: the next slab's block is
        // loaded before this slab's segments are summed, into two alternating
        // registers.  Every lane issues the load (a load guarded per lane is a
        // branch, at whose end the compiler waits for it, and so is a copy of
        // the prefetched value); a lane that will not use it reads past the
        // copy (zeros, no memory access), its offset formed with a mask.  At
        // 512^3 / 512 views: FP32 48.5 -> 46.5 ms, FP16 46.0 -> 42.7 ms; with
        // every lane reading its stale offset instead, the idle lanes' lines
        // cost FP32 what the prefetch gained.  The 64-bit index reads (copies
        // of 4 GiB or more) keep the direct form below: at 1024^3 the walk is
        // bound by the memory traffic, and the same prefetch (idle lanes
        // reading the copy's first block) took FP32 from 482 to 520 ms.
        auto next_off = [&](int outer_next) __attribute__((always_inline)) {
            const uint32_t keep = 0u - (uint32_t)(live && outer_next >= lag);
            return (e & keep) | (kSidOutside & ~keep);
        };
        float4 va = rd(next_off(0)), vb = va;
        for (int outer = 0; outer < trips; outer += 2) {
            Slab sl = advance(outer);
            vb = rd(next_off(outer + 1));
            accumulate(sl, va);
            if (outer + 1 < trips) {
                sl = advance(outer + 1);
                va = rd(next_off(outer + 2));
                accumulate(sl, vb);
            }
        }
    } else {
        for (int outer = 0; outer < trips; ++outer) {
            if (live && outer >= lag) {
                const float4 v = rd(e);
                const Slab sl = advance(outer);
                accumulate(sl, v);
            }
        }
    }
    acc_out = acc;
}

template <int A, bool ZNEG, typename QR>
__device__ __forceinline__ void sid_walk_quads_b(const SidRay& r, const QR& rd, int nA, int nB, int nz, bool live,
                                                 float& acc) {
    const bool pos_all = __ballot(live && r.step[1 - A] < 0) == 0ull;
    const bool neg_all = __ballot(live && r.step[1 - A] >= 0) == 0ull;
    if (pos_all) sid_walk_quads<A, ZNEG, 0>(r, rd, nA, nB, nz, live, acc);
    else if (neg_all) sid_walk_quads<A, ZNEG, 1>(r, rd, nA, nB, nz, live, acc);
    else sid_walk_quads<A, ZNEG, 2>(r, rd, nA, nB, nz, live, acc);
}

// Forward clean Siddon.  Each lane takes the segments of the per-ray
// traversal (same operations, same order), so results are bit-identical to
// it.  What changes is the schedule and the memory layout: when every ray of
// the wave has the same dominant in-plane axis A (|d_A| >= the other two
// components) and the same direction along it, the wave walks the slabs of A
// together.  The 64 rays of a wave are a 32-column x 2-row detector block, spread
// across the other in-plane axis, so at a given slab they read voxels that
// are contiguous when that other axis varies fastest: x-fastest (the caller's
// layout) for y-slabs, a y-fastest copy for x-slabs.  Waves that mix axes
// (views near 45 degrees) step per lane on the x-fastest volume.
// Grid: x = detector column group, y = view, z = detector row group.  Blocks
// are dispatched x fastest and spread over the XCDs round robin, so the chip
// works through every view of one row band before the next band: the rays of
// a row band stay in a thin z band of the volume, which the L2 and the
// Infinity Cache hold, instead of streaming the whole volume from HBM once
// per view.
template <typename TV, typename TS, bool BUF>
__device__ __forceinline__ void siddon_tile(const DevGeom& g, const TV* __restrict__ qx, const TV* __restrict__ qy,
                                            unsigned qx_bytes, unsigned qy_bytes, float vol_inv,
                                            TS* __restrict__ sino, float sino_scale, int bx, int p, int bz) {
    const int m = bx * BX + threadIdx.x;
    const int n = bz * BY + threadIdx.y;
    const bool inside = m < g.nu && n < g.nv;
    SidRay r;
    r.ok = false;
    if (inside) r = sid_setup(g, p, m, n);
    const bool ok = r.ok && r.tmax > 0.0f;
    // Slab code of this ray: 0/1 = x axis up/down, 2/3 = y axis up/down, when
    // the crossings of the other two axes are spaced farther apart than the
    // slabs by more than two ulps of tmax (inc = delta/|d|), and the first
    // slab is no longer than the others.  A FOV entry outside the grid along
    // the slab axis (the cylinder is wider than the grid when ny < nx) clamps
    // the first voxel and makes the first slab long; such rays step per lane.
    int code = -1;
    if (ok) {
        const float big = 3.0e38f, margin = 1.0e-6f * r.tmax;
        const float ix = r.step[0] ? r.inc[0] : big, iy = r.step[1] ? r.inc[1] : big;
        const float iz = r.step[2] ? r.inc[2] : big;
        if (r.step[0] != 0 && iy > ix + margin && iz > ix + margin && r.first[0] <= ix)
            code = r.step[0] > 0 ? 0 : 1;
        else if (r.step[1] != 0 && ix > iy + margin && iz > iy + margin && r.first[1] <= iy)
            code = r.step[1] > 0 ? 2 : 3;
    }
    const unsigned long long okmask = __ballot(ok);
    int mode = -1;
    if (okmask != 0ull) {
#pragma unroll
        for (int c = 0; c < 4; ++c)
            if (__ballot(ok && code == c) == okmask) mode = c;
    }
    // one z direction per wave (its rows on one side of the detector centre), or none
    const bool zneg_all = __ballot(ok && r.step[2] > 0) == 0ull;
    const bool zpos_all = __ballot(ok && r.step[2] < 0) == 0ull;
    if (!zneg_all && !zpos_all) mode = -1;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    float acc = 0.0f;
    if (mode < 0) {
        // Per-lane traversal on the y-slab copy, where voxel (x, y, z) is
        // element 2(x+1) + 2(nx+2)(z+1) + 2(nx+2)(nz+1) y (z-pair index of
        // corner (x, z) on plane y, first of the pair).
        const auto rd = make_reader<TV, BUF>(qy, qy_bytes, 1.0f);
        SidState s;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            s.tn[a] = r.first[a];
            s.cnt[a] = 0.0f;
        }
        // The copy holds fewer than 2^32 elements (checked by the launcher):
        // the index and its signed steps wrap correctly in 32-bit unsigned.
        const unsigned sz = 2u * (unsigned)(nx + 2), sy = sz * (unsigned)(nz + 1);
        s.lin = ok ? 2u * (unsigned)(r.idx[0] + 1) + sz * (unsigned)(r.idx[2] + 1) + sy * (unsigned)r.idx[1] : 0u;
        s.dlin[0] = 2u * (unsigned)r.step[0];
        s.dlin[1] = (unsigned)r.step[1] * sy;
        s.dlin[2] = (unsigned)r.step[2] * sz;
        s.t = 0.0f;
        s.acc = 0.0f;
        s.live = ok;
        // Every step moves one axis toward the grid exit that bounds tmax;
        // the cap only bounds the loop.
        int guard = nx + ny + nz + 4;
        while (s.live && guard-- > 0) sid_step(s, r, rd);
        acc = s.acc;
    } else if (mode < 2) {
        const auto rq = make_quad_reader<TV, BUF>(qx, qx_bytes, 1.0f);
        if (zneg_all && !zpos_all) sid_walk_quads_b<0, true>(r, rq, nx, ny, nz, ok, acc);
        else sid_walk_quads_b<0, false>(r, rq, nx, ny, nz, ok, acc);
    } else {
        const auto rq = make_quad_reader<TV, BUF>(qy, qy_bytes, 1.0f);
        if (zneg_all && !zpos_all) sid_walk_quads_b<1, true>(r, rq, ny, nx, nz, ok, acc);
        else sid_walk_quads_b<1, false>(r, rq, ny, nx, nz, ok, acc);
    }
    // 16-bit volumes are read unscaled; the sum is linear in them, so the
    // storage scale applies once here (FP32 storage has no scale).
    if (sizeof(TV) == 2) acc *= vol_inv;
    if (inside) Storage<TS>::store(sino, (size_t)m + (size_t)g.nu * ((size_t)n + (size_t)g.nv * p), acc, sino_scale);
}

// ---------------------------------------------------------------------------
// LDS-staged clean Siddon (FP32 volumes).
//
// The workgroup (32 x 8 detector cells of one view) walks the slabs of A in
// lockstep when all its rays share A and its direction: at iteration k every
// live lane is in slab I + k sA, I the workgroup's first slab.  Iterations go
// in windows of up to kSidWmax slabs [wlo, whi] (fewer when the box is too
// tall for a buffer); for each window the workgroup copies the voxels its
// rays can touch there into LDS (stage_box.hpp: corner rays at the window's
// boundary planes wlo - 1/2 and whi + 1/2, the voxel round(B) and its
// neighbours, likewise in z), and each slab's quad -- the current
// voxel, its B neighbour, its z neighbour and the diagonal one -- is two
// ds_read2_b32 there instead of one 16-byte gather from a z-pair copy.  The
// per-slab arithmetic is sid_walk_quads', so every lane takes the segments
// of the per-ray traversal in its order: results are bit-identical to it.
// A wave whose rays go up and down in z copies its share of rows and keeps
// the barriers, then steps per lane; a window whose box exceeds the buffer
// reads the quads from the volume; a workgroup whose rays do not share A and
// its direction steps per lane on the volume.
constexpr int kSidWmax = 8;  // slabs per window at most

// Per-lane general traversal on the volume (x fastest), as sid_step.
template <typename RD>
__device__ __forceinline__ float sid_general(const SidRay& r, const RD& rd, bool ok, int nx, int ny, int nz) {
    SidState s;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        s.tn[a] = r.first[a];
        s.cnt[a] = 0.0f;
    }
    const unsigned sy = (unsigned)nx, sz = (unsigned)nx * (unsigned)ny;
    s.lin = ok ? (unsigned)r.idx[0] + sy * (unsigned)r.idx[1] + sz * (unsigned)r.idx[2] : 0u;
    s.dlin[0] = (unsigned)r.step[0];
    s.dlin[1] = (unsigned)r.step[1] * sy;
    s.dlin[2] = (unsigned)r.step[2] * sz;
    s.t = 0.0f;
    s.acc = 0.0f;
    s.live = ok;
    int guard = nx + ny + nz + 4;
    while (s.live && guard-- > 0) sid_step(s, r, rd);
    return s.acc;
}

// The quad of corner (iA, bc, zc) straight from the volume, taps off the
// grid 0: (b, z), (b, z+1), (b+1, z), (b+1, z+1).
template <int A>
__device__ __forceinline__ float4 sid_quad_direct(const float* __restrict__ vol, int nx, int ny, int nz, int iA, int bc,
                                                  int zc) {
    auto at = [&](int b, int z) {
        const int x = A == 0 ? iA : b, y = A == 0 ? b : iA;
        const bool ok = (unsigned)x < (unsigned)nx && (unsigned)y < (unsigned)ny && (unsigned)z < (unsigned)nz;
        return ok ? vol[(size_t)x + (size_t)nx * ((size_t)y + (size_t)ny * (size_t)z)] : 0.0f;
    };
    return make_float4(at(bc, zc), at(bc, zc + 1), at(bc + 1, zc), at(bc + 1, zc + 1));
}

// A window of slabs [wlo, whi] staged as planes wlo .. whi (s.lo = wlo).
// All fields are wave-uniform.
struct SidWindow {
    int wlo, whi;
    StageWin s;
};

// The window that starts at slab `from` in the walk's direction: the largest
// of kSidWmax, 4, 2, 1 slabs whose box fits a buffer.
__device__ __forceinline__ SidWindow sid_plan(const WgCorners& wc, int from, int sAw, int max_nzr) {
    SidWindow w{};
    for (int W = kSidWmax;; W >>= 1) {
        const int wlo = sAw > 0 ? from : from - W + 1, whi = wlo + W - 1;
        const StageBox b = stage_box(wc, (float)wlo - 0.5f, (float)whi + 0.5f, 0.5f, 1, 1);
        w = SidWindow{wlo, whi, stage_window(b, wlo, W, max_nzr)};
        if (w.s.fits || W == 1) break;
    }
    return w;
}

// walks: the wave takes part in the lockstep walk (otherwise it only copies
// its rows and keeps the barriers).  I: the workgroup's first slab, sAw its
// direction along A, trips its iteration count: at iteration k every live
// lane is in slab I + k sAw.  The staging is double buffered: the copy of
// window i + 1 is in flight while the workgroup walks window i, one barrier
// per window.
template <int A, bool ZNEG, int BNEG>
__device__ __forceinline__ float sid_walk_staged(const DevGeom& g, const SidRay& r, bool live, bool walks, int I,
                                                 int sAw, int trips, const float* __restrict__ src, int nBp,
                                                 const float* __restrict__ vol, const WgCorners& wc,
                                                 float* __restrict__ st0, float* __restrict__ st1, int lane, int wv) {
    constexpr int B = 1 - A;
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const int nA = A == 0 ? nx : ny;
    const size_t src_elems = (size_t)nBp * (size_t)nA * (size_t)nz;
    const int max_nzr = stage_max_rows(nA, nBp);
    const int sB = r.step[B];
    const bool bneg = BNEG == 2 ? sB < 0 : BNEG == 1;
    float tA = r.first[A], tB = r.first[B], tz = r.first[2];
    float cB = 0.0f, cz = 0.0f, cA = 0.0f;
    float t = 0.0f, acc = 0.0f;
    const int lag = live ? abs(r.idx[A] - I) : 0;
    const int b0 = bneg ? r.idx[B] - 1 : r.idx[B], z0 = ZNEG ? r.idx[2] - 1 : r.idx[2];
    // the quad corner's (z, B) part of its offset in the window's buffer, (z - zlo) P + B - blo4
    // with P floats per staged row (set per window); the plane's part is wave-uniform
    int ez = 0;
    const int dB = sB;
    int dZ = 0;
    auto body = [&](const float4 v) {
        // v = (b, z), (b, z+1), (b+1, z), (b+1, z+1) of the block's corner
        const float V0 = bneg ? (ZNEG ? v.w : v.z) : (ZNEG ? v.y : v.x);
        const float Vb = bneg ? (ZNEG ? v.y : v.x) : (ZNEG ? v.w : v.z);
        const float Vz = bneg ? (ZNEG ? v.z : v.w) : (ZNEG ? v.x : v.y);
        const float Vbz = bneg ? (ZNEG ? v.x : v.y) : (ZNEG ? v.z : v.w);
        const bool stepB = tB <= tA, stepZ = tz <= tA;
        const float V1 = tB < tz ? Vb : Vz;
        const float V2 = stepB ? (stepZ ? Vbz : Vb) : (stepZ ? Vz : V0);
        const float c3 = min_nn(tA, r.tmax);
        const float c1 = min_nn(min_nn(tB, tz), c3), c2 = min_nn(max_nn(tB, tz), c3);
        acc = fmaf(c1 - t, V0, acc);
        acc = fmaf(c2 - c1, V1, acc);
        acc = fmaf(c3 - c2, V2, acc);
        t = c3;
        ez += (stepB ? dB : 0) + (stepZ ? dZ : 0);
        cB += stepB ? 1.0f : 0.0f;
        cz += stepZ ? 1.0f : 0.0f;
        tB = fmaf(cB, r.inc[B], r.first[B]);
        tz = fmaf(cz, r.inc[2], r.first[2]);
        cA += 1.0f;
        tA = fmaf(cA, r.inc[A], r.first[A]);
        live = t < r.tmax;
    };
    const int last = I + sAw * (trips - 1);  // the slab of the last iteration
    SidWindow cur = sid_plan(wc, I, sAw, max_nzr);
    if (cur.s.fits) stage_copy(src, src_elems, nA, nBp, nz, cur.s, st0, lane, wv);
    // One window: window `cur` is in `stage`, the next one is copied into
    // `other`; false after the last window.  The windows alternate between
    // the two LDS arrays and each call names them directly, so the compiler
    // knows that the copy in flight does not write what the reads read and
    // does not wait for it before them.
    auto window = [&](const float* __restrict__ stage, float* __restrict__ other) __attribute__((always_inline)) {
        // this wave's copies of window `cur` have landed; every wave is done with the other buffer
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        const int from = sAw > 0 ? cur.whi + 1 : cur.wlo - 1;
        const bool more = sAw > 0 ? from <= last : from >= last;
        SidWindow nxt = cur;
        if (more) {
            nxt = sid_plan(wc, from, sAw, max_nzr);
            if (nxt.s.fits) stage_copy(src, src_elems, nA, nBp, nz, nxt.s, other, lane, wv);
        }
        if (walks) {
            // iterations of the window's slabs, in the walk's order
            const int k0 = sAw > 0 ? cur.wlo - I : I - cur.whi;
            const int k1 = min(trips, k0 + (cur.whi - cur.wlo + 1));
            if (cur.s.fits) {
                const int P = 4 * cur.s.nch, ps = cur.s.nzr * P;
                dZ = r.step[2] * P;
                // the current quad corner, from the crossings taken so far
                const int bc = b0 + sB * (int)cB, zc = z0 + r.step[2] * (int)cz;
                ez = (zc - cur.s.zlo) * P + (bc - cur.s.blo4);
                for (int outer = k0; outer < k1; ++outer) {
                    if (live && outer >= lag) {
                        const int slab = I + sAw * outer;
                        const float* q = stage + (slab - cur.wlo) * ps + ez;
                        body(make_float4(q[0], q[P], q[1], q[P + 1]));
                    }
                }
            } else {
                for (int outer = k0; outer < k1; ++outer) {
                    if (live && outer >= lag) {
                        const int iA = I + sAw * outer;
                        const int bc = b0 + sB * (int)cB, zc = z0 + r.step[2] * (int)cz;
                        body(sid_quad_direct<A>(vol, nx, ny, nz, iA, bc, zc));
                    }
                }
            }
        }
        cur = nxt;
        return more;
    };
    while (window(st0, st1) && window(st1, st0)) {
    }
    return acc;
}

template <int A, bool ZNEG>
__device__ __forceinline__ float sid_walk_staged_b(const DevGeom& g, const SidRay& r, bool live, bool walks, int I,
                                                   int sAw, int trips, const float* __restrict__ src, int nBp,
                                                   const float* __restrict__ vol, const WgCorners& wc,
                                                   float* __restrict__ st0, float* __restrict__ st1, int lane,
                                                   int wv) {
    const bool pos_all = __ballot(live && r.step[1 - A] < 0) == 0ull;
    const bool neg_all = __ballot(live && r.step[1 - A] >= 0) == 0ull;
    if (pos_all) return sid_walk_staged<A, ZNEG, 0>(g, r, live, walks, I, sAw, trips, src, nBp, vol, wc, st0, st1, lane, wv);
    if (neg_all) return sid_walk_staged<A, ZNEG, 1>(g, r, live, walks, I, sAw, trips, src, nBp, vol, wc, st0, st1, lane, wv);
    return sid_walk_staged<A, ZNEG, 2>(g, r, live, walks, I, sAw, trips, src, nBp, vol, wc, st0, st1, lane, wv);
}

// Forward clean Siddon with LDS-staged quads (FP32 volume `vol`).  volT: its
// y-fastest copy with rows of nyp floats, volP: rows of nxp floats along x
// (the volume itself when nxp = nx); nxp, nyp are nx, ny rounded up to
// multiples of four.  Grid: x = column group, y = view, z = row group.
template <typename TS>
__global__ __launch_bounds__(BX* BY, 5) void siddon_staged_kernel(DevGeom g, const float* __restrict__ vol,
                                                                 const float* __restrict__ volT,
                                                                 const float* __restrict__ volP, int nxp, int nyp,
                                                                 unsigned vol_bytes, TS* __restrict__ sino,
                                                                 float sino_scale) {
    __shared__ float st0[kSlotFloats];
    __shared__ float st1[kSlotFloats];
    __shared__ int wsh[4][3];
    const int tid = threadIdx.y * BX + threadIdx.x;
    const int lane = tid & 63, wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int bx = blockIdx.x, p = blockIdx.y, bz = blockIdx.z;
    const int m = bx * BX + threadIdx.x;
    const int n = bz * BY + threadIdx.y;
    const bool inside = m < g.nu && n < g.nv;
    SidRay r;
    r.ok = false;
    if (inside) r = sid_setup(g, p, m, n);
    const bool ok = r.ok && r.tmax > 0.0f;
    // slab code as in siddon_tile
    int code = -1;
    if (ok) {
        const float big = 3.0e38f, margin = 1.0e-6f * r.tmax;
        const float ix = r.step[0] ? r.inc[0] : big, iy = r.step[1] ? r.inc[1] : big;
        const float iz = r.step[2] ? r.inc[2] : big;
        if (r.step[0] != 0 && iy > ix + margin && iz > ix + margin && r.first[0] <= ix)
            code = r.step[0] > 0 ? 0 : 1;
        else if (r.step[1] != 0 && ix > iy + margin && iz > iy + margin && r.first[1] <= iy)
            code = r.step[1] > 0 ? 2 : 3;
    }
    const unsigned long long okmask = __ballot(ok);
    int wmode = 4;  // 0..3: the wave's common code; -1: codes differ; 4: no ray
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
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    float acc = 0.0f;
    if (mode == -1) {
        if (vol_bytes != 0u)
            acc = sid_general(r, make_reader<float, true>(vol, vol_bytes, 1.0f), ok, nx, ny, nz);
        else
            acc = sid_general(r, make_reader<float, false>(vol, 0u, 1.0f), ok, nx, ny, nz);
    } else if (mode != 4) {
        const int A = mode < 2 ? 0 : 1;
        const int sAw = (mode & 1) ? -1 : 1;
        const int nA = A == 0 ? nx : ny;
        // the workgroup's first slab, then its iteration count (as sid_walk_quads, with the workgroup's I)
        const int iA = ok ? r.idx[A] : 0;
        int I = sAw > 0 ? wave_min(ok ? iA : 0x7fffffff) : wave_max(ok ? iA : -0x7fffffff);
        __syncthreads();  // every wave has read wsh[*][0]
        if (lane == 0) wsh[wv][1] = I;
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 4; ++k) I = sAw > 0 ? min(I, wsh[k][1]) : max(I, wsh[k][1]);
        I = __builtin_amdgcn_readfirstlane(I);
        int span = 0;
        if (ok) {
            const int lag = abs(iA - I);
            const float kf = ceilf((r.tmax - r.first[A]) * __builtin_amdgcn_rcpf(r.inc[A]));
            span = lag + (int)fminf(fmaxf(kf, 0.0f), 1.0e6f) + 3;
        }
        span = wave_max(span);
        if (lane == 0) wsh[wv][2] = span;
        __syncthreads();
        int trips = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) trips = max(trips, wsh[k][2]);
        trips = __builtin_amdgcn_readfirstlane(min(trips, nA + 2));
        // one z direction per wave, or none (then the wave steps per lane after the windows)
        const bool zneg_all = __ballot(ok && r.step[2] > 0) == 0ull;
        const bool zpos_all = __ballot(ok && r.step[2] < 0) == 0ull;
        const bool walks = zneg_all || zpos_all;
        const int mlo = bx * BX, mhi = min(bx * BX + BX - 1, g.nu - 1);
        const int nlo = bz * BY, nhi = min(bz * BY + BY - 1, g.nv - 1);
        const WgCorners wc = wg_corners(g, A, p, mlo, mhi, nlo, nhi);
        const float* src = A == 0 ? volT : volP;
        const int nBp = A == 0 ? nyp : nxp;
        const bool zneg = zneg_all && !zpos_all;
        if (A == 0) {
            acc = zneg ? sid_walk_staged_b<0, true>(g, r, ok, walks, I, sAw, trips, src, nBp, vol, wc, st0, st1, lane, wv)
                       : sid_walk_staged_b<0, false>(g, r, ok, walks, I, sAw, trips, src, nBp, vol, wc, st0, st1, lane, wv);
        } else {
            acc = zneg ? sid_walk_staged_b<1, true>(g, r, ok, walks, I, sAw, trips, src, nBp, vol, wc, st0, st1, lane, wv)
                       : sid_walk_staged_b<1, false>(g, r, ok, walks, I, sAw, trips, src, nBp, vol, wc, st0, st1, lane, wv);
        }
        if (!walks) {
            if (vol_bytes != 0u)
                acc = sid_general(r, make_reader<float, true>(vol, vol_bytes, 1.0f), ok, nx, ny, nz);
            else
                acc = sid_general(r, make_reader<float, false>(vol, 0u, 1.0f), ok, nx, ny, nz);
        }
    }
    if (inside) Storage<TS>::store(sino, (size_t)m + (size_t)g.nu * ((size_t)n + (size_t)g.nv * p), acc, sino_scale);
}

template <typename TV, typename TS, bool BUF>
__global__ __launch_bounds__(BX* BY) void siddon_kernel(DevGeom g, const TV* __restrict__ qx,
                                                      const TV* __restrict__ qy, unsigned qx_bytes,
                                                      unsigned qy_bytes, float vol_inv, TS* __restrict__ sino,
                                                      float sino_scale) {
    siddon_tile<TV, TS, BUF>(g, qx, qy, qx_bytes, qy_bytes, vol_inv, sino, sino_scale, blockIdx.x, blockIdx.y,
                             blockIdx.z);
}

// Shared launch: tiles are the single-problem grid's workgroups (column
// group fastest, then view, then row group); FP32 or BF16 (kItemBF16) z-pair
// copies, FP32 sinograms.
__global__ __launch_bounds__(BX* BY) void siddon_batch_kernel(const BatchTable* __restrict__ t, uint32_t* __restrict__ next) {
    TOMO_BATCH_CLAIMED(t, next) {
        const BatchItem& it = t->item[batch_find(t, tile)];
        const uint32_t local = tile - it.begin;
        const uint32_t bx = local % it.gx, rest = local / it.gx;
        const int by = (int)(rest % it.gy), bz = (int)(rest / it.gy);
        const bool bf = (it.flags & kItemBF16) != 0;
        const uint64_t es = bf ? 2ull : 4ull;
        const uint64_t qxb = zpair_elems_x(it.g.nx, it.g.ny, it.g.nz) * es;
        const uint64_t qyb = zpair_elems_y(it.g.nx, it.g.ny, it.g.nz) * es;
        const bool buf = qxb < kSidOutside && qyb < kSidOutside;
        using BF = __hip_bfloat16;
        if (bf && buf)
            siddon_tile<BF, float, true>(it.g, (const BF*)it.in, (const BF*)it.in2, (unsigned)qxb, (unsigned)qyb, 1.0f,
                                         (float*)it.out, 1.0f, (int)bx, by, bz);
        else if (bf)
            siddon_tile<BF, float, false>(it.g, (const BF*)it.in, (const BF*)it.in2, 0u, 0u, 1.0f, (float*)it.out,
                                          1.0f, (int)bx, by, bz);
        else if (buf)
            siddon_tile<float, float, true>(it.g, (const float*)it.in, (const float*)it.in2, (unsigned)qxb,
                                            (unsigned)qyb, 1.0f, (float*)it.out, 1.0f, (int)bx, by, bz);
        else
            siddon_tile<float, float, false>(it.g, (const float*)it.in, (const float*)it.in2, 0u, 0u, 1.0f,
                                             (float*)it.out, 1.0f, (int)bx, by, bz);
    }
}

template <typename TS>
__global__ __launch_bounds__(BX* BY) void siddon_adjoint_kernel(DevGeom g, const TS* __restrict__ sino,
                                                              float sino_inv, float* __restrict__ vol) {
    const int m = blockIdx.x * BX + threadIdx.x;
    const int n = blockIdx.z * BY + threadIdx.y;
    const int p = blockIdx.y;
    if (m >= g.nu || n >= g.nv) return;
    const float gv = Storage<TS>::load(sino, (size_t)m + (size_t)g.nu * ((size_t)n + (size_t)g.nv * p), sino_inv);
    if (gv == 0.0f) return;
    const SidRay r = sid_setup(g, p, m, n);
    if (r.ok) sid_trace<1>(g, r, vol, gv, nullptr);
}

__global__ __launch_bounds__(BX* BY) void siddon_count_kernel(DevGeom g, unsigned long long* total) {
    const int m = blockIdx.x * BX + threadIdx.x;
    const int n = blockIdx.z * BY + threadIdx.y;
    const int p = blockIdx.y;
    unsigned c = 0;
    if (m < g.nu && n < g.nv) {
        const SidRay r = sid_setup(g, p, m, n);
        if (r.ok) sid_trace<2>(g, r, nullptr, 0.0f, &c);
    }
    unsigned long long cc = c;
    for (int off = 32; off > 0; off >>= 1) cc += __shfl_down(cc, off, 64);
    if (((threadIdx.y * BX + threadIdx.x) & 63) == 0) atomicAdd(total, cc);  // first lane of the wave
}

// Reference C path, statement for statement in its arithmetic types.
template <typename TV, typename TS>
__global__ __launch_bounds__(BX* BY) void siddon_compat_kernel(DevGeom g, const TV* __restrict__ vol, float vol_inv,
                                                             TS* __restrict__ sino, float sino_scale) {
    const unsigned long un = blockIdx.x * BX + threadIdx.x;
    const unsigned long vn = blockIdx.z * BY + threadIdx.y;
    const int phi = blockIdx.y;
    const unsigned long nx = g.nx, ny = g.ny, nz = g.nz, nu = g.nu, nv = g.nv;
    if (un >= nu || vn >= nv) return;
    const float fdd = (float)g.fdd, fod = (float)g.fod, dl = g.delta;
    const float xlr = g.cx * dl, ylr = g.cy * dl, zlr = g.cz * dl;
    const float al = g.view_f[2 * phi], be = g.view_f[2 * phi + 1];
    const float xs = xlr + fod * al, ys = ylr + fod * be, zs = zlr;
    const float s = ((float)un - g.cu) * g.du;
    const float t = ((float)vn - g.cv) * g.dv;
    const float xd = (xlr - al * (fdd - fod)) - s * be;
    const float yd = (ylr - be * (fdd - fod)) + s * al;
    const float zd = zlr + t;
    const float L = sqrtf((fdd * fdd + s * s) + t * t);
    const float A = (fdd * fdd + s * s) / (L * L);
    const float B = (-fod * fdd) / L;
    const float C = (float)((double)(fod * fod) - 0.25 * (double)(float)nx * (double)(float)nx * (double)dl * (double)dl);
    const float dlam = B * B - A * C;
    float ray = 0.0f;
    if (dlam > 0.0f) {
        const float lmin = (-B - sqrtf(dlam)) / A;
        const float lmax = (-B + sqrtf(dlam)) / A;
        const float P[3] = {(xs + (lmin * (xd - xs) / L)) / dl, (ys + (lmin * (yd - ys) / L)) / dl,
                            (zs + (lmin * (zd - zs) / L)) / dl};
        const float S[3] = {xs, ys, zs}, D[3] = {xd, yd, zd};
        const unsigned long N[3] = {nx, ny, nz};
        unsigned long idx[3];
        float lam[3], inc[3];
        bool up[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            up[a] = S[a] <= D[a];
            const float pm = up[a] ? (float)((double)P[a] - 0.5) : (float)((double)P[a] + 0.5);
            if (pm < 0.0f) idx[a] = 0;
            else if (pm > (float)(N[a] - 1)) idx[a] = N[a] - 1;
            else idx[a] = up[a] ? (unsigned long)floorf(pm) : (unsigned long)ceilf(pm);
            const float span = D[a] - S[a];
            if (fabsf(span) > dl) {
                const double off = up[a] ? 0.5 : -0.5;
                lam[a] = (float)((double)L * (((double)(float)idx[a] + off) * (double)dl - (double)S[a]) / (double)span);
            } else {
                lam[a] = L;
            }
            inc[a] = (L / fabsf(span)) * dl;
        }
        float lambda = lmin;
        int it = 0;
        const int cap = 4 * (g.nx + g.ny + g.nz) + 16;
        while (lambda <= lmax && it++ < cap) {
            float lk = lam[0];
            if (lam[1] < lk) lk = lam[1];
            if (lam[2] < lk) lk = lam[2];
            const float chord = lk - lambda;
            const unsigned long o = (idx[0] % 256) + (idx[1] % 256) * nx + (idx[2] % 256) * nx * ny;
            const float prod = chord * Storage<TV>::load(vol, o, vol_inv);
            ray = ray + prod;
            lambda = lk;
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                if (lam[a] <= lk) {
                    lam[a] = lam[a] + inc[a];
                    idx[a] = up[a] ? idx[a] + 1 : idx[a] - 1;
                }
            }
        }
    }
    Storage<TS>::store(sino, (un % 256) + (vn % 256) * nu + (unsigned long)phi * nu * nv, ray, sino_scale);
}

}  // namespace

int launch_siddon(const tomo_geom* g, int compat, Buf vol, MutBuf sino, hipStream_t s, const ProjExtra& x) {
    const DevGeom& d = g->dev;
    if (compat && (d.nx != 256 || d.ny != 256 || d.nz != 256 || d.nu > 256 || d.nv > 256))
        return fail(TOMO_ERR_UNSUPPORTED,
                    "compat Siddon is defined for a 256^3 volume and at most 256x256 detector cells "
                    "(the reference indexes modulo 256)");
    // Grid: column group, view, row group (every view of a row band in turn).
    const dim3 grid(ceil_div(d.nu, BX), d.nphi, ceil_div(d.nv, BY));
    const float inv = 1.0f / vol.scale;
    if (compat) {
        TOMO_DISPATCH_STORAGE(vol.dtype, TV, {
            TOMO_DISPATCH_STORAGE(sino.dtype, TS, {
                hipLaunchKernelGGL((siddon_compat_kernel<TV, TS>), grid, dim3(BX, BY), 0, s, d, (const TV*)vol.p, inv,
                                   (TS*)sino.p, sino.scale);
            });
        });
        TOMO_HIP(hipGetLastError());
        return TOMO_OK;
    }
    if (x.staged && vol.dtype == TOMO_F32 && !x.index_reads) {
        // LDS-staged quads, from copies of the volume with rows along y (x
        // slabs) and, when nx is not a multiple of four, along x (y slabs)
        if (g->nvox() >= (1ull << 32))
            return fail(TOMO_ERR_UNSUPPORTED, "volume too large for the 32-bit traversal index of the Siddon projector");
        StageSources src;
        hipError_t le = src.make((const float*)vol.p, d.nx, d.ny, d.nz, s);
        if (le == hipSuccess) {
            const uint64_t vb = (uint64_t)g->nvox() * 4ull;
            const unsigned vol_bytes = vb < (1ull << 32) ? (unsigned)vb : 0u;  // 0: index reads
            TOMO_DISPATCH_STORAGE(sino.dtype, TS, {
                hipLaunchKernelGGL((siddon_staged_kernel<TS>), grid, dim3(BX, BY), 0, s, d, (const float*)vol.p,
                                   (const float*)src.vt, src.rows_x, src.nxp, src.nyp, vol_bytes, (TS*)sino.p,
                                   sino.scale);
            });
            le = hipGetLastError();
        }
        const hipError_t fe = src.release(s);
        TOMO_HIP(le);
        TOMO_HIP(fe);
        return TOMO_OK;
    }
    // z-pair copies of the volume for x slabs and y slabs (stream-ordered
    // pool allocations, released after the projection on the same stream).
    const size_t es = vol.dtype == TOMO_F32 ? 4 : 2;
    const size_t qx_elems = zpair_elems_x(d.nx, d.ny, d.nz), qy_elems = zpair_elems_y(d.nx, d.ny, d.nz);
    if (qx_elems >= (1ull << 32) || qy_elems >= (1ull << 32))
        return fail(TOMO_ERR_UNSUPPORTED, "volume too large for the 32-bit traversal index of the Siddon projector");
    const size_t qx_bytes = qx_elems * es, qy_bytes = qy_elems * es;
    ZPairs zp;
    hipError_t le = zp.alloc(vol.p, es, d.nx, d.ny, d.nz, s, x);
    if (le == hipSuccess) {
        const bool buf = qx_bytes < kSidOutside && qy_bytes < kSidOutside && !x.index_reads;
        TOMO_DISPATCH_STORAGE(vol.dtype, TV, {
            TOMO_DISPATCH_STORAGE(sino.dtype, TS, {
                if (buf)
                    hipLaunchKernelGGL((siddon_kernel<TV, TS, true>), grid, dim3(BX, BY), 0, s, d, (const TV*)zp.qx,
                                       (const TV*)zp.qy, (unsigned)qx_bytes, (unsigned)qy_bytes, inv, (TS*)sino.p,
                                       sino.scale);
                else
                    hipLaunchKernelGGL((siddon_kernel<TV, TS, false>), grid, dim3(BX, BY), 0, s, d, (const TV*)zp.qx,
                                       (const TV*)zp.qy, 0u, 0u, inv, (TS*)sino.p, sino.scale);
            });
        });
        le = hipGetLastError();
    }
    const hipError_t fe = zp.release(s);
    TOMO_HIP(le);
    TOMO_HIP(fe);
    return TOMO_OK;
}

hipError_t launch_batch_siddon(BatchTable* t, hipStream_t s) {
    hipLaunchKernelGGL(siddon_batch_kernel, dim3(kBatchGrid), dim3(BX, BY), 0, s, t, t->next);
    return hipGetLastError();
}

int launch_siddon_adjoint(const tomo_geom* g, Buf sino, MutBuf vol, hipStream_t s) {
    if (vol.dtype != TOMO_F32) return fail(TOMO_ERR_UNSUPPORTED, "adjoint output must be F32");
    const DevGeom& d = g->dev;
    int rc = launch_fill((float*)vol.p, 0.0f, g->nvox(), s);
    if (rc) return rc;
    const dim3 grid(ceil_div(d.nu, BX), d.nphi, ceil_div(d.nv, BY));
    const float inv = 1.0f / sino.scale;
    TOMO_DISPATCH_STORAGE(sino.dtype, TS, {
        hipLaunchKernelGGL((siddon_adjoint_kernel<TS>), grid, dim3(BX, BY), 0, s, d, (const TS*)sino.p, inv,
                           (float*)vol.p);
    });
    TOMO_HIP(hipGetLastError());
    return TOMO_OK;
}

int count_siddon_segments(tomo_ctx* ctx, const tomo_geom* g, uint64_t* out, hipStream_t s) {
    const DevGeom& d = g->dev;
    unsigned long long* dt = nullptr;
    TOMO_HIP(hipMallocAsync((void**)&dt, sizeof(unsigned long long), s));
    TOMO_HIP(hipMemsetAsync(dt, 0, sizeof(unsigned long long), s));
    const dim3 grid(ceil_div(d.nu, BX), d.nphi, ceil_div(d.nv, BY));
    hipLaunchKernelGGL(siddon_count_kernel, grid, dim3(BX, BY), 0, s, d, dt);
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
