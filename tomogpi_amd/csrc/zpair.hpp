// zpair.hpp - the z-pair copies of a volume read by the forward projectors.
//
// For slabs along A (0: x, 1: y; B is the other in-plane axis) element
// (a, zz, bb), zz in [0, nz], bb in [0, nB+1], holds f(A=a, B=bb-1, z=zz-1)
// and f(A=a, B=bb-1, z=zz), 0 outside the grid, with B varying fastest.  Two
// neighbouring elements are the 2x2 (B, z) block with lower corner
// (bb-1, zz-1): one 16-byte (FP32) or 8-byte (16-bit) load,
// Storage<T>::load4 at zpair_index(a, z+1, b+1), returns
//   (b, z), (b, z+1), (b+1, z), (b+1, z+1)
// for any corner b in [-1, nB-1], z in [-1, nz-1], with the voxels outside
// the grid already 0.  qx (A = x) and qy (A = y) each hold
// nA * (nz+1) * (nB+2) * 2 elements, about twice the volume.
//
// The A-pair copies of a 16-bit volume (read by the RS projector, whose
// samples interpolate between two A planes) add the plane pair: group
// (ap, zz, bb), ap in [0, nA], holds four elements,
//   f(ap-1, bb-1, zz-1), f(ap-1, bb-1, zz), f(ap, bb-1, zz-1), f(ap, bb-1, zz)
// (planes -1 and nA are 0), B varying fastest; two neighbouring groups are the
// eight taps of a sample with lower corner (ap-1, bb-1, zz-1), one 16-byte
// load.  Each copy holds (nA+1) * (nz+1) * (nB+2) * 4 elements.
#pragma once

#include "device_common.hpp"

namespace tomo {

__device__ __forceinline__ size_t zpair_index(int a, int zz, int bb, int nz, int nB) {
    return (((size_t)a * (nz + 1) + zz) * (nB + 2) + bb) * 2;
}

// First element of A-pair group (ap, zz, bb).
__device__ __forceinline__ size_t apair_index(int ap, int zz, int bb, int nz, int nB) {
    return (((size_t)ap * (nz + 1) + zz) * (nB + 2) + bb) * 4;
}

// Fills qx and qy from an x-fastest volume of es-byte elements (es 4 or 2).
hipError_t launch_zpair(const void* vol, void* qx, void* qy, size_t es, int nx, int ny, int nz, hipStream_t s);

// Fills the A-pair copies qx and qy from an x-fastest 16-bit volume.
hipError_t launch_apair(const void* vol, void* qx, void* qy, int nx, int ny, int nz, hipStream_t s);

// out[y + nyp (x + nx z)] = vol[x + nx (y + ny z)] for y < ny and 0 for
// ny <= y < nyp (nyp >= ny), FP32: the volume with y fastest and rows padded
// with zeros, read by the LDS-staged projectors for slabs along x.
hipError_t launch_xy_transpose(const float* vol, float* out, int nx, int ny, int nyp, int nz, hipStream_t s);

// out[x + nxp r] = vol[x + nx r] for x < nx and 0 for nx <= x < nxp, r over
// the ny nz rows, FP32: rows along x padded with zeros.
hipError_t launch_row_pad(const float* vol, float* out, int nx, int nxp, size_t rows, hipStream_t s);

// The rows read by the LDS-staged projectors for the duration of one call:
// along y (vt, rows of nyp floats) for slabs along x, along x (the volume
// itself, or vp when nx is not a multiple of four) for slabs along y.  nxp,
// nyp: nx, ny rounded up to multiples of four.  make() then release() on the
// same stream; the copies come from the stream-ordered pool.
struct StageSources {
    float* vt = nullptr;
    float* vp = nullptr;
    const float* rows_x = nullptr;
    int nxp = 0, nyp = 0;
    hipError_t make(const float* vol, int nx, int ny, int nz, hipStream_t s) {
        nxp = (nx + 3) & ~3;
        nyp = (ny + 3) & ~3;
        hipError_t e = hipMallocAsync((void**)&vt, (size_t)nyp * nx * nz * sizeof(float), s);
        if (e == hipSuccess) e = launch_xy_transpose(vol, vt, nx, ny, nyp, nz, s);
        rows_x = vol;
        if (e == hipSuccess && nxp != nx) {
            e = hipMallocAsync((void**)&vp, (size_t)nxp * ny * nz * sizeof(float), s);
            if (e == hipSuccess) e = launch_row_pad(vol, vp, nx, nxp, (size_t)ny * nz, s);
            rows_x = vp;
        }
        return e;
    }
    hipError_t release(hipStream_t s) {
        hipError_t e = hipSuccess, e2 = hipSuccess;
        if (vp) e = hipFreeAsync(vp, s);
        if (vt) e2 = hipFreeAsync(vt, s);
        vp = vt = nullptr;
        return e != hipSuccess ? e : e2;
    }
};

// Both copies for the duration of one operator call, filled from the volume:
// alloc() then release() on the same stream.  The copies come from the
// stream-ordered pool, or are the caller's (ProjExtra::qx, qy, sized for the
// largest form), which release() leaves alone.  apair: the A-pair form
// (16-bit volumes) instead of the z-pair form.
struct ZPairs {
    void* qx = nullptr;
    void* qy = nullptr;
    bool owned = false;
    hipError_t alloc(const void* vol, size_t es, int nx, int ny, int nz, hipStream_t s,
                     const ProjExtra& x = ProjExtra{}, bool apair = false) {
        hipError_t e = hipSuccess;
        if (x.qx && x.qy) {
            qx = x.qx;
            qy = x.qy;
        } else {
            owned = true;
            const size_t bx = (apair ? apair_elems_x(nx, ny, nz) : zpair_elems_x(nx, ny, nz)) * es;
            const size_t by = (apair ? apair_elems_y(nx, ny, nz) : zpair_elems_y(nx, ny, nz)) * es;
            e = hipMallocAsync(&qx, bx, s);
            if (e == hipSuccess) e = hipMallocAsync(&qy, by, s);
        }
        if (e == hipSuccess)
            e = apair ? launch_apair(vol, qx, qy, nx, ny, nz, s) : launch_zpair(vol, qx, qy, es, nx, ny, nz, s);
        return e;
    }
    hipError_t release(hipStream_t s) {
        hipError_t e = hipSuccess, e2 = hipSuccess;
        if (owned && qy) e = hipFreeAsync(qy, s);
        if (owned && qx) e2 = hipFreeAsync(qx, s);
        qx = qy = nullptr;
        owned = false;
        return e != hipSuccess ? e : e2;
    }
};

}  // namespace tomo
