// zpair.hpp - the z-pair and octet copies of a volume read by the forward
// projectors.
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
#pragma once

#include "device_common.hpp"

namespace tomo {

__device__ __forceinline__ size_t zpair_index(int a, int zz, int bb, int nz, int nB) {
    return (((size_t)a * (nz + 1) + zz) * (nB + 2) + bb) * 2;
}

// Octet copies of a 16-bit volume (read by the RS projector).  For slabs
// along A (0: x, 1: y; B the other in-plane axis) element (a, zz, bb),
// a in [0, nA], zz in [0, nz], bb in [0, nB], B varying fastest, is the
// cell whose lower corner is (A = a-1, B = bb-1, z = zz-1): its eight
// voxels, 0 outside the grid, in the order ox + 2 oy + 4 oz of their (x, y,
// z) offsets, 16 bytes in all.  A trilinear sample is one load.  The
// arguments of octet_index are the corner itself, each in [-1, n-1].
__device__ __forceinline__ size_t octet_index(int a, int z, int b, int nz, int nB) {
    return ((size_t)(a + 1) * (nz + 1) + (size_t)(z + 1)) * (nB + 1) + (size_t)(b + 1);
}

// Fills ox (A = x) and oy (A = y), octet_elems each, from an x-fastest
// 16-bit volume.
hipError_t launch_octet(const void* vol, void* ox, void* oy, int nx, int ny, int nz, hipStream_t s);

// Fills qx and qy from an x-fastest volume of es-byte elements (es 4 or 2).
hipError_t launch_zpair(const void* vol, void* qx, void* qy, size_t es, int nx, int ny, int nz, hipStream_t s);

// Both copies in the stream-ordered pool for the duration of one operator
// call: alloc() then release() on the same stream.
struct ZPairs {
    void* qx = nullptr;
    void* qy = nullptr;
    hipError_t alloc(const void* vol, size_t es, int nx, int ny, int nz, hipStream_t s) {
        hipError_t e = hipMallocAsync(&qx, zpair_elems_x(nx, ny, nz) * es, s);
        if (e == hipSuccess) e = hipMallocAsync(&qy, zpair_elems_y(nx, ny, nz) * es, s);
        if (e == hipSuccess) e = launch_zpair(vol, qx, qy, es, nx, ny, nz, s);
        return e;
    }
    hipError_t release(hipStream_t s) {
        hipError_t e = hipSuccess, e2 = hipSuccess;
        if (qy) e = hipFreeAsync(qy, s);
        if (qx) e2 = hipFreeAsync(qx, s);
        qx = qy = nullptr;
        return e != hipSuccess ? e : e2;
    }
};

// Both octet copies in the stream-ordered pool for one operator call.
struct Octets {
    void* ox = nullptr;
    void* oy = nullptr;
    hipError_t alloc(const void* vol, int nx, int ny, int nz, hipStream_t s) {
        hipError_t e = hipMallocAsync(&ox, octet_elems(nx, ny, nz) * 16, s);
        if (e == hipSuccess) e = hipMallocAsync(&oy, octet_elems(ny, nx, nz) * 16, s);
        if (e == hipSuccess) e = launch_octet(vol, ox, oy, nx, ny, nz, s);
        return e;
    }
    hipError_t release(hipStream_t s) {
        hipError_t e = hipSuccess, e2 = hipSuccess;
        if (oy) e = hipFreeAsync(oy, s);
        if (ox) e2 = hipFreeAsync(ox, s);
        ox = oy = nullptr;
        return e != hipSuccess ? e : e2;
    }
};

}  // namespace tomo
