// batch.hpp - problem tables of the shared (batched) launches.
//
// A shared launch runs one operator for up to kBatchMax independent problems
// of different sizes.  Its grid is fixed (kBatchGrid workgroups) and
// persistent: workgroup w takes tiles w, w + kBatchGrid, ... of the launch.
// The tiles of problem q are [item[q].begin, item[q].begin + item[q].tiles);
// a tile is one workgroup of the single-problem kernel's grid, numbered x
// fastest, then y, then z, and runs the same code as there.  A problem's
// results therefore do not depend on what else shares the launch.  Because
// the grid and the table address are fixed, a captured hipGraph of shared
// launches stays valid while the table contents change between replays.
#pragma once

#include "tomo_internal.hpp"

namespace tomo {

constexpr int kBatchMax = 64;
constexpr int kBatchGrid = 2048;
constexpr int kVecTile = 4096;  // elements per tile of the element-wise ops

struct BatchItem {
    DevGeom g;
    const void* in;   // operator input / first element-wise operand
    const void* in2;  // second input (z-pair qy, element-wise operand b)
    const void* in3;  // per-element weight of the element-wise ops (SIRT), or null
    void* out;        // output
    void* aux;        // second output (z-pair qy) or FP64 partial sums (one per tile)
    double* slot;     // per-problem FP64 scalars
    uint64_t n;       // element count (element-wise ops)
    uint32_t begin, tiles;
    uint32_t gx, gy;  // tile grid of the problem (x, y); z = tiles / (gx * gy)
    int32_t flags;
    float a;          // op constant
    double d;         // op constant (FP64)
};

struct BatchTable {
    int32_t count;
    uint32_t total;
    uint32_t pad[2];
    BatchItem item[kBatchMax];
};

// Per-problem scalars of the gradient loop, FP64, in `slot`.
enum GdSlot {
    kSlotResidual = 0,  // sum (y - P x)^2
    kSlotRegSq = 1,     // sum (L dJ)^2
    kSlotNormDj = 2,    // sum dJ^2
    kSlotProjSq = 3,    // sum (P dJ)^2
    kSlotAlpha = 4,
    kSlotAlphaMin = 5,
    kSlotNormDjProj = 6,
    kSlotCount = 8
};

#ifdef __HIP__
// Tile -> problem: the last item whose begin <= tile (block-uniform).
__device__ __forceinline__ int batch_find(const BatchTable* t, uint32_t tile) {
    int lo = 0, hi = t->count - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (t->item[mid].begin <= tile) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}
#endif

// Shared launches (stream-ordered; `t` is a device copy of the table).
// Storage is FP32 throughout.
hipError_t launch_batch_zpair_x(const BatchTable* t, hipStream_t s);      // in volume -> out qx
hipError_t launch_batch_zpair_y(const BatchTable* t, hipStream_t s);      // in volume -> out qy
hipError_t launch_batch_siddon(const BatchTable* t, hipStream_t s);       // in qx, in2 qy -> out sinogram
hipError_t launch_batch_rs(const BatchTable* t, hipStream_t s);           // in qx, in2 qy -> out; flags bit0 cube
hipError_t launch_batch_vi_quads(const BatchTable* t, hipStream_t s);     // in sinogram -> out tap quads
hipError_t launch_batch_vi(const BatchTable* t, hipStream_t s);           // in tap quads -> out volume * a
hipError_t launch_batch_diff_sq(const BatchTable* t, hipStream_t s);      // out = in - in2; aux partials of out^2
hipError_t launch_batch_sq(const BatchTable* t, hipStream_t s);           // aux partials of in^2
hipError_t launch_batch_laplacian(const BatchTable* t, hipStream_t s);    // out = L in; aux partials of out^2 if set
hipError_t launch_batch_axpy(const BatchTable* t, hipStream_t s);         // out += alpha * in (see kernels_vec.hip)
hipError_t launch_batch_fill(const BatchTable* t, hipStream_t s);         // out = a
hipError_t launch_batch_finish(const BatchTable* t, hipStream_t s);       // slot[flags] = sum of aux[0..n)
hipError_t launch_batch_step(const BatchTable* t, hipStream_t s);         // step rule on slot

}  // namespace tomo
