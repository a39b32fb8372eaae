// batch.hpp - problem tables of the shared (batched) launches.
//
// A shared launch runs one operator for up to kBatchMax independent problems
// of different sizes.  Its grid is fixed (kBatchGrid workgroups; RS: the
// workgroups the device holds at once, batch_grids_init) and
// persistent.  The tiles of problem q are [item[q].begin, item[q].begin +
// item[q].tiles); a tile is one workgroup of the single-problem kernel's
// grid, numbered x fastest, then y, then z, and runs the same code as there.
// A problem's results therefore do not depend on what else shares the launch
// or on which workgroup runs a tile.  Workgroup w starts with tile w.  The
// element-wise operators then take tiles w + kBatchGrid, w + 2 kBatchGrid, ...
// (equal-cost tiles); the projectors and the backprojector, whose tiles differ
// in cost with the problem's views and the rays' lengths, claim their next
// tile from a counter of the table instead (one vector atomic per tile, on a
// cache line apart from the items, broadcast through LDS), which balances them
// as hardware dispatch would.  The dispatcher deals a launch's workgroups to
// the kBatchXcds XCDs round robin, so workgroup w runs on XCD w mod kBatchXcds;
// the claims keep that mapping for tiles too: the workgroups of XCD x claim
// from counter x the tiles t = x (mod kBatchXcds), in order.  A single-problem
// launch puts the same tiles on the same XCD (tile t is its workgroup t, up to
// a rotation by the problem's first tile), so a detector column group meets
// the same XCD's L2 at every view, as it does there (one shared counter, which
// sends a tile to whichever XCD has a workgroup free, measured the same
// time).  Because the grid and the table address
// are fixed, a captured hipGraph of shared launches stays valid while the
// table contents change between replays; the host's copy of the table before
// each launch sequence resets the counter, so a table serves one launch per
// copy.
#pragma once

#include "tomo_internal.hpp"

namespace tomo {

constexpr int kBatchMax = 64;
constexpr int kBatchXcds = 8;     // XCDs of the chip: claim counters per table
constexpr int kBatchGrid = 2048;  // a multiple of kBatchXcds
constexpr int kVecTile = 4096;  // elements per tile of the element-wise ops
// VI workgroup (single problems and shared launches): kViStageX x kViStageY
// voxel columns of kViTileZ voxels, the taps staged in LDS.  Its transpose
// takes kViTileX x kViTileY columns (a wave is 64 consecutive x columns).
constexpr int kViTileX = 64, kViTileY = 4, kViTileZ = 16;
constexpr int kViStageX = 32, kViStageY = 8;

// Detector box of one (staged VI column tile, view): the extreme clamped
// cell corners floor(u) and the extreme row step A of the tile's columns.
struct ViBox {
    int32_t u0, u1;
    float amin, amax;
};
inline size_t vi_box_count(const DevGeom& d) {
    return (size_t)((d.nx + kViStageX - 1) / kViStageX) * ((d.ny + kViStageY - 1) / kViStageY) * d.nphi;
}
// Projector workgroup (RS, Siddon): kProjTileU detector columns x kProjTileV
// rows of one view; a wave is 32 columns x 2 rows, rays close enough to
// enter and leave the FOV within a few slabs of each other (the waves walk
// slabs or samples in lockstep).
constexpr int kProjTileU = 32, kProjTileV = 8;

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
    // Tiles claimed beyond the first kBatchGrid, per XCD (0 when the host
    // writes the table); their own cache line keeps the atomics off the items.
    alignas(128) uint32_t next[kBatchXcds];
};
static_assert(kBatchGrid % kBatchXcds == 0, "every XCD starts with the same number of workgroups");

// Item flag bit of the projector launches: the operator copies (z-pairs) are
// BF16, converted from the FP32 state by the copy launches.  RS keeps bit 0
// for the cube FOV.
constexpr int32_t kItemBF16 = 0x10;
// Item flag bit of the BF16 copy and RS launches: the copies are A-pair
// copies (zpair.hpp), read with one load per RS sample.
constexpr int32_t kItemApair = 0x20;

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
// Static order: tiles blockIdx.x, blockIdx.x + gridDim.x, ... (`tile` in the body).
#define TOMO_BATCH_STRIDED(t) for (uint32_t tile = blockIdx.x; tile < (t)->total; tile += gridDim.x)

// The workgroup's next claimed tile (every thread of the workgroup calls it):
// XCD x's k-th claim is tile x + kBatchXcds (gridDim.x / kBatchXcds + k), so
// the residue class x is covered once, after the first tiles blockIdx.x.
__device__ __forceinline__ uint32_t batch_claim(uint32_t* next, uint32_t* slot) {
    __syncthreads();  // the previous tile's readers of *slot are done
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        const uint32_t x = blockIdx.x % kBatchXcds;
        *slot = x + kBatchXcds * (gridDim.x / kBatchXcds + atomicAdd(next + x, 1u));
    }
    __syncthreads();
    return *slot;
}

// Dynamic order: tile blockIdx.x, then tiles claimed from the counters
// `next` (kBatchXcds of them).
#define TOMO_BATCH_CLAIMED(t, next)                                                         \
    __shared__ uint32_t tomo_batch_slot;                                                    \
    for (uint32_t tile = blockIdx.x; tile < (t)->total; tile = ::tomo::batch_claim((next), &tomo_batch_slot))

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
// The state is FP32 throughout; operator copies are FP32 or BF16 (kItemBF16).
hipError_t launch_batch_zpair_x(const BatchTable* t, hipStream_t s);   // in volume -> out qx (FP32/BF16)
hipError_t launch_batch_zpair_y(const BatchTable* t, hipStream_t s);   // in volume -> out qy (FP32/BF16)
hipError_t launch_batch_siddon(BatchTable* t, hipStream_t s);          // in qx, in2 qy -> out sinogram
hipError_t launch_batch_rs(BatchTable* t, hipStream_t s);              // in qx, in2 qy -> out; flags bit0 cube
// Grids of the shared launches whose workgroups take more registers than
// kBatchGrid resident workgroups allow (RS): the number the device holds at
// once, a multiple of kBatchXcds.  Called at context creation, before any
// launch is captured.
int batch_grids_init(int device);
hipError_t launch_batch_vi(BatchTable* t, hipStream_t s);              // in sinogram, aux boxes -> out volume * a
int launch_vi_boxes(const DevGeom& d, void* boxes, hipStream_t s);     // vi_box_count(d) ViBox of the geometry
hipError_t launch_batch_diff_sq(const BatchTable* t, hipStream_t s);   // out = in - in2; aux partials of out^2
hipError_t launch_batch_sq(const BatchTable* t, hipStream_t s);        // aux partials of in^2
hipError_t launch_batch_laplacian(const BatchTable* t, hipStream_t s); // out = L in; aux partials of out^2 if set
hipError_t launch_batch_axpy(const BatchTable* t, hipStream_t s);      // out += alpha * in (see kernels_vec.hip)
hipError_t launch_batch_fill(const BatchTable* t, hipStream_t s);      // out = a
hipError_t launch_batch_finish(const BatchTable* t, hipStream_t s);    // slot[flags] = sum of aux[0..n)
hipError_t launch_batch_step(const BatchTable* t, hipStream_t s);      // step rule on slot

}  // namespace tomo
