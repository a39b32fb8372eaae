// kernels_vec.hip - loop glue: element-wise volume / sinogram operations,
// FP64 reductions, the 7-point Laplacian and the ellipsoid phantoms.
//
// Element-wise ops follow Image3D_CPU / Sinogram3D_CPU (C/Image3D_CPU.cu:
// 108-205, C/Sinogram3D_CPU.cu:311-334): out = a - b, x *= s,
// y = y + (float)lambda * x with optional clamp at 0 (no fused multiply-add,
// as the reference host loops).  Squared norms are sum double(x)^2
// (C/Image3D_CPU.cu:363-390; "L2Norm" without the square root), reduced in a
// fixed order: per-block partials, then one block.
#include <mutex>
#include <type_traits>

#include "zpair.hpp"
#include "batch.hpp"

namespace tomo {
namespace {

constexpr int kBlock = 256;
constexpr int kReduceBlocks = 2048;

__global__ __launch_bounds__(kBlock) void fill_kernel(float* x, float v, size_t n) {
    for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) x[i] = v;
}
__global__ __launch_bounds__(kBlock) void diff_kernel(const float* a, const float* b, float* o, size_t n) {
    for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock)
        o[i] = a[i] - b[i];
}
__global__ __launch_bounds__(kBlock) void scale_kernel(float* x, float v, size_t n) {
    for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock)
        x[i] = x[i] * v;
}
__global__ __launch_bounds__(kBlock) void axpy_kernel(float* y, const float* x, float a, int positive, size_t n) {
    for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) {
        float v = y[i] + a * x[i];
        if (positive && v < 0.0f) v = 0.0f;
        y[i] = v;
    }
}

// SIRT element-wise steps (FP32, no fused multiply-add):
//   weights  w = s > 0 ? a / s : 0
//   residual d = y - e; out = d * w
//   update   x = x + c * b, clamped at 0 with positivity
__global__ __launch_bounds__(kBlock) void recip_weight_kernel(const float* s_in, float a, float* w, size_t n) {
    for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) {
        const float v = s_in[i];
        w[i] = v > 0.0f ? a / v : 0.0f;
    }
}
__global__ __launch_bounds__(kBlock) void mul_kernel(float* x, const float* w, size_t n) {
    for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock)
        x[i] = x[i] * w[i];
}
__global__ __launch_bounds__(kBlock) void waxpy_kernel(float* x, const float* c, const float* b, int positive,
                                                     size_t n) {
    for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) {
        float v = x[i] + c[i] * b[i];
        if (positive && v < 0.0f) v = 0.0f;
        x[i] = v;
    }
}

template <typename TI, typename TO>
__global__ __launch_bounds__(kBlock) void convert_kernel(const TI* in, float inv, TO* out, float scale, size_t n) {
    for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock)
        Storage<TO>::store(out, i, Storage<TI>::load(in, i, inv), scale);
}

__device__ __forceinline__ double block_sum(double v) {
    __shared__ double sh[kBlock / 64];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    double r = 0.0;
    if (threadIdx.x == 0)
        for (int w = 0; w < kBlock / 64; ++w) r += sh[w];
    return r;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void sq_partial_kernel(const T* x, float inv, size_t n, double* partial) {
    double s = 0.0;
    for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) {
        const double v = (double)Storage<T>::load(x, i, inv);
        s += v * v;
    }
    const double r = block_sum(s);
    if (threadIdx.x == 0) partial[blockIdx.x] = r;
}

__global__ __launch_bounds__(kBlock) void dot_partial_kernel(const float* a, const float* b, size_t n, double* partial) {
    double s = 0.0;
    for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock)
        s += (double)a[i] * (double)b[i];
    const double r = block_sum(s);
    if (threadIdx.x == 0) partial[blockIdx.x] = r;
}

__device__ __forceinline__ float block_max(float v) {
    __shared__ float shm[kBlock / 64];
    for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_down(v, off, 64));
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) shm[wid] = v;
    __syncthreads();
    float r = 0.0f;
    if (threadIdx.x == 0)
        for (int w = 0; w < kBlock / 64; ++w) r = fmaxf(r, shm[w]);
    return r;
}

// max |x| (NaN is ignored by fmaxf), partials then one final block.
__global__ __launch_bounds__(kBlock) void max_abs_partial_kernel(const float* x, size_t n, float* partial) {
    float m = 0.0f;
    for (size_t i = blockIdx.x * (size_t)kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock)
        m = fmaxf(m, fabsf(x[i]));
    const float r = block_max(m);
    if (threadIdx.x == 0) partial[blockIdx.x] = r;
}

__global__ __launch_bounds__(kBlock) void final_max_kernel(const float* partial, int count, float* out) {
    float m = 0.0f;
    for (int i = threadIdx.x; i < count; i += kBlock) m = fmaxf(m, partial[i]);
    const float r = block_max(m);
    if (threadIdx.x == 0) *out = r;
}

__global__ __launch_bounds__(kBlock) void final_sum_kernel(const double* partial, int count, double* out) {
    double s = 0.0;
    for (int i = threadIdx.x; i < count; i += kBlock) s += partial[i];
    const double r = block_sum(s);
    if (threadIdx.x == 0) *out = r;
}

// y = ((row + column) + plane), each pass (2x - previous) - next, zero
// outside the volume (C/Convolution3D_CPU_GPU.cu:146-148).
// One tile: 64 x 4 (x, y) voxels of plane k; returns the value written (0
// outside the volume).
__device__ __forceinline__ float laplacian_tile(const float* __restrict__ x, float* __restrict__ y, int nx, int ny,
                                                int nz, int bx, int by, int k) {
    const int i = bx * 64 + (threadIdx.x & 63);
    const int j = by * 4 + (threadIdx.x >> 6);
    if (i >= nx || j >= ny) return 0.0f;
    const size_t sy = (size_t)nx, sz = (size_t)nx * ny;
    const size_t o = (size_t)i + sy * j + sz * k;
    const float c = x[o], c2 = c + c;
    const float xm = i > 0 ? x[o - 1] : 0.0f, xp = i < nx - 1 ? x[o + 1] : 0.0f;
    const float ym = j > 0 ? x[o - sy] : 0.0f, yp = j < ny - 1 ? x[o + sy] : 0.0f;
    const float zm = k > 0 ? x[o - sz] : 0.0f, zp = k < nz - 1 ? x[o + sz] : 0.0f;
    const float h = (c2 - xm) - xp, v = (c2 - ym) - yp, p = (c2 - zm) - zp;
    const float r = (h + v) + p;
    y[o] = r;
    return r;
}

__global__ __launch_bounds__(kBlock) void laplacian_kernel(const float* __restrict__ x, float* __restrict__ y, int nx,
                                                         int ny, int nz) {
    (void)laplacian_tile(x, y, nx, ny, nz, blockIdx.x, blockIdx.y, blockIdx.z);
}

// Ellipsoid phantoms; 16 doubles per ellipsoid: rotation (9, row-major),
// centre (3), axis terms (3), density (1).
//   mode 0 (Shepp-Logan, Data3D_0256/phantom3d.m:76-107): r = normalised
//     voxel coordinates in [-1,1]; p = M r; q = sum (p_a - c_a)^2 / s_a^2
//     with axis terms s_a^2;
//   mode 1 (random ellipsoids): r = voxel offset from the volume centre;
//     l = R (r - c); q = sum l_a^2 * s_a with axis terms 1/s_a^2.
// A voxel takes the sum of the densities of the ellipsoids with q <= 1.
__global__ __launch_bounds__(kBlock) void ellipsoid_kernel(int nx, int ny, int nz, int mode, const double* __restrict__ tab,
                                                         int count, float* __restrict__ vol) {
    const int i = blockIdx.x * 64 + (threadIdx.x & 63);
    const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int k = blockIdx.z;
    if (i >= nx || j >= ny) return;
    const double hx = (nx - 1) / 2.0, hy = (ny - 1) / 2.0, hz = (nz - 1) / 2.0;
    double acc = 0.0;
    if (mode == 0) {
        const double x = ((double)i - hx) / hx, y = ((double)j - hy) / hy, z = ((double)k - hz) / hz;
        for (int e = 0; e < count; ++e) {
            const double* t = tab + 16 * e;
            const double px = (t[0] * x + t[1] * y) + t[2] * z;
            const double py = (t[3] * x + t[4] * y) + t[5] * z;
            const double pz = (t[6] * x + t[7] * y) + t[8] * z;
            const double tx = px - t[9], ty = py - t[10], tz = pz - t[11];
            const double q = ((tx * tx) / t[12] + (ty * ty) / t[13]) + (tz * tz) / t[14];
            if (q <= 1.0) acc += t[15];
        }
    } else {
        const double ox = (double)i - hx, oy = (double)j - hy, oz = (double)k - hz;
        for (int e = 0; e < count; ++e) {
            const double* t = tab + 16 * e;
            const double dx = ox - t[9], dy = oy - t[10], dz = oz - t[11];
            const double l0 = (t[0] * dx + t[1] * dy) + t[2] * dz;
            const double l1 = (t[3] * dx + t[4] * dy) + t[5] * dz;
            const double l2 = (t[6] * dx + t[7] * dy) + t[8] * dz;
            const double q = ((l0 * l0) * t[12] + (l1 * l1) * t[13]) + (l2 * l2) * t[14];
            if (q <= 1.0) acc += t[15];
        }
    }
    vol[(size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * k)] = (float)acc;
}

// Element conversion of the copies: bit copy, or FP32 to BF16 (nearest
// even) when the copy is BF16.
template <typename TO, typename TI>
__device__ __forceinline__ TO copy_cvt(TI v) {
    if constexpr (std::is_same<TO, TI>::value) return v;
    else return __bfloat16_as_ushort(__float2bfloat16(v));
}

// One z-pair (the two elements at an even index o) as one aligned store: the
// copies start on 256-byte boundaries, so out + o is aligned to the pair.
template <typename T>
__device__ __forceinline__ void store_pair(T* __restrict__ out, size_t o, T a, T b) {
    typedef T T2 __attribute__((ext_vector_type(2)));
    T2 v;
    v.x = a;
    v.y = b;
    *reinterpret_cast<T2*>(out + o) = v;
}

// z-pair copy for y slabs (A = y, B = x): B is also the input's fastest axis,
// a direct pass.  Grid (ceil((nx+2)/256), ny, nz+1).
template <typename TI, typename T>
__device__ __forceinline__ void zpair_y_tile(const TI* __restrict__ in, T* __restrict__ out, int nx, int ny, int nz,
                                             int bx, int y, int zz) {
    const int bb = bx * 256 + threadIdx.x;
    if (bb > nx + 1) return;
    const int x = bb - 1;
    T v[2];
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        const int z = zz - 1 + c;
        __builtin_memset(&v[c], 0, sizeof(T));
        if (x >= 0 && x < nx && z >= 0 && z < nz)
            v[c] = copy_cvt<T>(in[(size_t)x + (size_t)nx * ((size_t)y + (size_t)ny * z)]);
    }
    store_pair(out, zpair_index(y, zz, bb, nz, nx), v[0], v[1]);
}

// z-pair copy for x slabs (A = x, B = y): a 32x32 (x, B) tile transpose of
// two z planes through LDS.  Grid (ceil(nx/32), ceil((ny+2)/32), nz+1).
template <typename TI, typename T>
__device__ __forceinline__ void zpair_x_tile(const TI* __restrict__ in, T* __restrict__ out, int nx, int ny, int nz,
                                             int bx, int by, int zz, T (*tile)[32][33]) {
    const int x0 = bx * 32, b0 = by * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
    for (int r = ty; r < 32; r += 8) {
        const int x = x0 + tx, y = b0 + r - 1;
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const int z = zz - 1 + c;
            T v;
            __builtin_memset(&v, 0, sizeof(T));
            if (x < nx && y >= 0 && y < ny && z >= 0 && z < nz)
                v = copy_cvt<T>(in[(size_t)x + (size_t)nx * ((size_t)y + (size_t)ny * z)]);
            tile[c][r][tx] = v;
        }
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        const int x = x0 + r, bb = b0 + tx;
        if (x < nx && bb <= ny + 1) {
            store_pair(out, zpair_index(x, zz, bb, nz, ny), tile[0][tx][r], tile[1][tx][r]);
        }
    }
}

// Single-problem z-pair copy for y slabs: a thread walks kZChunk z pairs of
// one (y, bb) column, reading each input row once (a pair's upper voxel is
// the next pair's lower one), where a tile of one pair reads both.  Grid
// (ceil((nx+2)/256), ny, ceil((nz+1)/kZChunk)).
constexpr int kZChunk = 32;
template <typename T>
__global__ __launch_bounds__(256) void zpair_y_kernel(const T* __restrict__ in, T* __restrict__ out, int nx, int ny,
                                                    int nz) {
    const int bb = blockIdx.x * 256 + threadIdx.x;
    if (bb > nx + 1) return;
    const int x = bb - 1, y = blockIdx.y;
    const int z0 = blockIdx.z * kZChunk, z1 = min(z0 + kZChunk, nz + 1);
    const bool xin = x >= 0 && x < nx;
    const T* col = in + (size_t)(xin ? x : 0) + (size_t)nx * y;  // (x, y, z) at col[nx ny z]
    const size_t zs = (size_t)nx * ny;
    // Clamped address (always in the volume), then a mask (the element types
    // are integers): a conditional load, or a select the compiler turns into
    // one, is a branch per pair that waits for each load before the next.
    auto voxel = [&](int z) {
        const T v = col[zs * (size_t)min(max(z, 0), nz - 1)];
        return (T)(v & (T)(0u - (unsigned)(xin && z >= 0 && z < nz)));
    };
    T lo = voxel(z0 - 1);
#pragma unroll 8
    for (int zz = z0; zz < z1; ++zz) {
        const T hi = voxel(zz);
        store_pair(out, zpair_index(y, zz, bb, nz, nx), lo, hi);
        lo = hi;
    }
}

// out[y + nyp (x + nx z)] = in[x + nx (y + ny z)], 0 for ny <= y < nyp: the
// volume with y fastest and zero-padded rows, one 32 x 32 (x, y) tile of one
// z plane per workgroup, through LDS.
__global__ __launch_bounds__(256) void xy_transpose_kernel(const float* __restrict__ in, float* __restrict__ out,
                                                         int nx, int ny, int nyp) {
    __shared__ float tile[32][33];
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    const int x0 = blockIdx.x * 32, y0 = blockIdx.y * 32;
    const size_t plane_in = (size_t)nx * ny * blockIdx.z, plane_out = (size_t)nx * nyp * blockIdx.z;
    for (int r = ty; r < 32; r += 8) {
        const int x = x0 + tx, y = y0 + r;
        tile[r][tx] = (x < nx && y < ny) ? in[plane_in + (size_t)y * nx + x] : 0.0f;
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        const int y = y0 + tx, x = x0 + r;
        if (x < nx && y < nyp) out[plane_out + (size_t)x * nyp + y] = tile[tx][r];
    }
}

// out[x + nxp r] = in[x + nx r], 0 for nx <= x < nxp.
__global__ __launch_bounds__(256) void row_pad_kernel(const float* __restrict__ in, float* __restrict__ out, int nx,
                                                    int nxp, size_t n) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        const size_t r = i / (size_t)nxp;
        const int x = (int)(i - r * (size_t)nxp);
        out[i] = x < nx ? in[r * (size_t)nx + x] : 0.0f;
    }
}

// Single-problem z-pair copy for x slabs: a workgroup takes one 32 x 32
// (x, B) tile through kZChunk z pairs, each input plane of the tile read once
// into LDS (a pair's upper plane is the next pair's lower one); the next
// plane's loads are issued before the current pairs are written.  Grid
// (ceil(nx/32), ceil((ny+2)/32), ceil((nz+1)/kZChunk)).
template <typename T>
__global__ __launch_bounds__(256) void zpair_x_kernel(const T* __restrict__ in, T* __restrict__ out, int nx, int ny,
                                                    int nz) {
    __shared__ T tile[2][32][33];  // [plane][B row][x]
    const int x0 = blockIdx.x * 32, b0 = blockIdx.y * 32;
    const int z0 = blockIdx.z * kZChunk, z1 = min(z0 + kZChunk, nz + 1);
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
    const int x = x0 + tx;
    // rows r = ty + 8 j of plane z, as in zpair_y_kernel: clamped addresses
    // (the loads stay unconditional) and masks of the voxels in the grid,
    // applied when the rows go to LDS so that nothing waits for the loads
    // before then
    auto fetch = [&](int z, T (&v)[4], T (&m)[4]) {
        const int zc = min(max(z, 0), nz - 1);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int y = b0 + ty + 8 * j - 1;
            const bool ok = x < nx && y >= 0 && y < ny && z >= 0 && z < nz;
            const size_t o = (size_t)min(x, nx - 1) + (size_t)nx * ((size_t)min(max(y, 0), ny - 1) + (size_t)ny * zc);
            v[j] = in[o];
            m[j] = (T)(0u - (unsigned)ok);
        }
    };
    auto put = [&](const T (&v)[4], const T (&m)[4], int c) {
#pragma unroll
        for (int j = 0; j < 4; ++j) tile[c][ty + 8 * j][tx] = (T)(v[j] & m[j]);
    };
    T v[4], m[4];
    fetch(z0 - 1, v, m);
    put(v, m, 0);
    fetch(z0, v, m);
    for (int zz = z0; zz < z1; ++zz) {
        const int cur = (zz - z0 + 1) & 1, prev = cur ^ 1;  // planes zz and zz - 1
        put(v, m, cur);
        __syncthreads();
        if (zz + 1 < z1) fetch(zz + 1, v, m);
        for (int r = ty; r < 32; r += 8) {
            const int xo = x0 + r, bb = b0 + tx;
            if (xo < nx && bb <= ny + 1)
                store_pair(out, zpair_index(xo, zz, bb, nz, ny), tile[prev][tx][r], tile[cur][tx][r]);
        }
        __syncthreads();  // every pair of this z is written before its lower plane is overwritten
    }
}

// One A-pair group (four 16-bit elements at an index that is a multiple of
// four) as one aligned 8-byte store.
__device__ __forceinline__ void store_group(uint16_t* __restrict__ out, size_t o, uint16_t a, uint16_t b, uint16_t c,
                                            uint16_t d) {
    typedef uint16_t U4 __attribute__((ext_vector_type(4)));
    U4 v;
    v.x = a;
    v.y = b;
    v.z = c;
    v.w = d;
    *reinterpret_cast<U4*>(out + o) = v;
}

// A voxel of an x-fastest volume as a 16-bit copy element (0 outside the
// grid): a bit copy of a 16-bit volume, or an FP32 voxel rounded to BF16.
// Read at the clamped address (always in the volume), then masked: a load
// guarded by the bounds test is a branch, which waits for it before the
// next.  The empty statement keeps the compiler from seeing that the mask is
// 0 or ~0 (it would turn the AND back into a select, and that into the
// branch); it emits no instruction.
template <typename TI>
__device__ __forceinline__ uint16_t voxel16(const TI* __restrict__ in, int x, int y, int z, int nx, int ny, int nz) {
    const bool ok = (x >= 0) & (x < nx) & (y >= 0) & (y < ny) & (z >= 0) & (z < nz);
    const size_t o = (size_t)min(max(x, 0), nx - 1) +
                     (size_t)nx * ((size_t)min(max(y, 0), ny - 1) + (size_t)ny * min(max(z, 0), nz - 1));
    uint32_t m = 0u - (uint32_t)ok;
    asm("" : "+v"(m));
    return (uint16_t)(copy_cvt<uint16_t>(in[o]) & m);
}

// A-pair copy for x slabs (A = x, B = y): a tile of 32 plane pairs x 32 B
// columns of one z pair through LDS (the input runs along x, the copy along
// y).  Grid (ceil((nx+1)/32), ceil((ny+2)/32), nz+1).
template <typename TI>
__device__ __forceinline__ void apair_x_tile(const TI* __restrict__ in, uint16_t* __restrict__ out, int nx, int ny,
                                             int nz, int bx, int by, int zz, uint16_t (*tile)[32][34]) {
    // tile[z - (zz-1)][bb - b0][x - (ap0-1)]
    const int ap0 = bx * 32, b0 = by * 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
    for (int r = ty; r < 32; r += 8) {
        const int y = b0 + r - 1;
        for (int c = tx; c < 34; c += 32) {
            tile[0][r][c] = voxel16(in, ap0 - 1 + c, y, zz - 1, nx, ny, nz);
            tile[1][r][c] = voxel16(in, ap0 - 1 + c, y, zz, nx, ny, nz);
        }
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        const int ap = ap0 + r, bb = b0 + tx;
        if (ap <= nx && bb <= ny + 1)
            store_group(out, apair_index(ap, zz, bb, nz, ny), tile[0][tx][r], tile[1][tx][r], tile[0][tx][r + 1],
                        tile[1][tx][r + 1]);
    }
}

// A-pair copy for y slabs (A = y, B = x): B is the input's fastest axis, a
// direct pass.  Grid (ceil((nx+2)/256), ny+1, nz+1): group (ap, zz, bb) with
// bb = 256 bx + thread.
template <typename TI>
__device__ __forceinline__ void apair_y_tile(const TI* __restrict__ in, uint16_t* __restrict__ out, int nx, int ny,
                                             int nz, int bx, int ap, int zz) {
    const int bb = bx * 256 + threadIdx.x;
    if (bb > nx + 1) return;
    const int x = bb - 1;
    store_group(out, apair_index(ap, zz, bb, nz, nx), voxel16(in, x, ap - 1, zz - 1, nx, ny, nz),
                voxel16(in, x, ap - 1, zz, nx, ny, nz), voxel16(in, x, ap, zz - 1, nx, ny, nz),
                voxel16(in, x, ap, zz, nx, ny, nz));
}

__global__ __launch_bounds__(256) void apair_x_kernel(const uint16_t* __restrict__ in, uint16_t* __restrict__ out,
                                                    int nx, int ny, int nz) {
    __shared__ uint16_t tile[2][32][34];
    apair_x_tile<uint16_t>(in, out, nx, ny, nz, blockIdx.x, blockIdx.y, blockIdx.z, tile);
}

__global__ __launch_bounds__(256) void apair_y_kernel(const uint16_t* __restrict__ in, uint16_t* __restrict__ out,
                                                    int nx, int ny, int nz) {
    apair_y_tile<uint16_t>(in, out, nx, ny, nz, blockIdx.x, blockIdx.y, blockIdx.z);
}

// ---- shared launches of the gradient loop (FP32).  Element-wise tiles
// cover kVecTile consecutive elements, 16 per thread at stride 256; FP64
// partial sums are per tile, in that fixed order, so a problem's sums do not
// depend on the batch it runs in.

// Block sum for kernels that loop over tiles: as block_sum, with a trailing
// barrier so the next tile may reuse the shared slots.
__device__ __forceinline__ double block_sum_loop(double v) {
    const double r = block_sum(v);
    __syncthreads();
    return r;
}

__device__ __forceinline__ void batch_coords(const BatchItem& it, uint32_t tile, int& bx, int& by, int& bz) {
    const uint32_t local = tile - it.begin;
    bx = (int)(local % it.gx);
    const uint32_t rest = local / it.gx;
    by = (int)(rest % it.gy);
    bz = (int)(rest / it.gy);
}

// z-pair copies of FP32 volumes: FP32 (bit copy) or BF16 (kItemBF16), or
// BF16 A-pair copies (kItemApair, the item's grid is then the A-pair one).
__global__ __launch_bounds__(256) void zpair_x_batch_kernel(const BatchTable* __restrict__ t) {
    __shared__ uint32_t tile_lds[2][32][33];
    __shared__ uint16_t tile_lds16[2][32][33];
    __shared__ uint16_t tile_ap[2][32][34];
    TOMO_BATCH_STRIDED(t) {
        const BatchItem& it = t->item[batch_find(t, tile)];
        int bx, by, bz;
        batch_coords(it, tile, bx, by, bz);
        if (it.flags & kItemApair)
            apair_x_tile<float>((const float*)it.in, (uint16_t*)it.out, it.g.nx, it.g.ny, it.g.nz, bx, by, bz,
                                tile_ap);
        else if (it.flags & kItemBF16)
            zpair_x_tile<float, uint16_t>((const float*)it.in, (uint16_t*)it.out, it.g.nx, it.g.ny, it.g.nz, bx, by,
                                          bz, tile_lds16);
        else
            zpair_x_tile<uint32_t, uint32_t>((const uint32_t*)it.in, (uint32_t*)it.out, it.g.nx, it.g.ny, it.g.nz, bx,
                                             by, bz, tile_lds);
        __syncthreads();
    }
}

// A tile is a whole (y, z pair) row of nx + 2 pairs, in chunks of 256 (the
// single-problem grid's tiles of one row): tiles of one chunk left half the
// launch's workgroups nearly empty at 256^3 (258 pairs a row) and paid the
// table lookup for each.
__global__ __launch_bounds__(256) void zpair_y_batch_kernel(const BatchTable* __restrict__ t) {
    TOMO_BATCH_STRIDED(t) {
        const BatchItem& it = t->item[batch_find(t, tile)];
        int bx, by, bz;
        batch_coords(it, tile, bx, by, bz);
        const int nx = it.g.nx, ny = it.g.ny, nz = it.g.nz;
        for (int c = 0; c * 256 < nx + 2; ++c) {
            if (it.flags & kItemApair)  // by: the plane pair
                apair_y_tile<float>((const float*)it.in, (uint16_t*)it.out, nx, ny, nz, c, by, bz);
            else if (it.flags & kItemBF16)
                zpair_y_tile<float, uint16_t>((const float*)it.in, (uint16_t*)it.out, nx, ny, nz, c, by, bz);
            else
                zpair_y_tile<uint32_t, uint32_t>((const uint32_t*)it.in, (uint32_t*)it.out, nx, ny, nz, c, by, bz);
        }
    }
}

// d = in - in2; out = d (or d * in3 when in3 is set); aux[tile] = sum d^2.
__global__ __launch_bounds__(kBlock) void diff_sq_batch_kernel(const BatchTable* __restrict__ t) {
    TOMO_BATCH_STRIDED(t) {
        const BatchItem& it = t->item[batch_find(t, tile)];
        const uint32_t local = tile - it.begin;
        const float* a = (const float*)it.in;
        const float* b = (const float*)it.in2;
        const float* w = (const float*)it.in3;
        float* o = (float*)it.out;
        const uint64_t base = (uint64_t)local * kVecTile;
        double sq = 0.0;
#pragma unroll 4
        for (int k = 0; k < kVecTile / kBlock; ++k) {
            const uint64_t e = base + (uint64_t)k * kBlock + threadIdx.x;
            if (e < it.n) {
                const float v = a[e] - b[e];
                o[e] = w ? v * w[e] : v;
                sq += (double)v * (double)v;
            }
        }
        const double r = block_sum_loop(sq);
        if (threadIdx.x == 0) ((double*)it.aux)[local] = r;
    }
}

// aux[tile] = sum in^2.
__global__ __launch_bounds__(kBlock) void sq_batch_kernel(const BatchTable* __restrict__ t) {
    TOMO_BATCH_STRIDED(t) {
        const BatchItem& it = t->item[batch_find(t, tile)];
        const uint32_t local = tile - it.begin;
        const float* a = (const float*)it.in;
        const uint64_t base = (uint64_t)local * kVecTile;
        double sq = 0.0;
#pragma unroll 4
        for (int k = 0; k < kVecTile / kBlock; ++k) {
            const uint64_t e = base + (uint64_t)k * kBlock + threadIdx.x;
            if (e < it.n) {
                const double v = (double)a[e];
                sq += v * v;
            }
        }
        const double r = block_sum_loop(sq);
        if (threadIdx.x == 0) ((double*)it.aux)[local] = r;
    }
}

// out = L in over tiles of 64 x 4 voxels of one plane; aux[tile] = sum out^2
// when aux is set.
__global__ __launch_bounds__(kBlock) void laplacian_batch_kernel(const BatchTable* __restrict__ t) {
    TOMO_BATCH_STRIDED(t) {
        const BatchItem& it = t->item[batch_find(t, tile)];
        int bx, by, bz;
        batch_coords(it, tile, bx, by, bz);
        const float v = laplacian_tile((const float*)it.in, (float*)it.out, it.g.nx, it.g.ny, it.g.nz, bx, by, bz);
        if (it.aux) {
            const double r = block_sum_loop((double)v * (double)v);
            if (threadIdx.x == 0) ((double*)it.aux)[tile - it.begin] = r;
        }
    }
}

// out = out + a * in (multiply then add, as the reference host loop), a =
// (float) slot[alpha] when flags bit 0 is set, else the item's constant;
// flags bit 1 clamps the result at 0; flags bit 2 writes out = a instead;
// with in3 set, a is the element in3[e] (the SIRT volume weight).
__global__ __launch_bounds__(kBlock) void axpy_batch_kernel(const BatchTable* __restrict__ t) {
    TOMO_BATCH_STRIDED(t) {
        const BatchItem& it = t->item[batch_find(t, tile)];
        const uint32_t local = tile - it.begin;
        const float a = (it.flags & 1) ? (float)it.slot[kSlotAlpha] : it.a;
        const bool pos = (it.flags & 2) != 0;
        const bool fill = (it.flags & 4) != 0;
        const float* x = (const float*)it.in;
        const float* w = (const float*)it.in3;
        float* y = (float*)it.out;
        const uint64_t base = (uint64_t)local * kVecTile;
        for (int k = 0; k < kVecTile / kBlock; ++k) {
            const uint64_t e = base + (uint64_t)k * kBlock + threadIdx.x;
            if (e < it.n) {
                float v = fill ? a : y[e] + (w ? w[e] : a) * x[e];
                if (pos && v < 0.0f) v = 0.0f;
                y[e] = v;
            }
        }
    }
}

__global__ __launch_bounds__(kBlock) void fill_batch_kernel(const BatchTable* __restrict__ t) {
    TOMO_BATCH_STRIDED(t) {
        const BatchItem& it = t->item[batch_find(t, tile)];
        const uint32_t local = tile - it.begin;
        float* y = (float*)it.out;
        const uint64_t base = (uint64_t)local * kVecTile;
        for (int k = 0; k < kVecTile / kBlock; ++k) {
            const uint64_t e = base + (uint64_t)k * kBlock + threadIdx.x;
            if (e < it.n) y[e] = it.a;
        }
    }
}

// slot[flags] = sum of the n partials in aux (one tile per item).
__global__ __launch_bounds__(kBlock) void finish_batch_kernel(const BatchTable* __restrict__ t) {
    TOMO_BATCH_STRIDED(t) {
        const BatchItem& it = t->item[batch_find(t, tile)];
        const double* part = (const double*)it.aux;
        double s = 0.0;
        for (uint64_t i = threadIdx.x; i < it.n; i += kBlock) s += part[i];
        const double r = block_sum_loop(s);
        if (threadIdx.x == 0) it.slot[it.flags] = r;
    }
}

// Step rule of the simple-gradient loop on the problem's scalars (one tile
// per item; d = lambda, flags bit 0 = optimal step this iteration).  aux,
// when set, receives (residual, norm dJ, norm dJ proj, alpha) at entry n.
__global__ __launch_bounds__(64) void step_batch_kernel(const BatchTable* __restrict__ t) {
    TOMO_BATCH_STRIDED(t) {
        const BatchItem& it = t->item[batch_find(t, tile)];
        if (threadIdx.x != 0) continue;
        double* sl = it.slot;
        if (it.flags & 2) {  // SIRT: record the residual only
            if (it.aux) {
                double* h = (double*)it.aux + 4 * it.n;
                h[0] = sl[kSlotResidual];
                h[1] = h[2] = h[3] = 0.0;
            }
            continue;
        }
        double alpha;
        if (it.flags & 1) {
            const double norm_reg = it.d != 0.0 ? 2.0 * sl[kSlotRegSq] : 0.0;
            const double ndp = 2.0 * sl[kSlotProjSq] + it.d * norm_reg;
            alpha = -sl[kSlotNormDj] / ndp;
            sl[kSlotAlphaMin] = fmax(alpha, sl[kSlotAlphaMin]);
            sl[kSlotNormDjProj] = ndp;
        } else {
            alpha = sl[kSlotAlphaMin];
            sl[kSlotNormDj] = 0.0;
            sl[kSlotNormDjProj] = 0.0;
        }
        sl[kSlotAlpha] = alpha;
        if (it.aux) {
            double* h = (double*)it.aux + 4 * it.n;
            h[0] = sl[kSlotResidual];
            h[1] = sl[kSlotNormDj];
            h[2] = sl[kSlotNormDjProj];
            h[3] = alpha;
        }
    }
}

unsigned grid_for(size_t n) {
    size_t b = (n + kBlock - 1) / kBlock;
    if (b > 65536) b = 65536;
    if (b == 0) b = 1;
    return (unsigned)b;
}

}  // namespace

hipError_t launch_apair(const void* vol, void* qx, void* qy, int nx, int ny, int nz, hipStream_t s) {
    const dim3 gx(ceil_div((size_t)nx + 1, 32), ceil_div((size_t)ny + 2, 32), nz + 1);
    const dim3 gy(ceil_div((size_t)nx + 2, 256), ny + 1, nz + 1);
    hipLaunchKernelGGL(apair_x_kernel, gx, dim3(256), 0, s, (const uint16_t*)vol, (uint16_t*)qx, nx, ny, nz);
    hipLaunchKernelGGL(apair_y_kernel, gy, dim3(256), 0, s, (const uint16_t*)vol, (uint16_t*)qy, nx, ny, nz);
    return hipGetLastError();
}

hipError_t launch_zpair(const void* vol, void* qx, void* qy, size_t es, int nx, int ny, int nz, hipStream_t s) {
    const dim3 gx(ceil_div(nx, 32), ceil_div((size_t)ny + 2, 32), ceil_div((size_t)nz + 1, kZChunk));
    const dim3 gy(ceil_div((size_t)nx + 2, 256), ny, ceil_div((size_t)nz + 1, kZChunk));
    if (es == 4) {
        hipLaunchKernelGGL((zpair_x_kernel<uint32_t>), gx, dim3(256), 0, s, (const uint32_t*)vol, (uint32_t*)qx, nx,
                           ny, nz);
        hipLaunchKernelGGL((zpair_y_kernel<uint32_t>), gy, dim3(256), 0, s, (const uint32_t*)vol, (uint32_t*)qy, nx,
                           ny, nz);
    } else {
        hipLaunchKernelGGL((zpair_x_kernel<uint16_t>), gx, dim3(256), 0, s, (const uint16_t*)vol, (uint16_t*)qx, nx,
                           ny, nz);
        hipLaunchKernelGGL((zpair_y_kernel<uint16_t>), gy, dim3(256), 0, s, (const uint16_t*)vol, (uint16_t*)qy, nx,
                           ny, nz);
    }
    return hipGetLastError();
}

hipError_t launch_xy_transpose(const float* vol, float* out, int nx, int ny, int nyp, int nz, hipStream_t s) {
    const dim3 grid(ceil_div(nx, 32), ceil_div(nyp, 32), nz);
    hipLaunchKernelGGL(xy_transpose_kernel, grid, dim3(256), 0, s, vol, out, nx, ny, nyp);
    return hipGetLastError();
}

hipError_t launch_row_pad(const float* vol, float* out, int nx, int nxp, size_t rows, hipStream_t s) {
    const size_t n = rows * (size_t)nxp;
    hipLaunchKernelGGL(row_pad_kernel, dim3(grid_for(n)), dim3(256), 0, s, vol, out, nx, nxp, n);
    return hipGetLastError();
}

hipError_t launch_batch_zpair_x(const BatchTable* t, hipStream_t s) {
    hipLaunchKernelGGL(zpair_x_batch_kernel, dim3(kBatchGrid), dim3(256), 0, s, t);
    return hipGetLastError();
}
hipError_t launch_batch_zpair_y(const BatchTable* t, hipStream_t s) {
    hipLaunchKernelGGL(zpair_y_batch_kernel, dim3(kBatchGrid), dim3(256), 0, s, t);
    return hipGetLastError();
}
hipError_t launch_batch_diff_sq(const BatchTable* t, hipStream_t s) {
    hipLaunchKernelGGL(diff_sq_batch_kernel, dim3(kBatchGrid), dim3(kBlock), 0, s, t);
    return hipGetLastError();
}
hipError_t launch_batch_sq(const BatchTable* t, hipStream_t s) {
    hipLaunchKernelGGL(sq_batch_kernel, dim3(kBatchGrid), dim3(kBlock), 0, s, t);
    return hipGetLastError();
}
hipError_t launch_batch_laplacian(const BatchTable* t, hipStream_t s) {
    hipLaunchKernelGGL(laplacian_batch_kernel, dim3(kBatchGrid), dim3(kBlock), 0, s, t);
    return hipGetLastError();
}
hipError_t launch_batch_axpy(const BatchTable* t, hipStream_t s) {
    hipLaunchKernelGGL(axpy_batch_kernel, dim3(kBatchGrid), dim3(kBlock), 0, s, t);
    return hipGetLastError();
}
hipError_t launch_batch_fill(const BatchTable* t, hipStream_t s) {
    hipLaunchKernelGGL(fill_batch_kernel, dim3(kBatchGrid), dim3(kBlock), 0, s, t);
    return hipGetLastError();
}
hipError_t launch_batch_finish(const BatchTable* t, hipStream_t s) {
    hipLaunchKernelGGL(finish_batch_kernel, dim3(kBatchMax), dim3(kBlock), 0, s, t);
    return hipGetLastError();
}
hipError_t launch_batch_step(const BatchTable* t, hipStream_t s) {
    hipLaunchKernelGGL(step_batch_kernel, dim3(kBatchMax), dim3(64), 0, s, t);
    return hipGetLastError();
}

int launch_fill(float* x, float v, size_t n, hipStream_t s) {
    hipLaunchKernelGGL(fill_kernel, dim3(grid_for(n)), dim3(kBlock), 0, s, x, v, n);
    TOMO_HIP(hipGetLastError());
    return TOMO_OK;
}
int launch_diff(const float* a, const float* b, float* out, size_t n, hipStream_t s) {
    hipLaunchKernelGGL(diff_kernel, dim3(grid_for(n)), dim3(kBlock), 0, s, a, b, out, n);
    TOMO_HIP(hipGetLastError());
    return TOMO_OK;
}
int launch_scale(float* x, float v, size_t n, hipStream_t s) {
    hipLaunchKernelGGL(scale_kernel, dim3(grid_for(n)), dim3(kBlock), 0, s, x, v, n);
    TOMO_HIP(hipGetLastError());
    return TOMO_OK;
}
int launch_axpy(float* y, const float* x, float a, int positive, size_t n, hipStream_t s) {
    hipLaunchKernelGGL(axpy_kernel, dim3(grid_for(n)), dim3(kBlock), 0, s, y, x, a, positive, n);
    TOMO_HIP(hipGetLastError());
    return TOMO_OK;
}
int launch_recip_weight(const float* s_in, float a, float* w, size_t n, hipStream_t s) {
    hipLaunchKernelGGL(recip_weight_kernel, dim3(grid_for(n)), dim3(kBlock), 0, s, s_in, a, w, n);
    TOMO_HIP(hipGetLastError());
    return TOMO_OK;
}
int launch_mul(float* x, const float* w, size_t n, hipStream_t s) {
    hipLaunchKernelGGL(mul_kernel, dim3(grid_for(n)), dim3(kBlock), 0, s, x, w, n);
    TOMO_HIP(hipGetLastError());
    return TOMO_OK;
}
int launch_waxpy(float* x, const float* c, const float* b, int positive, size_t n, hipStream_t s) {
    hipLaunchKernelGGL(waxpy_kernel, dim3(grid_for(n)), dim3(kBlock), 0, s, x, c, b, positive, n);
    TOMO_HIP(hipGetLastError());
    return TOMO_OK;
}
int launch_convert(Buf in, MutBuf out, size_t n, hipStream_t s) {
    const float inv = 1.0f / in.scale;
    TOMO_DISPATCH_STORAGE(in.dtype, TI, {
        TOMO_DISPATCH_STORAGE(out.dtype, TO, {
            hipLaunchKernelGGL((convert_kernel<TI, TO>), dim3(grid_for(n)), dim3(kBlock), 0, s, (const TI*)in.p, inv,
                               (TO*)out.p, out.scale, n);
        });
    });
    TOMO_HIP(hipGetLastError());
    return TOMO_OK;
}

int launch_laplacian(const float* x, float* y, int nx, int ny, int nz, hipStream_t s) {
    dim3 grid(ceil_div(nx, 64), ceil_div(ny, 4), nz);
    hipLaunchKernelGGL(laplacian_kernel, grid, dim3(kBlock), 0, s, x, y, nx, ny, nz);
    TOMO_HIP(hipGetLastError());
    return TOMO_OK;
}

// Each reduction takes its partial sums from the stream-ordered pool for the
// duration of the call, so calls from several threads on one context (each
// on its own stream) never share them.  The final value is copied to the
// host and the stream synchronised before the call returns.
static int finish_reduce(double* part, int blocks, double* out, hipStream_t s) {
    hipLaunchKernelGGL(final_sum_kernel, dim3(1), dim3(kBlock), 0, s, part, blocks, part + kReduceBlocks);
    hipError_t e = hipGetLastError();
    double h = 0.0;
    if (e == hipSuccess) e = hipMemcpyAsync(&h, part + kReduceBlocks, sizeof(double), hipMemcpyDeviceToHost, s);
    const hipError_t f = hipFreeAsync(part, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    TOMO_HIP(e);
    TOMO_HIP(f);
    *out = h;
    return TOMO_OK;
}

// Per-block partials of sum double(x)^2 into part[0, blocks).
static int launch_sq_partial(Buf x, size_t n, unsigned blocks, double* part, hipStream_t s) {
    const float inv = 1.0f / x.scale;
    switch (x.dtype) {
        case TOMO_F32:
            hipLaunchKernelGGL((sq_partial_kernel<float>), dim3(blocks), dim3(kBlock), 0, s, (const float*)x.p, inv,
                               n, part);
            break;
        case TOMO_F16:
            hipLaunchKernelGGL((sq_partial_kernel<__half>), dim3(blocks), dim3(kBlock), 0, s, (const __half*)x.p,
                               inv, n, part);
            break;
        default:
            hipLaunchKernelGGL((sq_partial_kernel<__hip_bfloat16>), dim3(blocks), dim3(kBlock), 0, s,
                               (const __hip_bfloat16*)x.p, inv, n, part);
            break;
    }
    TOMO_HIP(hipGetLastError());
    return TOMO_OK;
}

static unsigned reduce_blocks(size_t n) {
    const unsigned b = grid_for(n);
    return b > (unsigned)kReduceBlocks ? (unsigned)kReduceBlocks : b;
}

int reduce_sq_norm(tomo_ctx* ctx, Buf x, size_t n, double* out, hipStream_t s) {
    (void)ctx;
    double* part = nullptr;
    TOMO_HIP(hipMallocAsync((void**)&part, sizeof(double) * (kReduceBlocks + 1), s));
    const unsigned blocks = reduce_blocks(n);
    const int rc = launch_sq_partial(x, n, blocks, part, s);
    if (rc) {
        (void)hipFreeAsync(part, s);
        return rc;
    }
    return finish_reduce(part, (int)blocks, out, s);
}

// ---- allocation-free squared norm (tomo_sq_norm_noalloc) ------------------
// The same kernels as reduce_sq_norm, with the partial sums and the final sum
// in `scratch`, a file-scope device array of kScratchDoubles doubles (16 KiB)
// that every GPU the library runs on holds until the process ends.  In host
// code the array's name is not a device address: noalloc_scratch_init, called
// once per device with that device current (tomo_ctx_create), looks the
// address up with hipGetSymbolAddress and keeps it per device, and only that
// pointer is ever passed to a kernel or a copy.  Every call on one GPU uses
// that GPU's one array, so calls that overlap on a GPU (two streams, two
// threads) race on it; nothing in it needs clearing, since every partial the
// final sum reads is written by the same call.
constexpr int kScratchDoubles = kReduceBlocks + 1;
static __device__ __align__(256) double scratch[kScratchDoubles];
constexpr int kMaxScratchDevices = 64;
static std::mutex g_scratch_mu;
static double* g_scratch_ptr[kMaxScratchDevices];

int noalloc_scratch_init(int device) {
    if (device < 0 || device >= kMaxScratchDevices)
        return fail(TOMO_ERR_UNSUPPORTED, "device index beyond the allocation-free scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return TOMO_OK;
    void* p = nullptr;
    TOMO_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(TOMO_ERR_HIP, "hipGetSymbolAddress returned no address for the reduction scratch");
    g_scratch_ptr[device] = static_cast<double*>(p);
    return TOMO_OK;
}

// The final sum copied to the host, the stream synchronised.
static int copy_result(const double* d, double* out, hipStream_t s) {
    double h = 0.0;
    TOMO_HIP(hipMemcpyAsync(&h, d, sizeof(double), hipMemcpyDeviceToHost, s));
    TOMO_HIP(hipStreamSynchronize(s));
    *out = h;
    return TOMO_OK;
}

int reduce_sq_norm_noalloc(int device, Buf x, size_t n, double* out, hipStream_t s) {
    double* part = nullptr;
    if (device >= 0 && device < kMaxScratchDevices) {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        part = g_scratch_ptr[device];
    }
    if (!part) return fail(TOMO_ERR_INVALID_ARG, "allocation-free scratch not set up on this device");
    const unsigned blocks = reduce_blocks(n);
    // blocks partials and the final sum; reduce_blocks keeps this within the array for every n.
    if ((size_t)blocks + 1 > (size_t)kScratchDoubles)
        return fail(TOMO_ERR_INVALID_ARG, "reduction needs more partial sums than the scratch holds");
    int rc = launch_sq_partial(x, n, blocks, part, s);
    if (rc) return rc;
    hipLaunchKernelGGL(final_sum_kernel, dim3(1), dim3(kBlock), 0, s, part, (int)blocks, part + kReduceBlocks);
    TOMO_HIP(hipGetLastError());
    return copy_result(part + kReduceBlocks, out, s);
}

int reduce_dot(tomo_ctx* ctx, const float* a, const float* b, size_t n, double* out, hipStream_t s) {
    (void)ctx;
    double* part = nullptr;
    TOMO_HIP(hipMallocAsync((void**)&part, sizeof(double) * (kReduceBlocks + 1), s));
    unsigned blocks = grid_for(n);
    if (blocks > (unsigned)kReduceBlocks) blocks = kReduceBlocks;
    hipLaunchKernelGGL(dot_partial_kernel, dim3(blocks), dim3(kBlock), 0, s, a, b, n, part);
    return finish_reduce(part, (int)blocks, out, s);
}

int reduce_max_abs(const float* x, size_t n, float* out, hipStream_t s) {
    float* part = nullptr;
    TOMO_HIP(hipMallocAsync((void**)&part, sizeof(float) * (kReduceBlocks + 1), s));
    unsigned blocks = grid_for(n);
    if (blocks > (unsigned)kReduceBlocks) blocks = kReduceBlocks;
    hipLaunchKernelGGL(max_abs_partial_kernel, dim3(blocks), dim3(kBlock), 0, s, x, n, part);
    hipLaunchKernelGGL(final_max_kernel, dim3(1), dim3(kBlock), 0, s, part, (int)blocks, part + kReduceBlocks);
    hipError_t e = hipGetLastError();
    float h = 0.0f;
    if (e == hipSuccess) e = hipMemcpyAsync(&h, part + kReduceBlocks, sizeof(float), hipMemcpyDeviceToHost, s);
    const hipError_t f = hipFreeAsync(part, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    TOMO_HIP(e);
    TOMO_HIP(f);
    *out = h;
    return TOMO_OK;
}

int launch_ellipsoids(const tomo_geom* g, int mode, const double* d_table, int count, float* vol, hipStream_t s) {
    const DevGeom& d = g->dev;
    dim3 grid(ceil_div(d.nx, 64), ceil_div(d.ny, 4), d.nz);
    hipLaunchKernelGGL(ellipsoid_kernel, grid, dim3(kBlock), 0, s, d.nx, d.ny, d.nz, mode, d_table, count, vol);
    TOMO_HIP(hipGetLastError());
    return TOMO_OK;
}

}  // namespace tomo
