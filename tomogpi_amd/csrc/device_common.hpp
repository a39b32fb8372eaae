// device_common.hpp - device helpers shared by the kernel files.
//
// Arithmetic contract: every file is compiled with -ffp-contract=off and
// correctly rounded FP32 division / square root, so each fused multiply-add
// below is an explicit fmaf()/fma() call.  The CPU oracle restates the same
// operations in the same order, which makes the FP32 forms bit-comparable.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <hip/hip_bf16.h>

#include <cstdint>

#include "tomo_internal.hpp"

namespace tomo {

// ---- storage types: value = (float)stored * inv_scale; stored = value*scale
// (FP32 storage is unscaled).  load4 reads elements i..i+3 with one load:
// 16 bytes for FP32, 8 for the 16-bit types; i must be even, which with the
// 256-byte aligned pool buffers keeps the load naturally aligned.
template <typename T> struct Storage;
template <> struct Storage<float> {
    __device__ __forceinline__ static float load(const float* p, uint64_t i, float) { return p[i]; }
    __device__ __forceinline__ static float4 load4(const float* p, uint64_t i, float) {
        float4 v;
        __builtin_memcpy(&v, p + i, sizeof(v));
        return v;
    }
    __device__ __forceinline__ static void store(float* p, uint64_t i, float v, float) { p[i] = v; }
};
template <> struct Storage<__half> {
    __device__ __forceinline__ static float load(const __half* p, uint64_t i, float inv) {
        return __half2float(p[i]) * inv;
    }
    __device__ __forceinline__ static float4 load4(const __half* p, uint64_t i, float inv) {
        __half h[4];
        __builtin_memcpy(h, p + i, sizeof(h));
        return make_float4(__half2float(h[0]) * inv, __half2float(h[1]) * inv, __half2float(h[2]) * inv,
                           __half2float(h[3]) * inv);
    }
    __device__ __forceinline__ static void store(__half* p, uint64_t i, float v, float scale) {
        p[i] = __float2half_rn(v * scale);
    }
};
template <> struct Storage<__hip_bfloat16> {
    __device__ __forceinline__ static float load(const __hip_bfloat16* p, uint64_t i, float inv) {
        return __bfloat162float(p[i]) * inv;
    }
    __device__ __forceinline__ static float4 load4(const __hip_bfloat16* p, uint64_t i, float inv) {
        unsigned short h[4];
        __builtin_memcpy(h, p + i, sizeof(h));
        return make_float4(__uint_as_float((unsigned)h[0] << 16) * inv, __uint_as_float((unsigned)h[1] << 16) * inv,
                           __uint_as_float((unsigned)h[2] << 16) * inv, __uint_as_float((unsigned)h[3] << 16) * inv);
    }
    __device__ __forceinline__ static void store(__hip_bfloat16* p, uint64_t i, float v, float scale) {
        p[i] = __float2bfloat16(v * scale);
    }
};

// Raw (untyped) buffer access: descriptor word 3 for gfx9-family parts.  A
// buffer resource covers `bytes` bytes (below 4 GiB) from `p`; loads take a
// per-lane 32-bit byte offset plus a wave-uniform one, and read 0 past the
// end.
constexpr int kBufferWord3 = 0x00020000;

__device__ __forceinline__ __amdgpu_buffer_rsrc_t buffer_rsrc(const void* p, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc((void*)p, (short)0, (int)bytes, kBufferWord3);
}

// Four consecutive elements at a byte offset, as float (value * inv).
template <typename T> struct BufLoad4;
template <> struct BufLoad4<float> {
    __device__ __forceinline__ static float4 load(__amdgpu_buffer_rsrc_t r, uint32_t off, uint32_t soff, float) {
        const auto v = __builtin_amdgcn_raw_buffer_load_b128(r, (int)off, (int)soff, 0);
        return make_float4(__uint_as_float(v[0]), __uint_as_float(v[1]), __uint_as_float(v[2]),
                           __uint_as_float(v[3]));
    }
};
template <> struct BufLoad4<__half> {
    __device__ __forceinline__ static float4 load(__amdgpu_buffer_rsrc_t r, uint32_t off, uint32_t soff, float inv) {
        const auto v = __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, (int)soff, 0);
        const __half h0 = __builtin_bit_cast(__half, (unsigned short)(v[0] & 0xffffu));
        const __half h1 = __builtin_bit_cast(__half, (unsigned short)(v[0] >> 16));
        const __half h2 = __builtin_bit_cast(__half, (unsigned short)(v[1] & 0xffffu));
        const __half h3 = __builtin_bit_cast(__half, (unsigned short)(v[1] >> 16));
        return make_float4(__half2float(h0) * inv, __half2float(h1) * inv, __half2float(h2) * inv,
                           __half2float(h3) * inv);
    }
};
template <> struct BufLoad4<__hip_bfloat16> {
    __device__ __forceinline__ static float4 load(__amdgpu_buffer_rsrc_t r, uint32_t off, uint32_t soff, float inv) {
        const auto v = __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, (int)soff, 0);
        return make_float4(__uint_as_float(v[0] << 16) * inv, __uint_as_float(v[0] & 0xffff0000u) * inv,
                           __uint_as_float(v[1] << 16) * inv, __uint_as_float(v[1] & 0xffff0000u) * inv);
    }
};

// The two 16-bit values of a packed word as float, low half first.
template <typename T> struct Unpack16;
template <> struct Unpack16<__half> {
    __device__ __forceinline__ static float lo(uint32_t w) {
        return __half2float(__builtin_bit_cast(__half, (unsigned short)(w & 0xffffu)));
    }
    __device__ __forceinline__ static float hi(uint32_t w) {
        return __half2float(__builtin_bit_cast(__half, (unsigned short)(w >> 16)));
    }
};
template <> struct Unpack16<__hip_bfloat16> {
    __device__ __forceinline__ static float lo(uint32_t w) { return __uint_as_float(w << 16); }
    __device__ __forceinline__ static float hi(uint32_t w) { return __uint_as_float(w & 0xffff0000u); }
};

// Predicated read: clamped address (always in bounds) then select.
template <typename T>
__device__ __forceinline__ float tap(const T* p, int64_t idx, bool ok, float inv) {
    const float v = Storage<T>::load(p, ok ? (uint64_t)idx : 0ull, inv);
    return ok ? v : 0.0f;
}

inline unsigned ceil_div(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

// Dispatch a templated launcher on a TOMO_F32/F16/BF16 code.
#define TOMO_DISPATCH_STORAGE(code, T, ...)                        \
    switch (code) {                                                \
        case TOMO_F32: { using T = float; __VA_ARGS__; break; }    \
        case TOMO_F16: { using T = __half; __VA_ARGS__; break; }   \
        case TOMO_BF16: { using T = __hip_bfloat16; __VA_ARGS__; break; } \
        default: return fail(TOMO_ERR_UNSUPPORTED, "unsupported element type"); \
    }

}  // namespace tomo
