// Shared device helpers for the gfx950 kernels (64-wide wavefronts, MFMA, LDS).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "lakshya_device.h"

#define LK_LAUNCH(kernel, grid, block, shmem, stream, ...)                                   \
    do {                                                                                       \
        if ((grid) > 0) hipLaunchKernelGGL(kernel, dim3(grid), dim3(block), shmem,             \
                                           (hipStream_t)(stream), __VA_ARGS__);                \
    } while (0)

static inline int lk_last_error() { return (int)hipGetLastError(); }

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

// One 16x16 output block, K = 4 per instruction.  Operand layout on gfx950 (both types):
//   A (16x4): lane l holds A[l % 16][l / 16];  B (4x16): lane l holds B[l / 16][l % 16].
// Result element e of lane l (measured on the device by tests/test_gpu_probe.py):
//   FP64 v_mfma_f64_16x16x4_f64: D[l / 16 + 4 e][l % 16]
//   FP32 v_mfma_f32_16x16x4_f32: D[4 (l / 16) + e][l % 16]
template <typename T> struct Mfma;
template <> struct Mfma<double> {
    using acc = v4d;
    static __device__ __forceinline__ acc mma(double a, double b, acc c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int e) { return (lane >> 4) + 4 * e; }
};
template <> struct Mfma<float> {
    using acc = v4f;
    static __device__ __forceinline__ acc mma(float a, float b, acc c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int e) { return 4 * (lane >> 4) + e; }
};

__device__ __forceinline__ double lk_shfl(double v, int lane) {
    return __shfl(v, lane, 64);
}
__device__ __forceinline__ float lk_shfl(float v, int lane) { return __shfl(v, lane, 64); }

// Wavefront sum / min / max (64 lanes).
__device__ __forceinline__ double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ double wave_min(double v) {
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ double wave_max(double v) {
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}
// Sum over groups of G consecutive lanes (G power of two <= 64); result valid in every lane.
template <int G>
__device__ __forceinline__ double group_sum(double v) {
    for (int o = G / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ void atomic_add(double* p, double v) { unsafeAtomicAdd(p, v); }
__device__ __forceinline__ void atomic_add(float* p, float v) { unsafeAtomicAdd(p, v); }
