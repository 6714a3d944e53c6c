// Device helpers shared by the soft-variant kernels (soft_features.hip,
// soft_cost.hip, soft_bp.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

#include "kernels.h"
#include "soft.h"

namespace deeds {
namespace soft_dev {

// exp2 / log2: the raw v_exp_f32 / v_log_f32 instructions for float (about
// 1 ulp, enough for a smooth relaxation), libm for the double checks.
template <typename R>
struct FastMath;
template <>
struct FastMath<float> {
    static __device__ __forceinline__ float exp2(float x) { return __builtin_amdgcn_exp2f(x); }
    static __device__ __forceinline__ float log2(float x) { return __builtin_amdgcn_logf(x); }
};
template <>
struct FastMath<double> {
    static __device__ __forceinline__ double exp2(double x) { return ::exp2(x); }
    static __device__ __forceinline__ double log2(double x) { return ::log2(x); }
};

constexpr double kLog2e = 1.4426950408889634;
constexpr double kLn2 = 0.6931471805599453;

template <typename R>
__device__ __forceinline__ R big();
template <>
__device__ __forceinline__ float big() { return 3.0e38f; }
template <>
__device__ __forceinline__ double big() { return 1.0e300; }

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// minimum as one instruction (v_min_f32 / v_min_f64)
__device__ __forceinline__ float rmin(float a, float b) { return __builtin_fminf(a, b); }
__device__ __forceinline__ double rmin(double a, double b) { return __builtin_fmin(a, b); }

// Minimum (is_min) or sum over the workgroup; every thread gets the result.
// red: one slot per wavefront (blockDim.x / 64 <= 16).
template <typename R>
__device__ R block_reduce(R v, R* red, bool is_min) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const R o = __shfl_xor(v, off, 64);
        v = is_min ? rmin(o, v) : v + o;
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[wid] = v;
    __syncthreads();
    R r = red[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); w++) r = is_min ? rmin(red[w], r) : r + red[w];
    __syncthreads();
    return r;
}

inline void set_smem(const void* fn, size_t bytes) {
    DEEDS_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
}

}  // namespace soft_dev
}  // namespace deeds
