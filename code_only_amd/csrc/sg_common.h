// Shared helpers of libsg: error reporting, launch checks and wavefront reductions (gfx950,
// 64-lane wavefronts).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>

#include "sg.h"

namespace sg {

void set_error(const char *fmt, ...);

#define SG_HIP(call)                                                                      \
    do {                                                                                  \
        hipError_t e_ = (call);                                                           \
        if (e_ != hipSuccess) {                                                           \
            ::sg::set_error("%s:%d: %s failed: %s", __FILE__, __LINE__, #call,            \
                            hipGetErrorString(e_));                                       \
            return SG_ERR_HIP;                                                            \
        }                                                                                 \
    } while (0)

#define SG_CHECK_LAUNCH() SG_HIP(hipGetLastError())

#define SG_REQUIRE(cond, ...)                                                             \
    do {                                                                                  \
        if (!(cond)) {                                                                    \
            ::sg::set_error(__VA_ARGS__);                                                 \
            return SG_ERR_ARG;                                                            \
        }                                                                                 \
    } while (0)

#define SG_TRY(expr)                                                                      \
    do {                                                                                  \
        int rc_ = (expr);                                                                 \
        if (rc_ != SG_OK) return rc_;                                                     \
    } while (0)

constexpr int kWave = 64;

__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, kWave));
    return v;
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

// Block-wide (max, sum) of per-thread values; result valid in thread 0.  blockDim.x <= 1024.
__device__ __forceinline__ void block_max_sum(double &mx, double &sm)
{
    __shared__ double smax[16], ssum[16];
    mx = wave_max(mx);
    sm = wave_sum(sm);
    const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
    const int nw = (blockDim.x + kWave - 1) / kWave;
    if (lane == 0) { smax[wid] = mx; ssum[wid] = sm; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = smax[0], b = ssum[0];
        for (int w = 1; w < nw; ++w) { a = fmax(a, smax[w]); b += ssum[w]; }
        mx = a; sm = b;
    }
    __syncthreads();
}

// atomic max of a non-negative double: IEEE order of non-negative doubles equals the order of
// their bit patterns read as unsigned integers.
__device__ __forceinline__ void atomic_max_nonneg(double *addr, double v)
{
    atomicMax(reinterpret_cast<unsigned long long *>(addr),
              static_cast<unsigned long long>(__double_as_longlong(v)));
}

// Block-wide max of per-thread non-negative values, folded into *addr with one atomic per block.
// Every thread of the block must call it (it contains a barrier).
__device__ __forceinline__ void block_atomic_max(double *addr, double v)
{
    __shared__ double smax[16];
    v = wave_max(v);
    const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
    const int nw = (blockDim.x + kWave - 1) / kWave;
    if (lane == 0) smax[wid] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = smax[0];
        for (int w = 1; w < nw; ++w) a = fmax(a, smax[w]);
        atomic_max_nonneg(addr, a);
    }
}

} // namespace sg
