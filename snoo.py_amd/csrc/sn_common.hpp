// Shared helpers of the libsnoo_hip kernels (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <string>

#include "../../include/snoo_hip.h"

namespace sn {

void set_error(const char *fmt, ...);

#define SN_HIP(call)                                                          \
    do {                                                                      \
        hipError_t e_ = (call);                                               \
        if (e_ != hipSuccess) {                                               \
            ::sn::set_error("%s:%d %s: %s", __FILE__, __LINE__, #call,        \
                            hipGetErrorString(e_));                           \
            return SN_ERR_HIP;                                                \
        }                                                                     \
    } while (0)

#define SN_CHECK_LAUNCH() SN_HIP(hipGetLastError())

#define SN_REQUIRE(cond, ...)                                                 \
    do {                                                                      \
        if (!(cond)) {                                                        \
            ::sn::set_error(__VA_ARGS__);                                     \
            return SN_ERR_ARG;                                                \
        }                                                                     \
    } while (0)

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

constexpr int kWave = 64;

// Sum over a 64-lane wavefront.
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

template <typename T>
__device__ __forceinline__ T wave_max(T v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        T o = __shfl_xor(v, off, kWave);
        v = o > v ? o : v;
    }
    return v;
}

// Block-wide sum; every thread gets the result. `lds` holds >= blockDim/64.
template <typename T>
__device__ __forceinline__ T block_sum(T v, T *lds) {
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = threadIdx.x / kWave;
    const int nw = blockDim.x / kWave;
    v = wave_sum(v);
    __syncthreads();
    if (lane == 0) lds[wid] = v;
    __syncthreads();
    T s = 0;
    for (int w = 0; w < nw; ++w) s += lds[w];  // fixed order
    return s;
}

inline int grid_for(long long n, int block, int max_blocks = 1 << 20) {
    long long g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > max_blocks) g = max_blocks;
    return (int)g;
}

// Number of compute units of the current device (cached per process).
int num_cus();

}  // namespace sn
