// Device helpers shared by the kernel files: wave64 and block reductions in a fixed order.
#pragma once
#include <hip/hip_runtime.h>

namespace moed {

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// block sum in a fixed order (wave tree, then waves in index order); result valid in thread 0
template <int NT>
__device__ __forceinline__ double block_sum(double v, double* red) {
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) red[wid] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0) {
        for (int w = 0; w < NT / 64; ++w) s += red[w];
    }
    __syncthreads();
    return s;
}

// Read-only, wave-uniform data is read through the constant address space so that the compiler
// emits scalar loads (scalar cache, SGPRs) instead of wave-wide vector loads, which would occupy
// the texture address/data units like any other vector memory instruction.
template <class T>
using cptr = const __attribute__((address_space(4))) T*;
template <class T>
__device__ __forceinline__ cptr<T> as_const(const T* p) {
    return (cptr<T>)p;
}

__device__ __forceinline__ uint32_t uniform_u32(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

// element i of a uint16 array read with a scalar load (the array must be readable up to the
// next 4-byte boundary)
__device__ __forceinline__ uint32_t scalar_u16(const uint16_t* a, int i) {
    const cptr<uint32_t> w = (cptr<uint32_t>)(const __attribute__((address_space(4))) void*)a;
    return (w[i >> 1] >> ((i & 1) * 16)) & 0xffffu;
}
__device__ __forceinline__ int32_t uniform_i32(int32_t v) { return __builtin_amdgcn_readfirstlane(v); }

}  // namespace moed
