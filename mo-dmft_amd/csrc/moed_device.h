// Device helpers shared by the kernel files: wave64 and block reductions in a fixed order.
#pragma once
#include <hip/hip_runtime.h>

namespace moed {

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// wave_sum's value in lane 0 (the other lanes hold partial sums), with the same additions in the
// same tree, built from VALU cross-lane moves instead of LDS permutes: lanes 0-31 take lane + 32
// (v_permlane32_swap), lanes 0-15 lane + 16 (v_permlane16_swap), then DPP row shifts by 8, 4, 2, 1
__device__ __forceinline__ double dbl_dpp_row_shl(double v, int n) {
    int lo = __double2loint(v), hi = __double2hiint(v);
    switch (n) {  // DPP row_shl:n (dpp_ctrl 0x100 + n): lane i takes lane i + n of its 16-lane row
        case 8:
            lo = __builtin_amdgcn_update_dpp(lo, lo, 0x108, 0xf, 0xf, false);
            hi = __builtin_amdgcn_update_dpp(hi, hi, 0x108, 0xf, 0xf, false);
            break;
        case 4:
            lo = __builtin_amdgcn_update_dpp(lo, lo, 0x104, 0xf, 0xf, false);
            hi = __builtin_amdgcn_update_dpp(hi, hi, 0x104, 0xf, 0xf, false);
            break;
        case 2:
            lo = __builtin_amdgcn_update_dpp(lo, lo, 0x102, 0xf, 0xf, false);
            hi = __builtin_amdgcn_update_dpp(hi, hi, 0x102, 0xf, 0xf, false);
            break;
        default:
            lo = __builtin_amdgcn_update_dpp(lo, lo, 0x101, 0xf, 0xf, false);
            hi = __builtin_amdgcn_update_dpp(hi, hi, 0x101, 0xf, 0xf, false);
            break;
    }
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double wave_sum0(double v) {
    {
        const int lo = __double2loint(v), hi = __double2hiint(v);
        const auto a = __builtin_amdgcn_permlane32_swap(lo, lo, false, false);
        const auto b = __builtin_amdgcn_permlane32_swap(hi, hi, false, false);
        v += __hiloint2double(b[1], a[1]);
    }
    {
        const int lo = __double2loint(v), hi = __double2hiint(v);
        const auto a = __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
        const auto b = __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
        v += __hiloint2double(b[1], a[1]);
    }
    v += dbl_dpp_row_shl(v, 8);
    v += dbl_dpp_row_shl(v, 4);
    v += dbl_dpp_row_shl(v, 2);
    v += dbl_dpp_row_shl(v, 1);
    return v;
}

// block sum in a fixed order (wave tree, then waves in index order); result valid in thread 0
template <int NT>
__device__ __forceinline__ double block_sum(double v, double* red) {
    v = wave_sum0(v);
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
