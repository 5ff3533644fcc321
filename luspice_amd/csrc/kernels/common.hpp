// common.hpp -- device helpers shared by the gfx950 kernels.
//
// Conventions
//   * wave64 everywhere: lane = threadIdx.x & 63, wave = threadIdx.x >> 6.
//   * Cross-workgroup hand-off uses epoch-tagged flags: a producer stores
//     flag[i] = epoch with release semantics at agent scope after its data
//     stores; a consumer spins on an acquire load until flag[i] == epoch.
//     Epochs increase per launch, so flags never need clearing.
//   * Every spin is bounded: after kSpinLimit polls the waiter raises the
//     shared abort word, every other waiter sees it and leaves, and the kernel
//     drains normally.  The host reports the abort as an error.
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "desc.hpp"

namespace lsp {

constexpr int kWave = 64;
constexpr int kSpinLimit = 1 << 22;   // ~ seconds with s_sleep(1) between polls

__device__ __forceinline__ int lane_id() { return threadIdx.x & (kWave - 1); }

__device__ __forceinline__ int ld_acquire(const int *p)
{
    return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int ld_relaxed(const int *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_release(int *p, int v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_relaxed(int *p, int v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Orders this wave's earlier LDS / global accesses before its later ones as
// seen by other lanes of the same wave.
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
}

// Spin until *flag == epoch.  Returns false when the wait was abandoned (spin
// limit hit here or elsewhere).
__device__ __forceinline__ bool wait_epoch(const int *flag, int epoch, int *abort_word)
{
    int spins = 0;
    while (ld_acquire(flag) != epoch) {
        if (ld_relaxed(abort_word) != 0) return false;
        __builtin_amdgcn_s_sleep(1);
        if (++spins > kSpinLimit) {
            st_relaxed(abort_word, kErrSpin);
            return false;
        }
    }
    return true;
}

// wave64 reductions (DPP / permute lowered by the compiler from shuffles)
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}
__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, kWave));
    return v;
}
__device__ __forceinline__ int wave_min_int(int v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = min(v, __shfl_xor(v, off, kWave));
    return v;
}

template <typename T> __device__ __forceinline__ double to_f64(T v) { return (double)v; }
template <> __device__ __forceinline__ double to_f64<__half>(__half v) { return (double)__half2float(v); }
template <typename T> __device__ __forceinline__ T from_f64(double v) { return (T)v; }
template <> __device__ __forceinline__ __half from_f64<__half>(double v) { return __float2half((float)v); }

}  // namespace lsp
