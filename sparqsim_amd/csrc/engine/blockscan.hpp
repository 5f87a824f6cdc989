// Wave64 / workgroup scan helpers (shuffles within a wave, LDS across waves).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace sparq {

template <class T>
__device__ __forceinline__ T wave_incl_scan(T v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        T t = __shfl_up(v, off, 64);
        if (lane >= off) v += t;
    }
    return v;
}

template <class T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Exclusive scan across the workgroup (blockDim.x a multiple of 64, <= 1024).
// `lds` must hold blockDim.x/64 elements.  Returns the exclusive prefix of v;
// `total` receives the workgroup sum.  Contains two barriers.
template <class T>
__device__ __forceinline__ T block_excl_scan(T v, T* lds, T& total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    T inc = wave_incl_scan(v);
    if (lane == 63) lds[wave] = inc;
    __syncthreads();
    T before = 0, tot = 0;
    for (int w = 0; w < nw; ++w) {
        T s = lds[w];
        before += (w < wave) ? s : T(0);
        tot += s;
    }
    __syncthreads();
    total = tot;
    return before + inc - v;
}

// Lanes below the calling lane, as a 64-bit mask.
__device__ __forceinline__ uint64_t lanemask_lt() {
    const int lane = threadIdx.x & 63;
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

}  // namespace sparq
