// Wave / block helpers shared by the nlhip kernels (gfx950, wave64).
#pragma once
#include "nlhip_internal.h"

namespace nlh {

// ---------------------------------------------------------------------------
// wave / block helpers
// ---------------------------------------------------------------------------

__device__ __forceinline__ int lane_id() { return threadIdx.x & (kWave - 1); }

__device__ __forceinline__ int wave_incl_scan(int v)
{
    const int lane = lane_id();
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        const int t = __shfl_up(v, off, kWave);
        if (lane >= off) v += t;
    }
    return v;
}

__device__ __forceinline__ int64_t wave_incl_scan64(int64_t v)
{
    const int lane = lane_id();
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        const int64_t t = __shfl_up(v, off, kWave);
        if (lane >= off) v += t;
    }
    return v;
}

__device__ __forceinline__ int64_t readlane64(int64_t v, int l)
{
    const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)(uint64_t)v, l);
    const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)((uint64_t)v >> 32), l);
    return (int64_t)(((uint64_t)hi << 32) | lo);
}

__device__ __forceinline__ uint32_t mbcnt64(uint64_t m)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Exclusive scan over an NT-thread block; returns this thread's exclusive
// prefix and the block total.
template <int NT>
__device__ __forceinline__ int64_t block_excl_scan_n(int64_t v, int64_t &total)
{
    __shared__ int64_t wsum[NT / kWave];
    const int lane = lane_id(), wv = threadIdx.x / kWave;
    const int64_t inc = wave_incl_scan64(v);
    if (lane == kWave - 1) wsum[wv] = inc;
    __syncthreads();
    int64_t before = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < NT / kWave; ++w) {
        const int64_t s = wsum[w];
        if (w < wv) before += s;
        tot += s;
    }
    __syncthreads();
    total = tot;
    return before + inc - v;
}

__device__ __forceinline__ int64_t block_excl_scan(int64_t v, int64_t &total)
{
    return block_excl_scan_n<kScanThreads>(v, total);
}

__device__ __forceinline__ double readlane_f64(double v, int l)
{
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)b, l);
    const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(b >> 32), l);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// Blocks are dealt to the 8 XCDs round-robin; remap so every XCD works on a
// contiguous range of cells and keeps its neighbour-cell data in its own L2.
__device__ __forceinline__ uint32_t xcd_remap(uint32_t b, uint32_t nb)
{
    const uint32_t q = nb / 8, r = nb % 8;
    const uint32_t xcd = b % 8, idx = b / 8;
    return xcd < r ? xcd * (q + 1) + idx : r * (q + 1) + (xcd - r) * q + idx;
}

template <typename TI>
__device__ __forceinline__ void store_ti(void *p, int64_t idx, int64_t v)
{
    reinterpret_cast<TI *>(p)[idx] = (TI)v;
}

}  // namespace nlh
